! Driver for the reference build (oracle/_ref): links the SIMOX routines that
! compile from their own sources without the absent FFTE library and runs them
! on one input record.  TEST INFRASTRUCTURE ONLY: it writes golden vectors for
! tests/golden and is never part of the product.
!
! The reference's own main program (main.f90) is not used: its RNG is unseeded
! and its config file cannot express the inputs the tests need (main.f90:28,
! 122-180).  This driver sets Nset/sys exactly as READ_INPUT does
! (main.f90:137-201) and then calls, in order:
!   CELL_HIERARCHY, CELL_SORT            cellsort.f90:3-80
!   POTENTIAL_SETUP                      datafmt.f90:142-262
!   PME_SELF                             force.f90:467-517
!   FORCE                                force.f90:3-361
!   VV_NVE1 / VV_NVE2 (on copies)        vverlet.f90:74-157
!   VV_NVT2_BEREND (on a copy after VV_NVE1, T_given)   vverlet.f90:160-208
!   VV_NVT1_STOCH / VV_NVT2_STOCH (on a copy, T_given = 0, so the random
!     force is zero and the result deterministic)        vverlet.f90:237-296
!   VV_NVE1_REFLECT then REMOVE_RIGID_MOTION (on a copy with z shifted by
!     zshift, so that the top atoms cross the reflecting wall) vverlet.f90:100-129,211-234
!   dUdq (on matrices read from input)   ctip.f90:307-333
! PME, PHI_MESH, CTIP_SOLVER and BRENT_SOLVER call FFTE's ZFFT3D and are not
! linked (link-time optimisation drops them because nothing here calls them).
!
! With a third argument R > 0 it then times R repetitions of the per-step pair
! work the reference can run here, CELL_SORT + FORCE (main.f90:41-61 without
! the CTIP solve and PME, which need FFTE), and prints "REF_TIME R seconds".
!
! Input  (stream file, native endian):
!   int32 n, fft(3), nnb; real64 box(3), rcut, dt, t_given (K), alpha, zshift
!   int32 id(n); real64 x(3,n), v(3,n), f(3,n), q(n)
!   int32 nlist(n), list(n,500); real64 asym(n,500), sym(n,500), phi(n)
!   (nnb = 0 skips the dUdq block)
! Output (stream file):
!   int32 ncell(3), npt(ncell_all), link(500,ncell_all), pair(13,ncell_all)
!   real64 f(3,n), energies(4) = Epair, Edir, Ectip, Egroup, eself, psi(fft)
!   real64 x1(3,n), v1(3,n) after VV_NVE1, v2(3,n), mv2 after VV_NVE2 (from v1)
!   int32 npt_new; int32 rdf(5,5,201)
!   real64 vb(3,n), mv2b after VV_NVT2_BEREND (from the VV_NVE1 state)
!   real64 xs(3,n), vs(3,n) after VV_NVT1_STOCH; vs2(3,n), fs2(3,n), mv2s
!     after VV_NVT2_STOCH with T_given = 0
!   real64 xr(3,n), vr(3,n) after VV_NVE1_REFLECT; xr2(3,n), vr2(3,n) after
!     REMOVE_RIGID_MOTION
!   real64 g(n) (if nnb > 0)
! Optional trailing input: int32 n_o2, seed, ncrit; real64 z_new, q_crit(ncrit)
!   -> int32 new_o2(ncrit) of OXIDE_CHECK (util.f90:152-188) per q_crit, then
!   NEW_O2 (datafmt.f90:98-139) after RANDOM_SEED(put = (seed, 1, 1, ...)):
!   real64 x(3, 2 n_o2), v(3, 2 n_o2); int32 id(2 n_o2); real64 q(2 n_o2)
program ref_driver
use datafmt
implicit none
type(amnt) :: nset
type(stat) :: sys
type(metl) :: eam(nmetal)
type(potn) :: param(natom)
type(oxid) :: oxygn
type(corr) :: pair
type(ptcl), allocatable :: q(:), q1(:), qb(:), qs(:), qr(:)
type(sort), allocatable :: cell(:)
real(dp), allocatable :: psi(:,:,:), xyz(:,:), vel(:,:), frc(:,:), chg(:)
real(dp), allocatable :: xjz(:,:), asym(:,:), sym(:,:), phi(:), qmx(:,:), g(:)
integer(gi), allocatable :: ids(:), list(:,:), nlist(:)
integer(gi) :: n, fft(3), nnb, i, m
real(dp) :: box(3), rcut, dt, t_given, alpha, zshift
character(len=512) :: fin, fout, arg
integer(8) :: c0, c1, crate
integer :: reps, ios, nseed
integer(gi) :: n_o2, seedv, ncrit
real(dp) :: z_new
real(dp), allocatable :: q_crit(:)
integer, allocatable :: seed(:)
type(tckt) :: tag
type(ptcl), pointer :: qp(:)

call get_command_argument(1, fin)
call get_command_argument(2, fout)
open(10, file=trim(fin), access='stream', form='unformatted', status='old')
read(10) n, fft, nnb
read(10) box, rcut, dt, t_given, alpha, zshift
allocate(ids(n), xyz(3,n), vel(3,n), frc(3,n), chg(n))
read(10) ids, xyz, vel, frc, chg

nset%npt_all = n
nset%box(:) = box(:)
nset%v = box(1)*box(2)*box(3)
nset%fft(:) = fft(:)
sys%box(:) = box(:)
sys%rcut = rcut
sys%rcut2 = rcut**2
sys%a = 3.8d0/rcut
nset%ncell(:) = int(box(:)/rcut)
nset%lcell(:) = box(:)/dble(nset%ncell(:))
nset%ncell_all = nset%ncell(1)*nset%ncell(2)*nset%ncell(3)
pair%dr = rcut/dble(nrdf)
pair%v = nset%v
pair%rdf(:,:,:) = 0

allocate(q(n), q1(n), qb(n), qs(n), qr(n), cell(nset%ncell_all))
nset%alpha = alpha
sys%t_given = t_given*kb
allocate(psi(fft(1),fft(2),fft(3)))
psi(:,:,:) = 0.0d0
do i = 1, n
   q(i)%id = ids(i)
   q(i)%xx(:) = xyz(:,i)
   q(i)%xv(:) = vel(:,i)
   q(i)%ff(:) = frc(:,i)
   q(i)%q = chg(i)
end do

call cell_hierarchy(nset, cell)
call cell_sort(nset, cell, q)
call potential_setup(eam, param, oxygn)
call pme_self(nset, sys, param, q, psi)

! integrator on copies, using the input forces
q1(:) = q(:)
qs(:) = q(:)
qr(:) = q(:)
do i = 1, n
   qr(i)%xx(3) = qr(i)%xx(3) + zshift
end do
call vv_nve1(nset, param, q1, dt)
qb(:) = q1(:)

open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
write(11) nset%ncell
write(11) (cell(m)%npt, m = 1, nset%ncell_all)
write(11) (cell(m)%link(:), m = 1, nset%ncell_all)
write(11) (cell(m)%pair(:), m = 1, nset%ncell_all)

call force(nset, eam, oxygn, sys, param, pair, q, cell)
write(11) (q(i)%ff(:), i = 1, n)
write(11) sys%epair, sys%edir, sys%ectip, sys%egroup, sys%eself
write(11) psi
write(11) (q1(i)%xx(:), i = 1, n)
write(11) (q1(i)%xv(:), i = 1, n)
call vv_nve2(nset, param, sys, q1, dt)
write(11) (q1(i)%xv(:), i = 1, n)
write(11) sys%mv2
write(11) pair%npt_new
write(11) pair%rdf
call vv_nvt2_berend(nset, param, sys, qb, dt)
write(11) (qb(i)%xv(:), i = 1, n)
write(11) sys%mv2
call vv_nvt1_stoch(nset, param, qs, dt)
write(11) (qs(i)%xx(:), i = 1, n)
write(11) (qs(i)%xv(:), i = 1, n)
sys%t_given = 0.0d0
call vv_nvt2_stoch(nset, param, sys, qs, dt)
write(11) (qs(i)%xv(:), i = 1, n)
write(11) (qs(i)%ff(:), i = 1, n)
write(11) sys%mv2
call vv_nve1_reflect(nset, param, qr, dt)
write(11) (qr(i)%xx(:), i = 1, n)
write(11) (qr(i)%xv(:), i = 1, n)
call remove_rigid_motion(nset, qr)
write(11) (qr(i)%xx(:), i = 1, n)
write(11) (qr(i)%xv(:), i = 1, n)

if (nnb > 0) then
   allocate(xjz(n,3), asym(n,nnb_max), sym(n,nnb_max), phi(n), qmx(n,2), g(n))
   allocate(list(n,nnb_max), nlist(n))
   read(10) nlist, list, asym, sym, phi
   do i = 1, n
      xjz(i,1) = param(ids(i))%chi
      xjz(i,2) = param(ids(i))%j - kc*2.d0*sys%a/sqrtpi
      xjz(i,3) = param(ids(i))%z
      qmx(i,1) = param(ids(i))%qmin
      qmx(i,2) = param(ids(i))%qmax
   end do
   call dudq(n, xjz, asym, sym, list, nlist, chg, qmx, phi, g)
   write(11) g
end if

! optional OXIDATION block: OXIDE_CHECK for each q_crit on the input record
! (N_o2 last atoms), then NEW_O2 after RANDOM_SEED(put = (seed, 1, 1, ...))
read(10, iostat=ios) n_o2, seedv, ncrit
if (ios == 0) then
   allocate(q_crit(ncrit))
   read(10) z_new, q_crit
   nset%n_o2 = n_o2
   nset%z_new = z_new
   do m = 1, ncrit
      nset%q_crit = q_crit(m)
      call oxide_check(nset, tag, q)
      write(11) merge(1_gi, 0_gi, tag%new_o2)
   end do
   call random_seed(size=nseed)
   allocate(seed(nseed))
   seed = 1
   seed(1) = seedv
   call random_seed(put=seed)
   sys%t_given = t_given*kb
   allocate(qp(n))
   qp(:) = q(:)
   call new_o2(nset, sys, qp)
   write(11) (qp(i)%xx(:), i = n + 1, n + 2*n_o2)
   write(11) (qp(i)%xv(:), i = n + 1, n + 2*n_o2)
   write(11) (qp(i)%id, i = n + 1, n + 2*n_o2)
   write(11) (qp(i)%q, i = n + 1, n + 2*n_o2)
   nset%npt_all = n
end if
close(10)
close(11)

if (command_argument_count() >= 3) then
   call get_command_argument(3, arg)
   read(arg, *) reps
   if (reps > 0) then
      call system_clock(c0, crate)
      do m = 1, reps
         call cell_sort(nset, cell, q)
         call force(nset, eam, oxygn, sys, param, pair, q, cell)
      end do
      call system_clock(c1)
      write(*, '(A,I8,ES16.8)') 'REF_TIME ', reps, dble(c1 - c0)/dble(crate)
   end if
end if
end program ref_driver
