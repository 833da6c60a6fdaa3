! simox_md: Fortran host program for SIMOX runs on one MI355X.  Reads the reference's input files (config.prm layout of
! main.f90:122-180; restart.xyz "El x y z vx vy vz fx fy fz q" or input.xyz
! "El x y z q", main.f90:205-245) and runs the step loop of main.f90:41-61
! device-resident through the C-ABI (simox_step), with TAG_SET's cadence
! (main.f90:296-334): charges re-equilibrated when MOD(Nloop, freq_new) == 1,
! cells re-sorted when MOD(Nloop, freq_sort) == 0.
!
! Output: temporal_status.dat in PRINT_STAT's format (util.f90:12-18), plus
! temporal_status_full.dat with the same columns at full precision, and
! snapshotNNN.xyz restart files in PRINT_SNAPSHOT's format (util.f90:23-81).
!
! Thermostat YES selects Berendsen (VV_NVT2_BEREND) or, with freq_new < 0,
! Langevin (VV_NVT1/2_STOCH) as main.f90:44-58 does; the pair histogram is
! written every time_rdf as PRINT_RDF does (util.f90:246-307).
!
! OXIDATION runs follow main.f90:62-86 call by call: VV_NVE1_REFLECT, cell
! sort, and on charge steps REMOVE_RIGID_MOTION, OXIDE_CHECK (util.f90:
! 152-188) and NEW_O2 (datafmt.f90:98-139, positions of RANDOM_POSITION_O2,
! util.f90:104-149) before the charge solve; the device context is rebuilt
! for the grown atom count and the new atoms are binned into cells at once.
!
! A non-restart run initialises velocities as VINIT does (vverlet.f90:2-71):
! uniform random components, rigid motion removed, each species scaled to
! T_given, and the first forces evaluated at x - dt*v.  The generator (VINIT,
! new O2 positions) is seeded from SIMOX_SEED (default 1) where the
! reference's RANDOM_SEED is unseeded.
!
! Usage: simox_md [mode]   (mode = faithful | fp64 | mixed; default faithful)
program simox_md
  use iso_c_binding
  use simox_gpu
  use simox_oxid, only: oxide_check, o2_dimers
  implicit none
  real(c_double), parameter :: tfm = 10.180505306645389d0, kb = 8.617343d-5
  character(len=256) :: dummy, runtype, tstat, rstart, fname, modearg
  real(c_double) :: time_max, time_snap, time_rdf, time_stat, dt, box(3), t_given, alpha_l
  real(c_double) :: rcut, z_new, q_crit, stats(8)
  integer(c_int) :: fft(3), freq_new, freq_sort, n_o2, mode, rc, n
  integer :: freq_max, freq_snap, freq_stat, freq_rdf, nloop, i, ios
  logical :: restart, do_ctip, do_sort, do_stat, thermostat, regular, new_o2
  real(c_double) :: e_self, e_rec, e4(4), mv2
  integer(c_long_long) :: hist_carry(5, 5, 201)
  integer :: npt_old
  integer(c_long_long), target :: hist(5, 5, 201)
  real(c_double), allocatable, target :: x(:,:), v(:,:), f(:,:), q(:)
  integer(c_int), allocatable, target :: id(:)
  character(len=3), allocatable :: el(:)
  type(c_ptr) :: ctx
  real(c_double), target :: st(8)

  mode = SIMOX_MODE_FAITHFUL
  if (command_argument_count() >= 1) then
    call get_command_argument(1, modearg)
    if (trim(modearg) == 'fp64') mode = SIMOX_MODE_FP64
    if (trim(modearg) == 'mixed') mode = SIMOX_MODE_MIXED
  end if

  open(10, file='config.prm', status='old', action='read', iostat=ios)
  if (ios /= 0) stop "#### cannot  open config.prm ###"
  read(10, *) dummy
  read(10, *) dummy
  read(10, *) runtype
  if (to_upper(runtype) == 'REGULAR') then
    regular = .true.
  else if (to_upper(runtype) == 'OXIDATION') then
    regular = .false.
  else
    stop "=== Simulation type is unknown ==="
  end if
  read(10, *) dummy
  read(10, *) time_max, time_snap, time_rdf, time_stat, dt
  freq_max = nint(time_max / dt)
  freq_snap = nint(time_snap / dt)
  freq_stat = nint(time_stat / dt)
  freq_rdf = nint(time_rdf / dt)
  dt = dt / tfm
  read(10, *) dummy
  read(10, *) box
  read(10, *) dummy
  read(10, *) fft
  read(10, *) dummy
  read(10, *) tstat, t_given, alpha_l
  if (to_upper(tstat) == 'YES') then
    thermostat = .true.
  else if (to_upper(tstat) == 'NO') then
    thermostat = .false.
  else
    stop "=== Temperature control is unknown ==="
  end if
  read(10, *) dummy
  read(10, *) rstart
  restart = to_upper(rstart) == 'YES'
  read(10, *) dummy
  read(10, *) rcut
  read(10, *) dummy
  read(10, *) freq_new, freq_sort
  read(10, *) dummy
  read(10, *) z_new, n_o2
  read(10, *) dummy
  read(10, *) q_crit
  close(10)

  fname = 'input.xyz'
  if (restart) fname = 'restart.xyz'
  open(11, file=trim(fname), status='old', action='read', iostat=ios)
  if (ios /= 0) stop "=== cannot open the particle file ==="
  read(11, *) n
  read(11, *) dummy
  allocate(x(3, n), v(3, n), f(3, n), q(n), id(n), el(n))
  v = 0.0d0
  f = 0.0d0
  do i = 1, n
    if (restart) then
      read(11, *) el(i), x(:, i), v(:, i), f(:, i), q(i)
    else
      read(11, *) el(i), x(:, i), q(i)
    end if
    select case (trim(el(i)))
    case ('Al'); id(i) = 1
    case ('Ni'); id(i) = 2
    case ('Fe'); id(i) = 3
    case ('O'); id(i) = 4
    case ('H'); id(i) = 5
    case default; stop "=== Particle type error ==="
    end select
  end do
  close(11)

  rc = simox_create(n, box, fft, rcut, mode, c_null_ptr, ctx)
  call simox_check(ctx, rc)
  rc = simox_set_state(ctx, c_loc(id), c_loc(x), c_loc(v), c_loc(f), c_loc(q))
  call simox_check(ctx, rc)
  rc = simox_pme_self(ctx, e_self)   ! main.f90:33
  call simox_check(ctx, rc)
  call seed_rng()
  hist_carry = 0
  npt_old = n
  if (.not. restart) call vinit()
  if (thermostat) then
    if (freq_new < 0 .and. regular) then
      rc = simox_set_thermostat(ctx, 2, t_given, alpha_l, 1_c_long_long)
    else
      rc = simox_set_thermostat(ctx, 1, t_given, alpha_l, 1_c_long_long)
    end if
    call simox_check(ctx, rc)
  end if
  rc = simox_set_rdf(ctx, 1)
  call simox_check(ctx, rc)

  open(15, file='temporal_status.dat')
  open(16, file='temporal_status_full.dat')
  write(15, *) "# time(fs) - T(K) - E_total -   E_Ewald   -   E_eam  -   ", &
       & "Epair   -  Egroup (eV)"
  write(16, '(A)') "# time(fs) T(K) E_total E_es E_eam Epair Egroup (full precision)"
  if (.not. regular) then
    rc = simox_set_reflect(ctx, 1)
    call simox_check(ctx, rc)
  end if
  do nloop = 1, freq_max
    do_ctip = freq_new >= 0 .and. mod(nloop, freq_new) == 1
    do_sort = mod(nloop, freq_sort) == 0
    do_stat = mod(nloop, freq_stat) == 0
    if (.not. regular) then
      call oxidation_step()
    else if (do_stat) then
      rc = simox_step(ctx, 1, dt, merge(1, 0, do_sort), merge(1, 0, do_ctip), c_loc(st))
    else
      rc = simox_step(ctx, 1, dt, merge(1, 0, do_sort), merge(1, 0, do_ctip), c_null_ptr)
    end if
    if (regular) call simox_check(ctx, rc)
    if (mod(nloop, freq_snap) == 0) call print_snapshot(nloop / freq_snap)
    if (do_stat) call print_stat(nloop)
    if (mod(nloop, freq_rdf) == 0) call print_rdf(nloop)
  end do
  close(15)
  close(16)
  rc = simox_destroy(ctx)

contains

  subroutine seed_rng()
    integer :: nseed
    integer, allocatable :: seed(:)
    character(len=32) :: sv
    call random_seed(size=nseed)
    allocate(seed(nseed))
    seed = 1
    call get_environment_variable('SIMOX_SEED', sv)
    if (len_trim(sv) > 0) read(sv, *) seed(1)
    call random_seed(put=seed)
  end subroutine seed_rng

  ! One OXIDATION step (main.f90:63-80), call by call; fills st(1:7).
  subroutine oxidation_step()
    integer(c_int) :: it, nev
    rc = simox_vv1(ctx, dt)                 ! VV_NVE1_REFLECT
    call simox_check(ctx, rc)
    if (do_sort) then
      rc = simox_cell_sort(ctx)
      call simox_check(ctx, rc)
    end if
    if (do_ctip) then
      rc = simox_remove_rigid_motion(ctx)
      call simox_check(ctx, rc)
      rc = simox_get_state(ctx, c_loc(x), c_null_ptr, c_null_ptr, c_loc(q))
      call simox_check(ctx, rc)
      new_o2 = oxide_check(n, id, x, q, n_o2, q_crit)
      if (new_o2) call add_o2()
      rc = simox_ctip_solve(ctx, e_self, it, nev)
      call simox_check(ctx, rc)
    end if
    rc = simox_force(ctx, e4)
    call simox_check(ctx, rc)
    rc = simox_pme(ctx, e_rec)
    call simox_check(ctx, rc)
    rc = simox_vv2(ctx, dt, mv2)            ! VV_NVE2 or VV_NVT2_BEREND
    call simox_check(ctx, rc)
    st(1:7) = [e_self, e4(2), e_rec, e4(3), e4(1), e4(4), mv2]
  end subroutine oxidation_step

  ! NEW_O2 (datafmt.f90:98-139): n_o2 new dimers (o2_dimers: positions of
  ! RANDOM_POSITION_O2, v = (0, 0, -sqrt(3 kB T / 16))) with f = 0, q = 0
  ! appended; the context is rebuilt for the new atom count.
  subroutine add_o2()
    real(c_double), allocatable, target :: xn(:,:), vn(:,:), fn(:,:), qn(:)
    integer(c_int), allocatable, target :: idn(:)
    character(len=3), allocatable :: eln(:)
    integer :: m, nn
    integer(c_long_long), target :: h(5, 5, 201)
    real(c_double), allocatable :: xo(:,:), vo(:,:)
    if (n_o2 <= 0) return
    allocate(xo(3, 2 * n_o2), vo(3, 2 * n_o2))
    call o2_dimers(n_o2, z_new, box(1), box(2), t_given * kb, xo, vo)
    rc = simox_get_state(ctx, c_loc(x), c_loc(v), c_loc(f), c_loc(q))
    call simox_check(ctx, rc)
    rc = simox_get_rdf(ctx, h, 1)
    call simox_check(ctx, rc)
    hist_carry = hist_carry + h
    nn = n + 2 * n_o2
    allocate(xn(3, nn), vn(3, nn), fn(3, nn), qn(nn), idn(nn), eln(nn))
    xn(:, 1:n) = x; vn(:, 1:n) = v; fn(:, 1:n) = f; qn(1:n) = q; idn(1:n) = id; eln(1:n) = el
    xn(:, n + 1:nn) = xo
    vn(:, n + 1:nn) = vo
    do m = n + 1, nn
      fn(:, m) = 0.0d0
      qn(m) = 0.0d0
      idn(m) = 4
      eln(m) = 'O'
    end do
    call move_alloc(xn, x); call move_alloc(vn, v); call move_alloc(fn, f)
    call move_alloc(qn, q); call move_alloc(idn, id); call move_alloc(eln, el)
    n = nn
    rc = simox_destroy(ctx)
    rc = simox_create(n, box, fft, rcut, mode, c_null_ptr, ctx)
    call simox_check(ctx, rc)
    rc = simox_set_state(ctx, c_loc(id), c_loc(x), c_loc(v), c_loc(f), c_loc(q))
    call simox_check(ctx, rc)
    rc = simox_set_reflect(ctx, 1)
    call simox_check(ctx, rc)
    if (thermostat) then
      rc = simox_set_thermostat(ctx, 1, t_given, alpha_l, 1_c_long_long)
      call simox_check(ctx, rc)
    end if
    rc = simox_set_rdf(ctx, 1)
    call simox_check(ctx, rc)
  end subroutine add_o2

  ! VINIT (vverlet.f90:2-71): random velocities at T_given, forces at x - dt*v
  subroutine vinit()
    real(c_double), allocatable, target :: x0(:,:), v0(:,:)
    real(c_double) :: sumv(3), mv2s(5), xm(5), r, lam(5), e4(4), erec
    integer :: nsp(5), k, j
    type(c_params) :: prm
    call simox_potential_setup(prm)
    xm = prm%prm(:)%xm
    sumv = 0.0d0
    do k = 1, n
      do j = 1, 3
        call random_number(r)
        v(j, k) = r - 0.5d0
        sumv(j) = sumv(j) + v(j, k)
      end do
    end do
    sumv = sumv / dble(n)
    mv2s = 0.0d0
    nsp = 0
    do k = 1, n
      v(:, k) = v(:, k) - sumv
      mv2s(id(k)) = mv2s(id(k)) + xm(id(k)) * sum(v(:, k)**2)
      nsp(id(k)) = nsp(id(k)) + 1
    end do
    lam = 0.0d0
    do j = 1, 5
      if (nsp(j) > 0) lam(j) = sqrt(t_given * kb / (mv2s(j) / dble(3 * nsp(j))))
    end do
    do k = 1, n
      v(:, k) = v(:, k) * lam(id(k))
    end do
    f = 0.0d0
    x0 = x
    v0 = v
    rc = simox_set_state(ctx, c_null_ptr, c_null_ptr, c_loc(v), c_loc(f), c_null_ptr)
    call simox_check(ctx, rc)
    rc = simox_vv1(ctx, -dt)
    call simox_check(ctx, rc)
    rc = simox_force(ctx, e4)
    call simox_check(ctx, rc)
    rc = simox_pme(ctx, erec)
    call simox_check(ctx, rc)
    rc = simox_set_state(ctx, c_null_ptr, c_loc(x0), c_loc(v0), c_null_ptr, c_null_ptr)
    call simox_check(ctx, rc)
  end subroutine vinit

  function to_upper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: k, c
    u = s
    do k = 1, len(s)
      c = ichar(s(k:k))
      if (c >= 97 .and. c <= 122) u(k:k) = char(c - 32)
    end do
  end function to_upper

  ! E_total = Eself + Edir + Erec + Ectip + Epair + Egroup + mv2/2 (util.f90:12-15)
  subroutine print_stat(nl)
    integer, intent(in) :: nl
    real(c_double) :: e_es, e_eam, e_tot, temp
    e_es = st(1) + st(2) + st(3) + st(4)
    e_eam = st(5) + st(6)
    e_tot = e_es + e_eam + st(7) / 2.0d0
    temp = st(7) / dble(n * 3) / kb
    write(15, '(ES11.4, 1X, F6.1, 5(1X, ES11.4))') nl * tfm * dt, temp, e_tot, e_es, e_eam, &
         & st(5), st(6)
    write(16, '(7(ES24.16, 1X))') nl * tfm * dt, temp, e_tot, e_es, e_eam, &
         & st(5), st(6)
  end subroutine print_stat

  ! PRINT_RDF (util.f90:246-307): histogram of the FORCE calls since the last
  ! print, averaged over freq_rdf, normalised as there; then cleared.
  subroutine print_rdf(nl)
    integer, intent(in) :: nl
    real(c_double), parameter :: pi = 3.1415926535897931d0
    real(c_double) :: res(15, 200), dr, r, cst, loop_sum, xrdf
    character(len=64) :: fname
    integer :: a, b, m, k, npt_new
    rc = simox_get_rdf(ctx, hist, 1)
    call simox_check(ctx, rc)
    hist = hist + hist_carry
    hist_carry = 0
    xrdf = dble(freq_rdf)
    m = 0
    do a = 1, 5
      do b = a, 5
        m = m + 1
        if (a /= b) then
          res(m, :) = dble(hist(a, b, 1:200) + hist(b, a, 1:200)) / xrdf
        else
          res(m, :) = dble(hist(a, a, 1:200)) / xrdf
        end if
      end do
    end do
    dr = rcut / 200.0d0
    npt_new = n
    loop_sum = dble(npt_old * npt_new) * 4.d0 * pi / 3.0d0 / (box(1) * box(2) * box(3))
    write(fname, '("rdf", I3.3, ".dat")') nl / freq_rdf
    open(22, file=trim(fname), action='write')
    write(22, '("# Header: current loop = ",I7," Npt_old = ", I5, " Npt_new = ", I5)') nl, &
         & npt_old, npt_new
    write(22, '(A)') "# Radius(A) / Al-Al(2) / Al-Ni(3) / Al-Fe(4) / Al-O(5) / Al-H(6) " // &
         & "/ Ni-Ni(7) / Ni-Fe(8) / Ni-O(9) / Ni-H(10) / Fe-Fe(11) / Fe-O(12)" // &
         & " / Fe-H(13) / O-O(14) / O-H(15) / H-H(16) "
    do k = 1, 200
      r = dble(k) * dr
      cst = ((r + dr)**3 - dr**3) * loop_sum
      write(22, '(ES10.3, 1X)', advance='no') dr * (dble(k) - 0.5d0)
      do m = 1, 15
        write(22, '(ES10.3, 1X)', advance='no') res(m, k) / cst
      end do
      write(22, *)
    end do
    close(22)
    npt_old = npt_new
  end subroutine print_rdf

  subroutine print_snapshot(nframe)
    integer, intent(in) :: nframe
    character(len=64) :: rest, snap
    integer :: k
    rc = simox_get_state(ctx, c_loc(x), c_loc(v), c_loc(f), c_loc(q))
    call simox_check(ctx, rc)
    write(rest, '("snapshot", I3.3, ".xyz")') nframe
    write(snap, '("image", I3.3, ".xyz")') nframe
    open(20, file=trim(rest), action='write')
    open(21, file=trim(snap), action='write')
    write(20, *) n
    write(20, *) "Frame = ", nframe, " energy =  0.0 "
    write(21, *) n
    write(21, *) "Frame = ", nframe, " energy =  0.0 "
    do k = 1, n
      write(20, '(A3, 3(ES13.6, 1X), 6(ES11.4, 1X), ES13.6)') el(k), x(:, k), v(:, k), f(:, k), q(k)
      write(21, '(A3, 4(ES11.4, 1X))') el(k), x(:, k), q(k)
    end do
    close(20)
    close(21)
  end subroutine print_snapshot

end program simox_md
