! Drop-in replacement of SIMOX's force.f90 (FORCE, PME, PME_SELF), ctip.f90
! (CTIP_SOLVER), cellsort.f90 (CELL_HIERARCHY, ID_CELL, CELL_SORT) and the
! NVE integrator of vverlet.f90 (VV_NVE1, VV_NVE2): the same external names
! and dummy-argument lists, computed by the MI355X library through
! iso_c_binding.  Link this unit in place of force.o, ctip.o and cellsort.o
! (FFTE is then no longer needed), and ahead of vverlet.o with the linker's
! first-definition rule for the two integrator halves (vverlet.f90 keeps
! VINIT, the thermostats and OXIDATION's reflecting half); it uses the
! caller's DATAFMT module, so it is compiled against the reference's
! datafmt.f90.
!
! Device state: one context, created on the first call with the published
! parameters and re-created when Npt_all changes (OXIDATION mode appends
! atoms, datafmt.f90:135).  The cell lists live on the device and change only
! when the caller runs CELL_SORT, as the reference's FORCE and CTIP_SOLVER
! read the lists of the last CELL_SORT (main.f90:47); CELL_SORT also fills the
! caller's cell(:)%Npt and %link as cellsort.f90:53-80 does.  Every other call
! uploads the positions / charges (/ forces for PME) of q(:) and copies the
! results back into q(:) and sys.  FORCE and CTIP_SOLVER pass their eam /
! oxygn / param arguments to the device whenever they change.
! Mode: FAITHFUL by default; SIMOX_MODE=fp64 or mixed selects the converged
! solves.
module simox_dropin_state
  use iso_c_binding
  use datafmt
  use simox_gpu
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: n_ctx = -1
  type(c_params), target, save :: cur
  real(c_double), allocatable, target, save :: xb(:,:), fb(:,:), qb(:), vb(:,:)
  ! positions at a CELL_SORT that ran before the context existed
  real(c_double), allocatable, target, save :: xsort(:,:)
  integer(c_int), allocatable, target, save :: idb(:), npt_b(:), link_b(:,:)
contains
  integer(c_int) function env_mode()
    character(len=16) :: v
    integer :: l, st
    call get_environment_variable('SIMOX_MODE', v, l, st)
    env_mode = SIMOX_MODE_FAITHFUL
    if (st == 0) then
      if (v(1:l) == 'fp64') env_mode = SIMOX_MODE_FP64
      if (v(1:l) == 'mixed') env_mode = SIMOX_MODE_MIXED
    end if
  end function env_mode

  ! Creates the context on first use (or when the atom count changed).  The
  ! atoms are binned at the positions of the caller's last CELL_SORT when one
  ! ran before (kept in xsort), else at their current positions.
  subroutine simox_dropin_ensure(nset, sys, q)
    type(amnt) :: nset
    type(stat) :: sys
    type(ptcl) :: q(nset%npt_all)
    integer :: i, n
    integer(c_int) :: rc
    real(c_double) :: box(3)
    integer(c_int) :: fft(3)
    n = nset%npt_all
    if (c_associated(ctx) .and. n == n_ctx) return
    if (c_associated(ctx)) then
      rc = simox_destroy(ctx)
    else
      call simox_potential_setup(cur)
    end if
    box = nset%box
    fft = nset%fft
    rc = simox_create(int(n, c_int), box, fft, real(sys%rcut, c_double), env_mode(), &
                      c_loc(cur), ctx)
    call simox_check(ctx, rc)
    rc = simox_set_rdf(ctx, 1)   ! FORCE also fills pair%rdf
    call simox_check(ctx, rc)
    if (allocated(xb)) deallocate(xb, fb, qb, idb, vb)
    allocate(xb(3, n), fb(3, n), qb(n), idb(n), vb(3, n))
    n_ctx = n
    do i = 1, n
      xb(:, i) = q(i)%xx(:)
      fb(:, i) = q(i)%ff(:)
      qb(i) = q(i)%q
      idb(i) = q(i)%id
    end do
    rc = -1
    if (allocated(xsort)) then
      if (size(xsort, 2) == n) then
        ! bin at the last CELL_SORT's positions, then move to the current ones
        rc = simox_set_state(ctx, c_loc(idb), c_loc(xsort), c_null_ptr, c_loc(fb), c_loc(qb))
        call simox_check(ctx, rc)
        rc = simox_set_state(ctx, c_null_ptr, c_loc(xb), c_null_ptr, c_null_ptr, c_null_ptr)
        call simox_check(ctx, rc)
      end if
      deallocate(xsort)
    end if
    if (rc /= 0) then
      rc = simox_set_state(ctx, c_loc(idb), c_loc(xb), c_null_ptr, c_loc(fb), c_loc(qb))
      call simox_check(ctx, rc)
    end if
  end subroutine simox_dropin_ensure

  ! Uploads positions and charges (and forces if with_f) without re-binning.
  subroutine simox_dropin_upload(nset, q, with_f)
    type(amnt) :: nset
    type(ptcl) :: q(nset%npt_all)
    logical, intent(in) :: with_f
    integer :: i
    integer(c_int) :: rc
    do i = 1, nset%npt_all
      xb(:, i) = q(i)%xx(:)
      fb(:, i) = q(i)%ff(:)
      qb(i) = q(i)%q
    end do
    if (with_f) then
      rc = simox_set_state(ctx, c_null_ptr, c_loc(xb), c_null_ptr, c_loc(fb), c_loc(qb))
    else
      rc = simox_set_state(ctx, c_null_ptr, c_loc(xb), c_null_ptr, c_null_ptr, c_loc(qb))
    end if
    call simox_check(ctx, rc)
  end subroutine simox_dropin_upload

  ! Sends p to the device if it differs from the parameters in use.
  subroutine simox_dropin_params_update(p)
    type(c_params), target :: p
    integer(c_int) :: rc
    real(c_double), pointer :: a(:), b(:)
    integer :: nd
    nd = int(c_sizeof(p) / 8)
    call c_f_pointer(c_loc(p), a, [nd])
    call c_f_pointer(c_loc(cur), b, [nd])
    if (all(a == b)) return
    cur = p
    rc = simox_set_params(ctx, c_loc(cur))
    call simox_check(ctx, rc)
  end subroutine simox_dropin_params_update

  ! DATAFMT's POTN array into the C layout (fafb transposed).
  subroutine simox_dropin_potn(param, p)
    type(potn) :: param(natom)
    type(c_params) :: p
    integer :: m, k
    do m = 1, 5
      p%prm(m)%xm = param(m)%xm; p%prm(m)%qmin = param(m)%qmin; p%prm(m)%qmax = param(m)%qmax
      p%prm(m)%chi = param(m)%chi; p%prm(m)%jj = param(m)%j; p%prm(m)%xi = param(m)%xi
      p%prm(m)%zz = param(m)%z
      do k = 1, 4
        p%prm(m)%fafb(k, :) = param(m)%fafb(:, k)
      end do
    end do
  end subroutine simox_dropin_potn

  ! DATAFMT's METL / OXID / POTN into the C layout.
  subroutine simox_dropin_params(eam, oxygn, param, p)
    type(metl) :: eam(nmetal)
    type(oxid) :: oxygn
    type(potn) :: param(natom)
    type(c_params) :: p
    integer :: m
    do m = 1, 3
      p%eam(m)%r_e = eam(m)%r_e; p%eam(m)%f_e = eam(m)%f_e
      p%eam(m)%rho_e = eam(m)%rho_e; p%eam(m)%rho_s = eam(m)%rho_s
      p%eam(m)%alpha = eam(m)%alpha; p%eam(m)%beta = eam(m)%beta
      p%eam(m)%aa = eam(m)%a; p%eam(m)%bb = eam(m)%b
      p%eam(m)%kappa = eam(m)%kappa; p%eam(m)%lambda = eam(m)%lambda
      p%eam(m)%fn = eam(m)%fn; p%eam(m)%fm = eam(m)%fm; p%eam(m)%fp = eam(m)%fp
      p%eam(m)%eta = eam(m)%eta; p%eam(m)%fe = eam(m)%fe
    end do
    p%ox%r_e = oxygn%r_e; p%ox%alph = oxygn%alph; p%ox%beta = oxygn%beta
    p%ox%aa = oxygn%a; p%ox%bb = oxygn%b; p%ox%kapp = oxygn%kapp; p%ox%lamb = oxygn%lamb
    p%ox%f(:, 1) = oxygn%f0; p%ox%f(:, 2) = oxygn%f1; p%ox%f(:, 3) = oxygn%f2
    p%ox%f(:, 4) = oxygn%f3; p%ox%f(:, 5) = oxygn%f4
    p%ox%re = oxygn%re; p%ox%rmax = oxygn%rmax
    p%ox%fe = oxygn%fe; p%ox%gamm = oxygn%gamm; p%ox%nu = oxygn%nu
    call simox_dropin_potn(param, p)
  end subroutine simox_dropin_params

end module simox_dropin_state

! CELL_HIERARCHY (cellsort.f90:3-36): the 13 half-shell neighbour cells of
! every cell, periodic, in the reference's order (offsets dx fastest, then
! dy, then dz = -1, followed by the four dz = 0 cells).
subroutine cell_hierarchy(nset, cell)
  use datafmt
  implicit none
  type(amnt) :: nset
  type(sort) :: cell(nset%ncell_all)
  integer :: i, j, k, m, nc(3), id_cell
  integer, parameter :: off(3, 13) = reshape([ &
       -1, -1, -1,   0, -1, -1,   1, -1, -1,  -1, 0, -1,   0, 0, -1,   1, 0, -1, &
       -1,  1, -1,   0,  1, -1,   1,  1, -1,  -1, -1, 0,   0, -1, 0,   1, -1, 0, &
       -1,  0,  0], [3, 13])
  nc = nset%ncell
  do i = 1, nc(1)
    do j = 1, nc(2)
      do k = 1, nc(3)
        do m = 1, 13
          cell(id_cell(nc, i, j, k))%pair(m) = id_cell(nc, i + off(1, m), j + off(2, m), k + off(3, m))
        end do
      end do
    end do
  end do
end subroutine cell_hierarchy

! ID_CELL (cellsort.f90:39-50): 1-based periodic cell id of 1-based (i,j,k).
integer function id_cell(nc, i, j, k)
  implicit none
  integer :: nc(3), i, j, k
  id_cell = modulo(i - 1, nc(1)) + modulo(j - 1, nc(2)) * nc(1) + &
            modulo(k - 1, nc(3)) * nc(1) * nc(2) + 1
end function id_cell

! CELL_SORT (cellsort.f90:53-80): bins the atoms on the device and fills
! cell(:)%Npt and cell(:)%link (1-based, ascending atom order).  Before the
! context exists (main.f90:31 runs before the first call that carries the
! cutoff) the binning is done here on the host and the positions are kept, so
! the context starts from the same cells.
subroutine cell_sort(nset, cell, q)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(sort) :: cell(nset%ncell_all)
  type(ptcl) :: q(nset%npt_all)
  integer(c_int) :: rc
  integer :: c, i, ic(3)
  if (.not. c_associated(ctx) .or. nset%npt_all /= n_ctx) then
    if (allocated(xsort)) deallocate(xsort)
    allocate(xsort(3, nset%npt_all))
    cell(:)%npt = 0
    do i = 1, nset%npt_all
      xsort(:, i) = q(i)%xx(:)
      ic = modulo(int((q(i)%xx(:) + 0.5d0 * nset%box(:)) / nset%lcell(:)), nset%ncell(:))
      c = ic(1) + ic(2) * nset%ncell(1) + ic(3) * nset%ncell(1) * nset%ncell(2) + 1
      cell(c)%npt = cell(c)%npt + 1
      if (cell(c)%npt > npt_max) stop "=== Cell particle limit crashed !!! ==="
      cell(c)%link(cell(c)%npt) = i
    end do
    return
  end if
  call simox_dropin_upload(nset, q, .false.)
  rc = simox_cell_sort(ctx)
  call simox_check(ctx, rc)
  if (allocated(npt_b)) then
    if (size(npt_b) /= nset%ncell_all) deallocate(npt_b, link_b)
  end if
  if (.not. allocated(npt_b)) allocate(npt_b(nset%ncell_all), link_b(npt_max, nset%ncell_all))
  rc = simox_get_cells(ctx, c_loc(npt_b), c_loc(link_b), int(npt_max, c_int))
  call simox_check(ctx, rc)
  do c = 1, nset%ncell_all
    cell(c)%npt = npt_b(c)
    cell(c)%link(1:npt_b(c)) = link_b(1:npt_b(c), c) + 1
  end do
end subroutine cell_sort

! FORCE (force.f90:3): overwrites q%ff, sets Epair, Edir, Ectip, Egroup.
subroutine force(nset, eam, oxygn, sys, param, pair, q, cell)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(metl) :: eam(nmetal)
  type(oxid) :: oxygn
  type(stat) :: sys
  type(potn) :: param(natom)
  type(corr) :: pair
  type(ptcl) :: q(nset%npt_all)
  type(sort) :: cell(nset%ncell_all)
  type(c_params), target, save :: pc
  real(c_double) :: e(4)
  integer(c_int) :: rc
  integer :: i
  integer(c_long_long), target :: hist(5, 5, 201)
  call simox_dropin_ensure(nset, sys, q)
  call simox_dropin_params(eam, oxygn, param, pc)
  call simox_dropin_params_update(pc)
  call simox_dropin_upload(nset, q, .false.)
  rc = simox_force(ctx, e)
  call simox_check(ctx, rc)
  rc = simox_get_state(ctx, c_null_ptr, c_null_ptr, c_loc(fb), c_null_ptr)
  call simox_check(ctx, rc)
  do i = 1, nset%npt_all
    q(i)%ff(:) = fb(:, i)
  end do
  sys%epair = e(1); sys%edir = e(2); sys%ectip = e(3); sys%egroup = e(4)
  pair%npt_new = nset%npt_all
  ! pair histogram: counts per pair direction; PRINT_RDF reads rdf(i,j) + rdf(j,i)
  rc = simox_get_rdf(ctx, hist, 1)
  call simox_check(ctx, rc)
  pair%rdf(:, :, :) = pair%rdf(:, :, :) + int(hist(:, :, :), gi)
end subroutine force

! PME (force.f90:364): adds the reciprocal forces to q%ff, sets Erec.
subroutine pme(nset, sys, q, psi)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(stat) :: sys
  type(ptcl) :: q(nset%npt_all)
  real(dp) :: psi(nset%fft(1), nset%fft(2), nset%fft(3))
  real(c_double) :: erec
  integer(c_int) :: rc
  integer :: i
  call simox_dropin_ensure(nset, sys, q)
  call simox_dropin_upload(nset, q, .true.)
  rc = simox_pme(ctx, erec)
  call simox_check(ctx, rc)
  rc = simox_get_state(ctx, c_null_ptr, c_null_ptr, c_loc(fb), c_null_ptr)
  call simox_check(ctx, rc)
  do i = 1, nset%npt_all
    q(i)%ff(:) = fb(:, i)
  end do
  sys%erec = erec
end subroutine pme

! PME_SELF (force.f90:467): sets sys%a and Eself.  The device keeps its own
! influence function; psi is left untouched.
subroutine pme_self(nset, sys, param, q, psi)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(stat) :: sys
  type(potn) :: param(natom)
  type(ptcl) :: q(nset%npt_all)
  real(dp) :: psi(nset%fft(1), nset%fft(2), nset%fft(3))
  type(c_params), target, save :: pc
  real(c_double) :: e
  integer(c_int) :: rc
  sys%a = 3.8d0 / sys%rcut
  call simox_dropin_ensure(nset, sys, q)
  pc = cur
  call simox_dropin_potn(param, pc)
  call simox_dropin_params_update(pc)
  call simox_dropin_upload(nset, q, .false.)
  rc = simox_pme_self(ctx, e)
  call simox_check(ctx, rc)
  sys%eself = e
end subroutine pme_self

! CTIP_SOLVER (ctip.f90:3): overwrites q%q, sets Eself.
subroutine ctip_solver(nset, param, sys, q, cell, psi)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(potn) :: param(natom)
  type(stat) :: sys
  type(ptcl) :: q(nset%npt_all)
  type(sort) :: cell(nset%ncell_all)
  real(dp) :: psi(nset%fft(1), nset%fft(2), nset%fft(3))
  type(c_params), target, save :: pc
  real(c_double) :: e
  integer(c_int) :: rc, it, nev
  integer :: i
  call simox_dropin_ensure(nset, sys, q)
  pc = cur
  call simox_dropin_potn(param, pc)
  call simox_dropin_params_update(pc)
  call simox_dropin_upload(nset, q, .false.)
  rc = simox_ctip_solve(ctx, e, it, nev)
  call simox_check(ctx, rc)
  rc = simox_get_state(ctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(qb))
  call simox_check(ctx, rc)
  do i = 1, nset%npt_all
    q(i)%q = qb(i)
  end do
  sys%eself = e
end subroutine ctip_solver

! VV_NVE1 (vverlet.f90:74-96): v += dt/2 F/m, x += dt v, wrapped into the box,
! on the device; q%xx and q%xv are updated.  The context exists by then
! (PME_SELF runs before VINIT and the loop, main.f90:33-45).
subroutine vv_nve1(nset, param, q, dt)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(potn) :: param(natom)
  type(ptcl) :: q(nset%npt_all)
  real(dp) :: dt
  type(c_params), target, save :: pc
  integer(c_int) :: rc
  integer :: i
  if (.not. c_associated(ctx) .or. nset%npt_all /= n_ctx) then
    write(*, *) "=== VV_NVE1 drop-in: no device context for these atoms ==="
    stop
  end if
  pc = cur
  call simox_dropin_potn(param, pc)
  call simox_dropin_params_update(pc)
  do i = 1, nset%npt_all
    xb(:, i) = q(i)%xx(:)
    vb(:, i) = q(i)%xv(:)
    fb(:, i) = q(i)%ff(:)
  end do
  rc = simox_set_state(ctx, c_null_ptr, c_loc(xb), c_loc(vb), c_loc(fb), c_null_ptr)
  call simox_check(ctx, rc)
  rc = simox_vv1(ctx, real(dt, c_double))
  call simox_check(ctx, rc)
  rc = simox_get_state(ctx, c_loc(xb), c_loc(vb), c_null_ptr, c_null_ptr)
  call simox_check(ctx, rc)
  do i = 1, nset%npt_all
    q(i)%xx(:) = xb(:, i)
    q(i)%xv(:) = vb(:, i)
  end do
end subroutine vv_nve1

! VV_NVE2 (vverlet.f90:132-157): v += dt/2 F/m on the device, sys%mv2 = sum m v^2.
subroutine vv_nve2(nset, param, sys, q, dt)
  use datafmt
  use simox_dropin_state
  implicit none
  type(amnt) :: nset
  type(potn) :: param(natom)
  type(stat) :: sys
  type(ptcl) :: q(nset%npt_all)
  real(dp) :: dt
  type(c_params), target, save :: pc
  real(c_double) :: mv2
  integer(c_int) :: rc
  integer :: i
  call simox_dropin_ensure(nset, sys, q)
  pc = cur
  call simox_dropin_potn(param, pc)
  call simox_dropin_params_update(pc)
  do i = 1, nset%npt_all
    vb(:, i) = q(i)%xv(:)
    fb(:, i) = q(i)%ff(:)
  end do
  rc = simox_set_state(ctx, c_null_ptr, c_null_ptr, c_loc(vb), c_loc(fb), c_null_ptr)
  call simox_check(ctx, rc)
  rc = simox_vv2(ctx, real(dt, c_double), mv2)
  call simox_check(ctx, rc)
  rc = simox_get_state(ctx, c_null_ptr, c_loc(vb), c_null_ptr, c_null_ptr)
  call simox_check(ctx, rc)
  do i = 1, nset%npt_all
    q(i)%xv(:) = vb(:, i)
  end do
  sys%mv2 = mv2
end subroutine vv_nve2
