! iso_c_binding interface to libsimox_gpu.so (include/simox_gpu.h): the
! MI355X implementation of the SIMOX hot path.  Every function returns an
! integer status (0 = success); simox_error_text() returns the message the
! reference would print before its STOP.
module simox_gpu
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: SIMOX_MODE_FAITHFUL = 0
  integer(c_int), parameter :: SIMOX_MODE_FP64 = 1
  integer(c_int), parameter :: SIMOX_MODE_MIXED = 2

  ! simox_params (include/simox_gpu.h); the C array fafb[5][4] is fafb(4,5) here
  type, bind(c) :: c_metal
    real(c_double) :: r_e, f_e, rho_e, rho_s, alpha, beta, aa, bb, kappa, lambda
    real(c_double) :: fn(4), fm(4), fp(4), eta, fe
  end type c_metal
  type, bind(c) :: c_oxid
    real(c_double) :: r_e(5), alph(5), beta(5), aa(5), bb(5), kapp(5), lamb(5)
    real(c_double) :: f(4, 5), re(5), rmax(5), fe, gamm, nu
  end type c_oxid
  type, bind(c) :: c_potn
    real(c_double) :: xm, qmin, qmax, chi, jj, xi, zz, fafb(4, 5)
  end type c_potn
  type, bind(c) :: c_params
    type(c_metal) :: eam(3)
    type(c_oxid) :: ox
    type(c_potn) :: prm(5)
  end type c_params

  interface
    integer(c_int) function simox_create(n, box, fft, rcut, mode, params, ctx) &
        bind(c, name='simox_create')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n
      real(c_double), intent(in) :: box(3)
      integer(c_int), intent(in) :: fft(3)
      real(c_double), value :: rcut
      integer(c_int), value :: mode
      type(c_ptr), value :: params
      type(c_ptr), intent(out) :: ctx
    end function simox_create

    integer(c_int) function simox_destroy(ctx) bind(c, name='simox_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function simox_destroy

    type(c_ptr) function simox_last_error(ctx) bind(c, name='simox_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function simox_last_error

    integer(c_int) function simox_set_charge_tol(ctx, tol, maxit) &
        bind(c, name='simox_set_charge_tol')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: tol
      integer(c_int), value :: maxit
    end function simox_set_charge_tol

    ! any pointer may be c_null_ptr
    integer(c_int) function simox_set_state(ctx, id, x, v, f, q) bind(c, name='simox_set_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, id, x, v, f, q
    end function simox_set_state

    integer(c_int) function simox_get_state(ctx, x, v, f, q) bind(c, name='simox_get_state')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, v, f, q
    end function simox_get_state

    integer(c_int) function simox_cell_sort(ctx) bind(c, name='simox_cell_sort')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function simox_cell_sort

    ! npt(Ncell_all); link(link_stride, Ncell_all) 0-based caller indices, or c_null_ptr
    integer(c_int) function simox_get_cells(ctx, npt, link, link_stride) &
        bind(c, name='simox_get_cells')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, npt, link
      integer(c_int), value :: link_stride
    end function simox_get_cells

    integer(c_int) function simox_set_thermostat(ctx, kind, t_kelvin, alpha, seed) &
        bind(c, name='simox_set_thermostat')
      import :: c_int, c_double, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      real(c_double), value :: t_kelvin, alpha
      integer(c_long_long), value :: seed
    end function simox_set_thermostat

    integer(c_int) function simox_set_reflect(ctx, on) bind(c, name='simox_set_reflect')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function simox_set_reflect

    integer(c_int) function simox_remove_rigid_motion(ctx) bind(c, name='simox_remove_rigid_motion')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function simox_remove_rigid_motion

    integer(c_int) function simox_set_rdf(ctx, on) bind(c, name='simox_set_rdf')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function simox_set_rdf

    ! hist(5, 5, 201): counts per pair direction (see include/simox_gpu.h)
    integer(c_int) function simox_get_rdf(ctx, hist, reset) bind(c, name='simox_get_rdf')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: hist(5, 5, 201)
      integer(c_int), value :: reset
    end function simox_get_rdf

    integer(c_int) function simox_set_eam_cutoff(ctx, r_eam) bind(c, name='simox_set_eam_cutoff')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: r_eam
    end function simox_set_eam_cutoff

    integer(c_int) function simox_set_params(ctx, params) bind(c, name='simox_set_params')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, params
    end function simox_set_params

    integer(c_int) function simox_force(ctx, energies) bind(c, name='simox_force')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: energies(4)
    end function simox_force

    integer(c_int) function simox_pme_self(ctx, eself) bind(c, name='simox_pme_self')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: eself
    end function simox_pme_self

    integer(c_int) function simox_pme(ctx, erec) bind(c, name='simox_pme')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: erec
    end function simox_pme

    integer(c_int) function simox_ctip_solve(ctx, eself, iters, n_eval) &
        bind(c, name='simox_ctip_solve')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: eself
      integer(c_int), intent(out) :: iters, n_eval
    end function simox_ctip_solve

    integer(c_int) function simox_vv1(ctx, dt) bind(c, name='simox_vv1')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function simox_vv1

    integer(c_int) function simox_vv2(ctx, dt, mv2) bind(c, name='simox_vv2')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), intent(out) :: mv2
    end function simox_vv2

    ! stats: c_loc of real(c_double) :: s(8), or c_null_ptr
    integer(c_int) function simox_step(ctx, nsteps, dt, sort_every, ctip_every, stats) &
        bind(c, name='simox_step')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, sort_every, ctip_every
      real(c_double), value :: dt
      type(c_ptr), value :: stats
    end function simox_step

    subroutine simox_potential_setup(p) bind(c, name='simox_potential_setup')
      import :: c_params
      type(c_params), intent(out) :: p
    end subroutine simox_potential_setup
  end interface

contains

  ! message of the last failed call, as a Fortran string
  function simox_error_text(ctx) result(txt)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: txt
    character(kind=c_char), pointer :: s(:)
    integer :: k
    txt = ' '
    call c_f_pointer(simox_last_error(ctx), s, [256])
    do k = 1, 256
      if (s(k) == c_null_char) exit
      txt(k:k) = s(k)
    end do
  end function simox_error_text

  ! STOP with the library's text on a nonzero status (the reference STOPs)
  subroutine simox_check(ctx, rc)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
      write(*, '(A)') trim(simox_error_text(ctx))
      stop 1
    end if
  end subroutine simox_check

end module simox_gpu
