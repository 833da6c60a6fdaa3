! OXIDATION-mode host routines of simox_md: the new-dimer test and the dimer
! placement, on plain arrays (no device calls), so that the host program and
! its tests share them.
module simox_oxid
  use iso_c_binding, only: c_double, c_int
  implicit none
  private
  public :: oxide_check, o2_dimers

contains

  ! OXIDE_CHECK (util.f90:152-188): true when the highest Al (species 1) is
  ! above the highest of the last 2 n_o2 atoms that are O (species 4), or when
  ! the largest charge among those O is below q_crit.  Heights start at 0 and
  ! the largest charge at -10, as there.
  logical function oxide_check(n, id, x, q, n_o2, q_crit)
    integer, intent(in) :: n, n_o2
    integer(c_int), intent(in) :: id(n)
    real(c_double), intent(in) :: x(3, n), q(n), q_crit
    real(c_double) :: m_height, o_height, q_max
    integer :: k
    m_height = 0.0d0
    do k = 1, n
      if (id(k) == 1) m_height = max(x(3, k), m_height)
    end do
    o_height = 0.0d0
    q_max = -10.0d0
    do k = n, n - 2 * n_o2 + 1, -1
      if (id(k) == 4) then
        o_height = max(x(3, k), o_height)
        q_max = max(q(k), q_max)
      end if
    end do
    oxide_check = m_height > o_height .or. q_max < q_crit
  end function oxide_check

  ! NEW_O2 (datafmt.f90:98-139) with RANDOM_POSITION_O2 (util.f90:104-149):
  ! n_o2 dimers centred at height z_new, centres at least sqrt(Lx Ly / n_o2) / 2
  ! apart (rejection sampling), atoms at c +- d with d = L0 (sin(theta)
  ! cos(phi), sin(theta) cos(phi), cos(theta)), L0 = 2 A, as there (the y
  ! component repeats cos(phi), so |d| is not L0 in general), and
  ! v = (0, 0, -sqrt(3 T_given / 16)) with T_given in eV.  Random numbers are drawn in the reference's order:
  ! the centres first, then (phi, theta) per dimer.  xo(:, 2k-1) = c_k + d_k,
  ! xo(:, 2k) = c_k - d_k.  Stops with the reference's text when the dimers
  ! cannot be that far apart.
  subroutine o2_dimers(n_o2, z_new, lx, ly, t_given_ev, xo, vo)
    integer, intent(in) :: n_o2
    real(c_double), intent(in) :: z_new, lx, ly, t_given_ev
    real(c_double), intent(out) :: xo(3, 2 * n_o2), vo(3, 2 * n_o2)
    real(c_double), parameter :: pi = 3.1415926535897931d0, l0 = 2.0d0
    real(c_double) :: cx(n_o2, 3), r, rmin2, phi, theta, d(3)
    integer :: k, j
    logical :: close_pair
    if (n_o2 <= 0) return
    rmin2 = (sqrt(lx * ly / dble(n_o2)) * 0.5d0)**2
    if (rmin2 < l0**2) stop "=== Too many new oxygens ==="
    do k = 1, n_o2
      close_pair = .true.
      do while (close_pair)
        call random_number(r); cx(k, 1) = (r - 0.5d0) * lx
        call random_number(r); cx(k, 2) = (r - 0.5d0) * ly
        close_pair = .false.
        do j = 1, k - 1
          if ((cx(k, 1) - cx(j, 1))**2 + (cx(k, 2) - cx(j, 2))**2 < rmin2) close_pair = .true.
        end do
      end do
    end do
    cx(:, 3) = z_new
    do k = 1, n_o2
      call random_number(r); phi = r * pi * 2.0d0
      call random_number(r); theta = r * pi
      d = [l0 * sin(theta) * cos(phi), l0 * sin(theta) * cos(phi), l0 * cos(theta)]
      xo(:, 2 * k - 1) = cx(k, :) + d
      xo(:, 2 * k) = cx(k, :) - d
    end do
    do k = 1, 2 * n_o2
      vo(:, k) = [0.0d0, 0.0d0, -sqrt(t_given_ev * 3.0d0 / 16.0d0)]
    end do
  end subroutine o2_dimers

end module simox_oxid
