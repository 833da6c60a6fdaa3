! Test driver of simox_amd/fortran/simox_oxid.f90: OXIDE_CHECK for each q_crit
! and the NEW_O2 dimers after RANDOM_SEED(put = (seed, 1, 1, ...)), the way
! the reference driver (oracle/ref_driver.f90) runs the reference's own.
! Input (stream): int32 n, n_o2, seed, ncrit; real64 z_new, lx, ly, T (eV);
! real64 q_crit(ncrit); int32 id(n); real64 x(3, n), q(n).
! Output (stream): int32 new_o2(ncrit); real64 x(3, 2 n_o2), v(3, 2 n_o2).
program oxid_driver
  use iso_c_binding, only: c_int, c_double
  use simox_oxid, only: oxide_check, o2_dimers
  implicit none
  integer(c_int) :: n, n_o2, seedv, ncrit
  integer :: k, nseed
  integer(c_int), allocatable :: id(:)
  integer, allocatable :: seed(:)
  real(c_double), allocatable :: x(:,:), q(:), qc(:), xo(:,:), vo(:,:)
  real(c_double) :: z_new, lx, ly, t_ev
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n, n_o2, seedv, ncrit
  allocate(id(n), x(3, n), q(n), qc(ncrit), xo(3, 2 * n_o2), vo(3, 2 * n_o2))
  read(10) z_new, lx, ly, t_ev
  read(10) qc
  read(10) id
  read(10) x
  read(10) q
  close(10)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  do k = 1, ncrit
    write(11) merge(1_c_int, 0_c_int, oxide_check(int(n), id, x, q, int(n_o2), qc(k)))
  end do
  call random_seed(size=nseed)
  allocate(seed(nseed))
  seed = 1
  seed(1) = seedv
  call random_seed(put=seed)
  call o2_dimers(int(n_o2), z_new, lx, ly, t_ev, xo, vo)
  write(11) xo
  write(11) vo
  close(11)
end program oxid_driver
