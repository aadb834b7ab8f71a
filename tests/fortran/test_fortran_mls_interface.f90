! The reference's Fortran MLS test (tests/test_fortran_mls_interface.f90):
! a 20^3 grid onto itself through module pacman_mod, rel-L2 < 1e-4 against
! the Franke field; plus the nearest-source FE transfer on the same grid,
! which must return the values exactly with status 0 (OUTSIDE).
program test_fortran_mls_interface
  use, intrinsic :: iso_c_binding
  use pacman_mod
  implicit none
  integer, parameter :: g = 20, n = g * g * g
  real(c_double), allocatable :: pts(:, :), vals(:), out(:)
  integer(c_int), allocatable :: status(:), conn(:)
  integer :: i, j, k, p, failures
  integer(c_int) :: ierr
  real(c_double) :: h, x, y, z, err
  allocate(pts(3, n), vals(n), out(n), status(n), conn(1))
  conn = 0
  h = 1.0d0 / (g - 1)
  p = 0
  do i = 0, g - 1
    do j = 0, g - 1
      do k = 0, g - 1
        p = p + 1
        x = i * h; y = j * h; z = k * h
        pts(:, p) = [x, y, z]
        vals(p) = 0.75d0 * exp(-((9*x-2)**2 + (9*y-2)**2 + (9*z-2)**2) / 4d0) &
                + 0.75d0 * exp(-((9*x+1)**2 / 49d0 + (9*y+1) / 10d0 + (9*z+1) / 10d0)) &
                + 0.5d0 * exp(-((9*x-7)**2 + (9*y-3)**2 + (9*z-5)**2) / 4d0) &
                - 0.2d0 * exp(-((9*x-4)**2 + (9*y-7)**2 + (9*z-5)**2))
      end do
    end do
  end do
  call pacman_kokkos_initialize()
  failures = 0
  out = 0
  call pacman_mls_interpolate(3, PACMAN_HIP, pts, vals, pts, out, ierr)
  err = sqrt(sum((out - vals)**2) / sum(vals**2))
  if (ierr /= 0 .or. err >= 1d-4) then
    failures = failures + 1
    print '(a,es10.3)', 'FAIL mls same rel_l2=', err
  else
    print '(a,es10.3)', 'PASS mls same rel_l2=', err
  end if
  out = -1
  status = -1
  call pacman_fe_interpolate(3, PACMAN_HIP, PACMAN_FE_NEAREST_NEAREST, pts, &
                             vals, conn, conn, conn, pts, out, status, ierr)
  if (ierr /= 0 .or. any(out /= vals) .or. any(status /= 0)) then
    failures = failures + 1
    print '(a)', 'FAIL fe nearest same grid'
  else
    print '(a)', 'PASS fe nearest same grid'
  end if
  call pacman_mls_interpolate(3, PACMAN_SERIAL, pts, vals, pts, out, ierr)
  if (ierr == 0) then
    failures = failures + 1
    print '(a)', 'FAIL mls accepted an execution space it does not serve'
  end if
  call pacman_kokkos_finalize()
  if (failures /= 0) stop 1
end program
