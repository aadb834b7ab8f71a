! The reference's Fortran test (tests/test_fortran_rbf_interface.f90): 20^3
! grid interpolated onto itself through module pacman_mod, rel-L2 < 1e-8
! for every Wendland basis.
program test_fortran_rbf_interface
  use, intrinsic :: iso_c_binding
  use pacman_mod
  implicit none
  integer, parameter :: g = 20, n = g * g * g
  real(c_double), allocatable :: pts(:, :), vals(:), out(:)
  integer :: i, j, k, p, ierr, b, failures
  integer(c_int) :: bases(5)
  real(c_double) :: h, x, y, z, err
  bases = [PACMAN_WENDLANDC0, PACMAN_WENDLANDC2, PACMAN_WENDLANDC4, &
           PACMAN_WENDLANDC6, PACMAN_WENDLANDC8]
  allocate(pts(3, n), vals(n), out(n))
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
  if (pacman_best_execspace() /= PACMAN_HIP) failures = failures + 1
  do b = 1, 5
    out = 0
    call pacman_rbf_interpolate(3, PACMAN_HIP, bases(b), pts, vals, pts, out, ierr)
    err = sqrt(sum((out - vals)**2) / sum(vals**2))
    if (ierr /= 0 .or. err >= 1d-8) then
      failures = failures + 1
      print '(a,i3,a,es10.3)', 'FAIL basis ', bases(b), ' rel_l2=', err
    else
      print '(a,i3,a,es10.3)', 'PASS basis ', bases(b), ' rel_l2=', err
    end if
  end do
  call pacman_kokkos_finalize()
  if (failures /= 0) stop 1
end program
