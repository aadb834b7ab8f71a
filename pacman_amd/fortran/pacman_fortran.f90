! Fortran 2003 binding: module pacman_mod with the same constants and
! procedure names as the reference's src/pacman_fortran.f90 (constants
! :47-77, interface block :97-160, pacman_rbf_interpolate :215-235,
! pacman_fe_interpolate :254-293, pacman_mls_interpolate :310-330), bound
! through ISO_C_BINDING to the C ABI of libpacman_rbf.so
! (include/fortran_interface.h). The VTK cell-type helpers are not included;
! FE serves the method PACMAN_FE_NEAREST_NEAREST only (the others return an
! error code).
module pacman_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: PACMAN_SERIAL = 0
  integer(c_int), parameter, public :: PACMAN_OPENMP = 1
  integer(c_int), parameter, public :: PACMAN_THREADS = 2
  integer(c_int), parameter, public :: PACMAN_CUDA = 3
  integer(c_int), parameter, public :: PACMAN_HIP = 4
  integer(c_int), parameter, public :: PACMAN_SYCL = 5
  integer(c_int), parameter, public :: PACMAN_WENDLANDC0 = 16
  integer(c_int), parameter, public :: PACMAN_WENDLANDC2 = 17
  integer(c_int), parameter, public :: PACMAN_WENDLANDC4 = 18
  integer(c_int), parameter, public :: PACMAN_WENDLANDC6 = 19
  integer(c_int), parameter, public :: PACMAN_WENDLANDC8 = 20
  integer(c_int), parameter, public :: PACMAN_FE_NEAREST_NEAREST = 240
  integer(c_int), parameter, public :: PACMAN_FE_INTERP_CLAMP = 241
  integer(c_int), parameter, public :: PACMAN_FE_INTERP_NEAREST = 242
  integer(c_int), parameter, public :: PACMAN_FE_INTERP_ZEROFILL = 243
  integer(c_int), parameter, public :: PACMAN_FE_INTERP_EXTRAP = 244
  integer(c_int), parameter, public :: PACMAN_MLS = 246

  public :: pacman_kokkos_initialize, pacman_kokkos_finalize
  public :: pacman_best_execspace, pacman_rbf_interpolate
  public :: pacman_fe_interpolate, pacman_mls_interpolate

  interface
    subroutine c_initialize() bind(C, name="pacman_kokkos_initialize_c")
    end subroutine
    subroutine c_finalize() bind(C, name="pacman_kokkos_finalize_c")
    end subroutine
    function c_best() result(e) bind(C, name="pacman_best_execspace_c")
      import :: c_int
      integer(c_int) :: e
    end function
    function c_rbf(dim, es, fn, sp, ns, sv, tp, nt, tv) result(rc) &
        bind(C, name="pacman_rbf_interpolate_c")
      import :: c_int, c_ptr
      integer(c_int), value :: dim, es, fn, ns, nt
      type(c_ptr), value :: sp, sv, tp, tv
      integer(c_int) :: rc
    end function
    function c_fe(dim, es, method, sp, ns, sv, cv, ncv, co, nco, ct, tp, nt, &
                  tv, ts) result(rc) bind(C, name="pacman_fe_interpolate_c")
      import :: c_int, c_ptr
      integer(c_int), value :: dim, es, method, ns, ncv, nco, nt
      type(c_ptr), value :: sp, sv, cv, co, ct, tp, tv, ts
      integer(c_int) :: rc
    end function
    function c_mls(dim, es, sp, ns, sv, tp, nt, tv) result(rc) &
        bind(C, name="pacman_mls_interpolate_c")
      import :: c_int, c_ptr
      integer(c_int), value :: dim, es, ns, nt
      type(c_ptr), value :: sp, sv, tp, tv
      integer(c_int) :: rc
    end function
  end interface

contains

  subroutine pacman_kokkos_initialize()
    call c_initialize()
  end subroutine

  subroutine pacman_kokkos_finalize()
    call c_finalize()
  end subroutine

  integer function pacman_best_execspace()
    pacman_best_execspace = int(c_best())
  end function

  ! points are shaped (spaceDimension, n): column-major storage equals the C
  ! row-major [n][spaceDimension] layout.
  subroutine pacman_rbf_interpolate(spaceDimension, execSpace, rbfFunction, &
                                    sourcePoints, sourceValues, targetPoints, &
                                    targetValues, ierr)
    integer, intent(in) :: spaceDimension
    integer(c_int), intent(in) :: execSpace, rbfFunction
    real(c_double), contiguous, target, intent(in) :: sourcePoints(:, :)
    real(c_double), contiguous, target, intent(in) :: sourceValues(:)
    real(c_double), contiguous, target, intent(in) :: targetPoints(:, :)
    real(c_double), contiguous, target, intent(inout) :: targetValues(:)
    integer, optional, intent(out) :: ierr
    integer(c_int) :: rc
    rc = c_rbf(int(spaceDimension, c_int), execSpace, rbfFunction, &
               c_loc(sourcePoints), int(size(sourceValues), c_int), &
               c_loc(sourceValues), c_loc(targetPoints), &
               int(size(targetValues), c_int), c_loc(targetValues))
    if (present(ierr)) ierr = int(rc)
  end subroutine

  ! connVal / connOff / cellTypes are passed on but not read by the method
  ! served here (NEAREST_NEAREST).
  subroutine pacman_fe_interpolate(spaceDimension, execSpace, method, &
                                   sourcePoints, sourceValues, connVal, &
                                   connOff, cellTypes, targetPoints, &
                                   targetValues, targetStatus, ierr)
    integer, intent(in) :: spaceDimension
    integer(c_int), intent(in) :: execSpace, method
    real(c_double), contiguous, target, intent(in) :: sourcePoints(:, :)
    real(c_double), contiguous, target, intent(in) :: sourceValues(:)
    integer(c_int), contiguous, target, intent(in) :: connVal(:)
    integer(c_int), contiguous, target, intent(in) :: connOff(:)
    integer(c_int), contiguous, target, intent(in) :: cellTypes(:)
    real(c_double), contiguous, target, intent(in) :: targetPoints(:, :)
    real(c_double), contiguous, target, intent(out) :: targetValues(:)
    integer(c_int), contiguous, target, intent(out) :: targetStatus(:)
    integer(c_int), optional, intent(out) :: ierr
    integer(c_int) :: rc
    rc = c_fe(int(spaceDimension, c_int), execSpace, method, &
              c_loc(sourcePoints), int(size(sourceValues), c_int), &
              c_loc(sourceValues), c_loc(connVal), int(size(connVal), c_int), &
              c_loc(connOff), int(size(connOff), c_int), c_loc(cellTypes), &
              c_loc(targetPoints), int(size(targetValues), c_int), &
              c_loc(targetValues), c_loc(targetStatus))
    if (present(ierr)) ierr = rc
  end subroutine

  subroutine pacman_mls_interpolate(spaceDimension, execSpace, sourcePoints, &
                                    sourceValues, targetPoints, targetValues, &
                                    ierr)
    integer, intent(in) :: spaceDimension
    integer(c_int), intent(in) :: execSpace
    real(c_double), contiguous, target, intent(in) :: sourcePoints(:, :)
    real(c_double), contiguous, target, intent(in) :: sourceValues(:)
    real(c_double), contiguous, target, intent(in) :: targetPoints(:, :)
    real(c_double), contiguous, target, intent(out) :: targetValues(:)
    integer(c_int), optional, intent(out) :: ierr
    integer(c_int) :: rc
    rc = c_mls(int(spaceDimension, c_int), execSpace, c_loc(sourcePoints), &
               int(size(sourceValues), c_int), c_loc(sourceValues), &
               c_loc(targetPoints), int(size(targetValues), c_int), &
               c_loc(targetValues))
    if (present(ierr)) ierr = rc
  end subroutine

end module pacman_mod
