/*
 * Drop-in C ABI of the RBF-PUM transfer path (MI355X / gfx950 build).
 *
 * Same symbols, argument order, constants and return codes as the
 * reference's plain-C shim (src/fortran_interface.h), so the reference's
 * Fortran module (src/pacman_fortran.f90:97-113, 215-235) and any C caller
 * bind to this library unchanged. Of the FE entry point only the method
 * NEAREST_NEAREST is served (it shares the RBF path's point search); the
 * other FE methods return an error code. The MLS entry point is served.
 */
#ifndef PACMAN_AMD_FORTRAN_INTERFACE_H
#define PACMAN_AMD_FORTRAN_INTERFACE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Execution space selectors (src/fortran_interface.h:33-40). This build
 * implements PACMAN_HIP; the other selectors return an error code. */
enum PACMAN_ExecSpace {
  PACMAN_SERIAL = 0x00,
  PACMAN_OPENMP = 0x01,
  PACMAN_THREADS = 0x02,
  PACMAN_CUDA = 0x03,
  PACMAN_HIP = 0x04,
  PACMAN_SYCL = 0x05
};

/* RBF basis selectors (src/fortran_interface.h:43-49). */
enum PACMAN_RbfFunction {
  PACMAN_WENDLANDC0 = 0x10,
  PACMAN_WENDLANDC2 = 0x11,
  PACMAN_WENDLANDC4 = 0x12,
  PACMAN_WENDLANDC6 = 0x13,
  PACMAN_WENDLANDC8 = 0x14
};

/* FE method selectors (src/fortran_interface.h:51-59). This build serves
 * PACMAN_FE_NEAREST_NEAREST; the other selectors return an error code. */
enum PACMAN_FeMethod {
  PACMAN_FE_NEAREST_NEAREST = 0xF0,
  PACMAN_FE_INTERP_CLAMP = 0xF1,
  PACMAN_FE_INTERP_NEAREST = 0xF2,
  PACMAN_FE_INTERP_ZEROFILL = 0xF3,
  PACMAN_FE_INTERP_EXTRAP = 0xF4,
  PACMAN_MLS = 0xF6
};

/* Replaces pacman_kokkos_initialize_c (src/fortran_interface.h:69,
 * src/fortran_interface.cpp:29-33): selects and initialises the HIP device.
 * Idempotent. */
void pacman_kokkos_initialize_c(void);

/* Replaces pacman_kokkos_finalize_c (src/fortran_interface.h:75,
 * src/fortran_interface.cpp:35-39): releases cached device resources. */
void pacman_kokkos_finalize_c(void);

/* Replaces pacman_best_execspace_c (src/fortran_interface.h:83,
 * src/fortran_interface.cpp:41-57): returns PACMAN_HIP when a HIP device is
 * present, -1 otherwise. */
int pacman_best_execspace_c(void);

/* Replaces pacman_rbf_interpolate_c (src/fortran_interface.h:110-114,
 * src/fortran_interface.cpp:63-91). Row-major host arrays
 * points[n * spaceDimension]; targetValues is caller-allocated
 * (nTargetPoints). Uses the reference's fixed parameters: 50 nodes per
 * cluster, overlap 0.15, Wendland support radius 0.1
 * (src/rbf_pum/interpolator.hxx:39-41). Returns 0 on success, -1 on a
 * reported error (message on stderr and in pacman_rbf_last_error()), -2 on
 * an unknown failure. */
int pacman_rbf_interpolate_c(int spaceDimension, int execSpace, int rbfFunction,
                             const double *sourcePoints, int nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints, int nTargetPoints,
                             double *targetValues);

/* Replaces pacman_fe_interpolate_c (src/fortran_interface.h:148-153,
 * src/fortran_interface.cpp) for method PACMAN_FE_NEAREST_NEAREST: every
 * target takes the value of its nearest source (ties: lowest source index)
 * and status 0 (OUTSIDE); connectivity arguments are not read (NULL / 0
 * allowed). Caller-allocated outputs of length nTargetPoints. Returns 0, -1
 * on an error (message in pacman_rbf_last_error), -2 on an unknown failure. */
int pacman_fe_interpolate_c(int spaceDimension, int execSpace, int method,
                            const double *sourcePoints, int nSourcePoints,
                            const double *sourceValues, const int *connVal,
                            int connValSize, const int *connOff,
                            int connOffSize, const int *cellTypes,
                            const double *targetPoints, int nTargetPoints,
                            double *targetValues, int *targetStatus);

/* Replaces pacman_mls_interpolate_c (src/fortran_interface.h:175-179,
 * src/fortran_interface.cpp:135-164): moving least squares with the
 * reference's defaults (degree-2 basis, the basis-size nearest sources,
 * Wendland C0 weights over 1.1 times the farthest neighbour's distance).
 * Caller-allocated output of length nTargetPoints. Returns 0, -1 on an error
 * (message in pacman_rbf_last_error), -2 on an unknown failure. */
int pacman_mls_interpolate_c(int spaceDimension, int execSpace,
                             const double *sourcePoints, int nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints, int nTargetPoints,
                             double *targetValues);

#ifdef __cplusplus
}
#endif

#endif /* PACMAN_AMD_FORTRAN_INTERFACE_H */
