/*
 * Persistent MLS transfer plan: the nearest sources and the least-squares
 * weights of every target are computed once (setup), then any number of
 * fields is transferred by a gather (apply). The reference has no
 * counterpart: PACMAN::MLS_interpolate (src/mls_interface.cpp:77-106)
 * rebuilds everything on every call; pacman_mls_interpolate_c is a one-shot
 * plan.
 *
 * Conventions as in pacman_rbf_plan.h: points row-major
 * [n * spaceDimension], values field-major [nFields * nPoints], onDevice = 1
 * for HIP device pointers, 0 for host memory; 0 on success, a negative code
 * on error with the message in pacman_rbf_last_error().
 */
#ifndef PACMAN_AMD_MLS_PLAN_H
#define PACMAN_AMD_MLS_PLAN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pacman_mls_plan_s *pacman_mls_plan;

/* Needs at least k = (dim + 1)(dim + 2) / 2 sources (3, 6, 10). */
int pacman_mls_plan_create(pacman_mls_plan *plan, int spaceDimension,
                           const double *sourcePoints, long long nSourcePoints,
                           const double *targetPoints, long long nTargetPoints,
                           int onDevice, void *stream);

/* targetValues[f * nTarget + i] = sum_j w_ij sourceValues[f * nSource +
 * n_ij]. With onDevice = 1 the call is asynchronous on `stream`. */
int pacman_mls_plan_apply(pacman_mls_plan plan, const double *sourceValues,
                          int nFields, double *targetValues, int onDevice,
                          void *stream);

/* Neighbours per target (k, or -1 for a null plan). */
int pacman_mls_plan_k(pacman_mls_plan plan);

/* Host copies: idx[i * k + j] source index of the j-th nearest source of
 * target i, coef[i * k + j] its weight. */
int pacman_mls_plan_neighbours(pacman_mls_plan plan, int *idx, double *coef);

int pacman_mls_plan_destroy(pacman_mls_plan plan);

#ifdef __cplusplus
}
#endif

#endif /* PACMAN_AMD_MLS_PLAN_H */
