/*
 * Extended C ABI of the RBF-PUM transfer path: a persistent plan that splits
 * setup (patch radius, patch centres, membership, weights) from apply
 * (per-patch solve and partition-of-unity blend for F fields), with device-
 * or host-resident arrays and explicit algorithm parameters.
 *
 * The reference has no counterpart: it re-runs the whole setup on every
 * call and hard-codes the parameters (src/rbf_pum/interpolator.hxx:39-41,
 * src/rbf_interface.cpp:62-63). pacman_rbf_interpolate_c is a one-shot plan
 * with those defaults.
 *
 * Conventions: points are row-major [n * spaceDimension] doubles; values are
 * field-major [nFields * nPoints]. `onDevice` = 1 means the pointers are HIP
 * device pointers, 0 means host memory. `stream` is a hipStream_t (NULL for
 * the default stream); calls are synchronous with respect to the host unless
 * stated otherwise. Every function returns 0 on success and a negative code
 * on error; pacman_rbf_last_error() gives the message (thread-local).
 */
#ifndef PACMAN_AMD_RBF_PLAN_H
#define PACMAN_AMD_RBF_PLAN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pacman_rbf_plan_s *pacman_rbf_plan;

/* Builds a plan. supportRadius <= 0 selects the patch radius as support. */
int pacman_rbf_plan_create(pacman_rbf_plan *plan, int spaceDimension,
                           int rbfFunction, int nodesPerCluster,
                           double relativeOverlap, double supportRadius,
                           const double *sourcePoints, long long nSourcePoints,
                           const double *targetPoints, long long nTargetPoints,
                           int onDevice, void *stream);

/* targetValues[f * nTarget + i] = transfer of sourceValues[f * nSource + :].
 * With onDevice = 1 the call is asynchronous on `stream`. */
int pacman_rbf_plan_apply(pacman_rbf_plan plan, const double *sourceValues,
                          int nFields, double *targetValues, int onDevice,
                          void *stream);

/* Transpose of pacman_rbf_plan_apply (no reference counterpart; for the
 * gradient of a loss with respect to the source values):
 * sourceGrad[f * nSource + :] = Op^T targetGrad[f * nTarget + :], Op the
 * linear map source values -> target values of this plan. */
int pacman_rbf_plan_apply_transpose(pacman_rbf_plan plan,
                                    const double *targetGrad, int nFields,
                                    double *sourceGrad, int onDevice,
                                    void *stream);

/* The same apply replayed as a HIP graph, for repeated transfers: the launch
 * sequence is captured on the first call for these (device) pointers and
 * nFields and replayed on later calls (new contents of the same buffers are
 * picked up). Asynchronous on `stream`, ordered with it through events. */
int pacman_rbf_plan_apply_graph(pacman_rbf_plan plan,
                                const double *sourceValues, int nFields,
                                double *targetValues, void *stream);

/* pacman_rbf_plan_apply on device pointers, asynchronous on `stream`, that
 * allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The per-call scratch lives in a
 * static 64 MiB device array per GPU (held until the process ends); a call
 * that needs more (about 8 (F N + F_c pairs + 4 F_s K) bytes) returns -1
 * with the size in the error text. Field counts above 1 need one earlier
 * allocating apply of that count on the plan (it builds the per-count size
 * bins and the target permutation). The plan must belong to the current
 * device. Calls that overlap on one device share the scratch and race. */
int pacman_rbf_plan_apply_noalloc(pacman_rbf_plan plan,
                                  const double *sourceValues, int nFields,
                                  double *targetValues, void *stream);

/* Target-coordinate gradient (no reference counterpart). Doubles of solver
 * state that pacman_rbf_plan_apply_state saves for nFields fields (the
 * coefficients of every patch), or -1 on error. */
long long pacman_rbf_plan_state_size(pacman_rbf_plan plan, int nFields);

/* pacman_rbf_plan_apply (device pointers, asynchronous on `stream`) that also
 * writes the solver state into `state` (device, state_size doubles). */
int pacman_rbf_plan_apply_state(pacman_rbf_plan plan,
                                const double *sourceValues, int nFields,
                                double *targetValues, double *state,
                                void *stream);

/* coordGrad[i * spaceDimension + a] = d/dy_ia of
 * sum_f sum_j targetGrad[f * nTarget + j] u_f(y_j), u the transfer at the
 * plan's targets with the patch structure held fixed (device pointers,
 * asynchronous). `state` comes from pacman_rbf_plan_apply_state with the
 * same nFields and the source values whose transfer is differentiated. */
int pacman_rbf_plan_target_coord_grad(pacman_rbf_plan plan,
                                      const double *state, int nFields,
                                      const double *targetGrad,
                                      double *coordGrad, void *stream);

/* Stored operator for repeated transfers (no reference counterpart): keeps
 * the weight-scaled local operator of every patch in device memory (sum over
 * patches of targets x sources doubles; pacman_rbf_plan_operator_bytes), so
 * that later applies, graph replays included, and transpose applies read it
 * instead of solving every patch again. Synchronous. drop_operator frees it;
 * the target-coordinate gradient (apply_state) keeps using the solve. */
int pacman_rbf_plan_build_operator(pacman_rbf_plan plan, void *stream);
/* The same with the operator rounded to FP32 (half the memory and the bytes
 * each apply reads; products and sums stay FP64). Results differ from the
 * FP64 operator's by the rounding of its entries, about 1e-7 relative. */
int pacman_rbf_plan_build_operator_fp32(pacman_rbf_plan plan, void *stream);
int pacman_rbf_plan_drop_operator(pacman_rbf_plan plan);
long long pacman_rbf_plan_operator_bytes(pacman_rbf_plan plan);

int pacman_rbf_plan_destroy(pacman_rbf_plan plan);

/* Device memory of destroyed plans is cached for reuse by later plans of
 * the same sizes (no reference counterpart). This returns the cached blocks
 * to the HIP runtime (synchronises the device); 0 on success. Setting
 * PACMAN_AMD_NO_MEMORY_CACHE=1 in the environment disables the cache. */
int pacman_release_device_cache(void);
/* Bytes of device memory the cache holds. */
long long pacman_device_cache_bytes(void);

/* Introspection (host outputs). */
double pacman_rbf_plan_radius(pacman_rbf_plan plan);
double pacman_rbf_plan_support(pacman_rbf_plan plan);
long long pacman_rbf_plan_num_patches(pacman_rbf_plan plan);
long long pacman_rbf_plan_num_pairs(pacman_rbf_plan plan);
int pacman_rbf_plan_max_patch_size(pacman_rbf_plan plan);
int pacman_rbf_plan_lattice(pacman_rbf_plan plan, long long *shape3,
                            long long *nCandidates);
/* centers[K * 3] (padded to 3 coordinates), candidate[K] lattice index. */
int pacman_rbf_plan_centers(pacman_rbf_plan plan, double *centers,
                            long long *candidate);
/* which = 0: per patch, source indices within r; 1: per patch, target
 * indices within r - 1e-14; 2: per target, patches within r. Lists are in
 * ascending index order; offsets have (rows + 1) entries. */
long long pacman_rbf_plan_csr_size(pacman_rbf_plan plan, int which);
int pacman_rbf_plan_csr(pacman_rbf_plan plan, int which, long long *offsets,
                        long long *indices);
/* A plan on density-aware patches (no reference counterpart: the reference
 * derives one patch radius from FindRadius, clusters_radius.hpp:22-125,
 * which on strongly clustered clouds gives patches of one to ~1e5 points and
 * leaves sparse regions uncovered). A binary tree over Morton codes of the
 * points halves every cell holding more than leafPoints sources; each leaf
 * with targets is a patch centred on its cell, radius = max((1 +
 * relativeOverlap) * half the cell diagonal, distance to the minPoints-th
 * nearest source), patches holding more than maxPoints sources are halved
 * again; Wendland support = supportScale * radius, PU weights Wendland C2
 * over each patch's own radius. Every target lies inside some patch. The
 * target-coordinate gradient is not available on such plans. */
int pacman_rbf_plan_create_adaptive(
    pacman_rbf_plan *plan, int spaceDimension, int rbfFunction,
    int leafPoints, int minPoints, int maxPoints, double relativeOverlap,
    double supportScale, const double *sourcePoints, long long nSourcePoints,
    const double *targetPoints, long long nTargetPoints, int onDevice,
    void *stream);

/* radii[K]: the radius of every patch (the same value for all patches of a
 * plan built by pacman_rbf_plan_create). */
int pacman_rbf_plan_patch_radii(pacman_rbf_plan plan, double *radii);

/* Patches holding more than 264 source points (no reference counterpart:
 * the reference solves every patch the same way). Patches of up to 264
 * points are solved in registers (16-wave workgroups above 192 points);
 * larger ones by a persistent workspace kernel (blocked L D L^T in device
 * memory, any size). The transpose apply, the stored operator and the
 * target-coordinate gradient need every patch within 192 points and return
 * an error otherwise. */
long long pacman_rbf_plan_num_large_patches(pacman_rbf_plan plan);

/* Arithmetic of the workspace solve (no reference counterpart; the
 * reference is FP64 throughout, serial_solve.hpp:76-84). precision 0: FP64.
 * precision 1 (mixed): FP32 assembly and factorisation, FP64 iterative
 * refinement with a matrix-free FP64 residual (at most maxRefinements
 * steps) and a per-patch FP64 fallback when it does not converge; results
 * agree with FP64 to refinement accuracy (about 1e-13 relative).
 * precision 2 (auto): per field count, whichever of the two the library's
 * size-bin table (pacman_amd/autotune.py) measured faster. Patches of up to
 * 264 points are solved in FP64 either way. Takes effect at the next
 * apply. */
int pacman_rbf_plan_set_precision(pacman_rbf_plan plan, int precision,
                                  int maxRefinements);

/* Size-bin tuning hook (no reference counterpart; used by
 * pacman_amd/autotune.py): gathers `values` (device, F x N, 1 <= F <= 8),
 * fits the polynomials, then times `reps` launches of one class of the
 * per-patch solve on this plan's patches and writes the median in
 * milliseconds to *ms: bin R >= 1 of the register-resident solve (tile rows
 * R = ceil((n + F) / 16)) with `waves` waves per workgroup (8: bins 1..13,
 * 16: bins 11..17), or with R = 0 the workspace-path launches in FP64
 * (wsPrecision 0) or FP32 with FP64 refinement and its fallback pass (1).
 * *ms = -1 when the plan has no patch in that class or the variant does
 * not exist. Synchronises `stream`. */
int pacman_rbf_plan_time_solve(pacman_rbf_plan plan, const double *values,
                               int F, int R, int waves, int wsPrecision,
                               int reps, double *ms, void *stream);

/* Counters of the last apply (synchronises `stream`): counters3[0] patches
 * found numerically singular (pivot <= 1e-12 phi(0): duplicate or
 * near-duplicate sources; their values are not meaningful),
 * counters3[1] refinement steps summed over patches, counters3[2] patches
 * that fell back to FP64. */
int pacman_rbf_plan_apply_counters(pacman_rbf_plan plan, int *counters3,
                                   void *stream);

/* After an apply: active polynomial columns and pivot status per patch. */
int pacman_rbf_plan_patch_info(pacman_rbf_plan plan, int *polyColumns,
                               int *status);

/* Phase times in milliseconds (up to n entries; returns the count written):
 * setup phases 0 grids, 1 patch radius, 2 patch centres, 3 membership,
 * 4 setup total (host clock); apply phases of the last apply (HIP events on
 * its stream) 5 value gather, 6 per-patch solve (polynomial fit included),
 * 7 blend, 8 apply total, 9 polynomial fit of the first field chunk,
 * 10 first-pass workspace launches (patches above 264 sources) of the first
 * field chunk (0 when the plan has none), 11 the register-resident solve
 * launches of the first field chunk, 12 from the end of those to the end of
 * the chunked evaluation of the first field chunk (the patches whose solve
 * stored alpha; it runs beside the workspace launches, so 10 lies inside
 * 12). */
int pacman_rbf_plan_phase_times(pacman_rbf_plan plan, double *ms, int n);

/* Work counters (up to n entries; returns the count written): 0 patches K,
 * 1 sum n_k, 2 sum n_k^2, 3 sum n_k^3, 4 sum n_k m_k (member targets),
 * 5 target/patch pairs within r, 6 largest n_k; over the patches above 264
 * sources: 7 count, 8 sum n_k^3, 9 sum n_k^2, 10 sum n_k m_k of those the
 * workspace kernel evaluates (none: every patch is evaluated apart from its
 * solve); 11 sum n_k m_k of those evaluated in chunks (all); over the
 * patches of up to 264 sources: 12 sum n_k^3, 13 sum n_k^2, 14 sum n_k m_k
 * of those their solve evaluates (none). */
int pacman_rbf_plan_work(pacman_rbf_plan plan, double *work, int n);

const char *pacman_rbf_last_error(void);

/* Nearest-source transfer (FE method NEAREST_NEAREST) on device-resident
 * arrays: targetValues[i] = sourceValues[nearest source of target i] (ties:
 * lowest source index), targetStatus[i] = 0. No reference counterpart with
 * device pointers (the reference's entry is pacman_fe_interpolate_c, which
 * this library also exports). Synchronous; returns 0 or -1. */
int pacman_fe_nearest_device(int spaceDimension, const double *sourcePoints,
                             long long nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints,
                             long long nTargetPoints, double *targetValues,
                             int *targetStatus, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PACMAN_AMD_RBF_PLAN_H */
