// Internal host-side interfaces between the plan orchestration and the
// kernel translation units. Nothing here is exported from the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "rbf_common.hpp"

#define PACMAN_HIP_CHECK(expr)                                                 \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess)                                                      \
      throw std::runtime_error(std::string("HIP error: ") +                    \
                               hipGetErrorString(_e) + " at " + __FILE__ +     \
                               ":" + std::to_string(__LINE__) + " (" #expr ")"); \
  } while (0)

namespace pacman_amd {

// Named host-side range for tracers (rocprofv3 --marker-trace shows the
// plan's setup, apply and gradient calls and the setup phases); a no-op
// when no tracer is attached.
struct TraceRange {
  explicit TraceRange(const char *name);
  ~TraceRange();
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

// Pageable host array -> device memory through pinned staging buffers
// (host_copy.hip). With afterStream the copy follows the work queued on s;
// without, it is ordered after nothing. Returns once the data is on the
// device.
void copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t s,
                    bool afterStream = true);
// Device memory -> pageable host array after the work queued on s; returns
// once the data is in the host array.
void copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s);

// Largest patch (sources) of the 8-wave register-resident kernels: the
// transpose chain, the stored operator and the target-coordinate gradient
// take patches up to this size. The forward solve takes patches up to
// kTileMaxN in its 16-wave bins; larger ones go through the workspace solve
// (rbf_solve_ws.hip).
constexpr int kMaxPatch = 192;
// phi(0) of Wendland C0 .. C8 (rbf_functions.hpp:22-150, not normalised).
constexpr double kPhi0[5] = {1.0, 1.0, 3.0, 1.0, 15.0};

// Device memory through the library's cache (dev_alloc.hip): blocks freed
// by one plan are reused by the next allocation of the same rounded size.
void *dev_malloc(size_t bytes);
void dev_free(void *p, size_t bytes);
// Returns every cached block to the runtime (synchronises the device).
void dev_release_cache();
size_t dev_cached_bytes();

// Owning device buffer (dev_malloc / dev_free).
template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count)
      p = static_cast<T *>(dev_malloc(count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p)
      dev_free(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  T *get() const { return p; }
};

// Cell-sorted point set owned on the device.
struct DeviceGrid {
  GridView view{};
  DevBuf<int> cell_start, perm;
  DevBuf<double> x, y, z;
  int64_t ncell = 0;
};

// ---- geometry (rbf_geometry.hip) ------------------------------------------
void bbox_device(const double *pts, int64_t n, int dim, double *lo,
                 double *hi, hipStream_t s);
void build_grid(const double *pts, int64_t n, int dim, const double *lo,
                const double *hi, double pointsPerCell, DeviceGrid &g,
                hipStream_t s);
// Exclusive scan of n int32 counts into n+1 int64 offsets; returns the total.
int64_t exclusive_scan_i32_to_i64(const int *in, int64_t *out, int64_t n,
                                  hipStream_t s);
// The same without reading the total back (no host synchronisation).
void exclusive_scan_i32_to_i64_async(const int *in, int64_t *out, int64_t n,
                                     hipStream_t s);
// Exclusive scan of n int32 flags into n+1 int32 offsets; returns the total.
int exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s);

struct ProbeResult {
  double maxd2[7];
  int nearest[7];  // cell-order position of each projected probe
};
void probe_radius(const GridView &src, const double *probes, int nprobe,
                  int k, const double *lo, const double *hi, ProbeResult &res,
                  hipStream_t s);

struct Lattice {
  int dim;
  int64_t shape[3];
  double lower[3];
  double dist[3];
  int64_t nb;
  double r, r2, thr;
  std::vector<int64_t> posOffsets;  // strictly positive neighbour offsets
};
// Tag, snap, prune, re-tag and compact; returns the centres (SoA, device)
// and their candidate index.
int64_t make_centers(const Lattice &L, const GridView &src,
                     const GridView &tgt, DevBuf<double> &cx,
                     DevBuf<double> &cy, DevBuf<double> &cz,
                     DevBuf<int64_t> &cand, hipStream_t s);

// Patch membership CSR for sources (dist <= r).
void patch_sources(const GridView &src, const double *cx, const double *cy,
                   const double *cz, int64_t K, double r,
                   DevBuf<int64_t> &off, DevBuf<int> &idx, hipStream_t s);
// Patch/target pairs (dist <= r) with the member flag (dist <= r - eps),
// and the per-target transpose (slots in ascending patch order).
struct PairSet {
  DevBuf<int64_t> off;      // K + 1: pairs of patch k
  DevBuf<int> tgt;          // P: target cell-order position | member << 31
  DevBuf<int> slot;         // P: slot of the pair in its target's list
  DevBuf<int64_t> toff;     // M + 1: slots of target t
  DevBuf<int> slotPatch;    // P: patch of each slot
  DevBuf<unsigned char> slotMember;  // P
  int64_t P = 0;
};
void patch_targets(const GridView &tgt, const double *cx, const double *cy,
                   const double *cz, int64_t K, double r, PairSet &ps,
                   hipStream_t s);
// The per-target slots of a pair set whose off, tgt and P are filled
// (pairPatch[q] = patch of pair q).
void build_target_slots(PairSet &ps, const int *pairPatch, int64_t M,
                        hipStream_t s);

// ---- adaptive patches (rbf_adaptive.hip) ----------------------------------
struct AdaptiveParams {
  int leaf = 40;             // sources per tree leaf before it is halved
  int nmin = 20;             // sources every patch holds at least
  int nmax = 600;            // patches above this are halved again
  double overlap = 0.15;     // patch radius = (1 + overlap) half-diagonal
  double supportScale = 1.0; // Wendland support = supportScale * radius
};
struct AdaptivePatches {
  DeviceGrid src, tgt;       // Morton-ordered SoA point sets (no cells)
  DevBuf<double> cx, cy, cz, r, rinv, dinv;
  DevBuf<int64_t> key;       // cell key (prefix << (bits - depth))
  DevBuf<int> depth;
  DevBuf<int64_t> srcOff;
  DevBuf<int> srcIdx;        // positions in src
  PairSet pairs;             // positions in tgt
  int64_t K = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};
void build_adaptive(const AdaptiveParams &ap, int dim, const double *dSrc,
                    int64_t N, const double *dTgt, int64_t M,
                    AdaptivePatches &out, hipStream_t s);

// ---- solve (rbf_solve.hip) ------------------------------------------------
struct SolveArgs {
  const int *patches;       // patch ids of this bin
  int count;                // patches in this bin
  const int64_t *srcOff;    // K + 1
  const int *srcIdx;        // cell-order positions (source grid)
  const double *sx, *sy, *sz;
  const double *vals;       // F x N, cell order
  int64_t N;
  int F;
  const int64_t *pairOff;   // K + 1
  const int *pairTgt;       // target position | member << 31
  const int *pairSlot;
  const double *tx, *ty, *tz;  // target coordinates, cell order
  double *contrib;          // P x F
  const double *cx, *cy, *cz;  // patch centres (polynomial basis origin)
  double rinv;              // 1 / patch radius (polynomial basis scale)
  double dinv;              // 1 / support radius
  int dim;
  int *pvOut;               // K (optional) active polynomial columns
  int *status;              // K (optional) nonzero on a zero pivot
  double *alphaOut;         // (sum n) x F (optional): alpha per patch source
  // transpose (adjoint) apply only
  const double *wslot;      // P: Shepard weight of each member slot, else 0
  const double *gSorted;    // F x M upstream values, target cell order
  int64_t M;
  double *gammaOut;         // (sum n) x F: A^-1 E^T (w g) per patch source
  // Stored-operator build (transpose chain with unit upstream values): when
  // >= 0, field f's upstream value is the Shepard weight at pair opRow0 + f
  // of each patch and 0 at its other pairs (gSorted is not read).
  int opRow0;
  // Per-patch 1 / support radius and 1 / patch radius (adaptive patches);
  // null: dinv and rinv for all.
  const double *patchDinv;
  const double *patchRinv;
  // Pivots d <= 1e-12 phi(0) (or not finite) mark a numerically singular
  // patch; each one found adds 1 to *singular (when set).
  double phi0;
  int *singular;
  // Alpha records of the patches evaluated apart from their solve (every
  // patch of the plans built here): patch k's at alphaHeavy + 8 srcOff[k]
  // + alpha_stride(F) e + f (8 doubles reserved per member slot); the solve
  // kernels store them and skip the evaluation, which launch_wave_eval
  // spreads over many waves. Null: the workspace kernel evaluates itself.
  double *alphaHeavy;
  // x y z 0 of the source in each member slot (4 per slot, slot order);
  // read by the one-wave-per-chunk evaluation
  const double *slotXyz;
};
constexpr int kHeavyPairs = 4096;
// Doubles per source in the alpha records of the patches evaluated apart
// (zero past F): 1 for one field, 4 for up to four, 8 above.
__host__ __device__ inline int alpha_stride(int F) {
  return F == 1 ? 1 : (F <= 4 ? 4 : 8);
}

// ---- workspace solve (rbf_solve_ws.hip), patches of any size -------------
constexpr int kWsBlock = 32;   // block edge of the workspace factorisation
constexpr int kWsSlotsPerCu = 1;  // resident workgroups of the kernel per CU
// counters[]: work queue of the first pass, fallback list length, singular
// patches, refinement iterations, work queue of the FP64 fallback pass,
// fallbacks of the whole apply
constexpr int kWsQueue = 0, kWsFallbacks = 1, kWsSingular = 2,
              kWsIrIters = 3, kWsQueue2 = 4, kWsFallbackTotal = 5,
              kWsCounters = 6;
struct WsArgs {
  SolveArgs a;              // patches = queue (largest first), count
  const double *coef;       // K x F x 4 polynomial coefficients
  int nbMax;                // block rows per slot (covers n + F)
  void *work;               // per slot: lower block triangle + Linv blocks
  int64_t slotBlocks;       // elements (of the storage type) per slot
  int64_t linvOffset;       // first Linv block within a slot
  double *vec;              // per slot: coordinates, pivots, 4 x F vectors
  int64_t slotVec;          // doubles per slot
  int *counters;            // kWsCounters ints
  int *fallback;            // patches the FP32 refinement hands to FP64
  const int *countPtr;      // non-null: queue length read on the device
  int queueIdx;             // counter that serves as this launch's queue
  int maxIR;                // refinement iteration limit (FP32)
  double phi0;              // phi(0) of the basis (pivot threshold scale)
};
int64_t ws_slot_blocks(int nbMax);
int64_t ws_linv_offset(int nbMax);
int64_t ws_slot_vec(int nbMax);
// Block rows up to which a slot keeps coordinates and alpha in LDS.
int ws_lds_block_rows();
// all = true: every counter (start of an apply); false: the queues and the
// fallback list only (next chunk of fields).
// keepFallbacks: reset the first-pass queue only (the next queue of the
// same chunk of fields).
void launch_ws_reset(int *counters, bool all, hipStream_t s,
                     bool keepFallbacks = false);
// One persistent launch of `grid` workgroups over the queue: FP64 storage
// (fp32 = false; singular pivots counted) or FP32 storage with FP64
// refinement (non-converging patches appended to the fallback list).
void launch_ws_solve(bool fp32, int kind, const WsArgs &w, int grid,
                     hipStream_t s);
// Evaluation of the patches whose alpha the solve stored (alpha of source
// t at a.alphaHeavy + a.heavyOff[k] + alpha_stride(F) t + f, zero past F
// within the record; polynomial from
// coef; source coordinates from a.slotXyz): chunks[c] = (patch, first pair)
// of up to 64 pairsPerLane (1 to 3) pairs, one wave each.
void launch_wave_eval(int kind, const SolveArgs &a, const double *coef,
                      const int2 *chunks, int nChunks, int pairsPerLane,
                      hipStream_t s);
// slotXyz of the member slots [0, total): 4 doubles each.
void launch_slot_xyz(const int *srcIdx, int64_t total, const double *sx,
                     const double *sy, const double *sz, double *xyz,
                     hipStream_t s);
// The evaluation's chunk list on the device: per (group g, class i) a run of
// count[g][i] chunks from first[g][i] (group 1: patches above tileMaxN
// sources; class i: chunks of i + 1 units of 64 pairs), patches in order.
// Synchronises once (the class sizes).
void plan_eval_chunks(const int64_t *srcOff, const int64_t *pairOff, int64_t K,
                      int tileMaxN, DevBuf<int2> &chunks, int64_t first[2][3],
                      int count[2][3], hipStream_t s);
// Fields handled by one poly/solve launch.
int solve_fields_per_launch();
// Polynomial fit of every patch (one wave per patch): coef[K][F][4], pvOut.
void launch_poly(const SolveArgs &a, int64_t K, double *coef, hipStream_t s);
// Size bin (16-row tile blocks R) of a patch of n sources solved for F
// fields in one launch: room for the patch and the F right-hand-side rows.
int solve_bin_rows(int n, int F);
// Launches the FP64 per-patch solve for the patches of bin R with a
// workgroup of `waves` waves (8: bins 1 .. 13; 16: bins 11 .. kTileMaxR).
void launch_solve_f64(int R, int waves, int kind, const SolveArgs &a,
                      const double *coef, hipStream_t s);
// Whether bin R has a variant with this many waves.
bool solve_tile_variant(int R, int waves);
// Largest size bin of the register-resident solve, and the largest patch it
// takes (room for kFMax right-hand-side rows in that bin).
constexpr int kTileMaxR = 17;
constexpr int kTileMaxN = 16 * kTileMaxR - 8;
// The same for one bin (explicitly instantiated in rbf_solve_tile_*.hip).
template <int R, int NW>
void launch_solve_tile(int kind, const SolveArgs &a, const double *coef,
                       hipStream_t s);
// Transpose of the per-patch chain for the bin with R blocks: right-hand
// side E^T (w g) at the member targets, gammaOut = A^-1 of it.
void launch_solve_f64_adjoint(int R, int kind, const SolveArgs &a,
                              hipStream_t s);
template <int R>
void launch_solve_tile_adjoint(int kind, const SolveArgs &a, hipStream_t s);
// Adjoint of the polynomial part (one wave per patch): contribS[e * F + f] =
// gamma + q(x)^T G^-1 (V^T w g - Q^T gamma) at every patch source entry e.
void launch_adjoint_poly(const SolveArgs &a, int64_t K, const double *gamma,
                         double *contribS, hipStream_t s);
// Storage of the stored operator: FP64, or FP32 applied with FP64
// accumulation. Exactly one of the pointers is set.
struct OpStore {
  double *f64 = nullptr;
  float *f32 = nullptr;
};
// Stored operator: op[opOff[k] + q * n_k + e] = the weight-scaled local
// operator of patch k, row q (its pair q), column e (its source e). From the
// transpose chain's per-source output of launches with opRow0 = q - f.
void launch_op_store(const SolveArgs &a, int64_t K, const double *contribS,
                     const int64_t *opOff, const OpStore &op, hipStream_t s);
// contrib[slot(pair q of patch k) * F + f] = op row q . values of patch k
// (a.vals: F x N, cell order), member pairs only; up to kOpFMax fields per
// pass over the operator (FP64 MFMA, 16 fields per output tile).
constexpr int kOpFMax = 32;
void launch_op_apply(const SolveArgs &a, int64_t K, const int64_t *opOff,
                     const OpStore &op, hipStream_t s);
// Source-major transpose of the patch source lists: srcTOff[N + 1] offsets,
// srcTEnt entry indices (into srcIdx) per source in ascending patch order.
void build_source_transpose(const int *srcIdx, int64_t total, int64_t N,
                            int64_t *srcTOff, int *srcTEnt, hipStream_t s);
// out[f * N + perm[p]] = sum of contribS[e * F + f] over the entries e of
// source position p, in order.
void launch_source_reduce(const int64_t *srcTOff, const int *srcTEnt,
                          int64_t N, int F, const double *contribS,
                          const int *perm, double *out, hipStream_t s);

struct BlendArgs {
  const int64_t *toff;      // M + 1
  const int *slotPatch;
  const unsigned char *slotMember;
  const double *contrib;    // P x F
  const double *tx, *ty, *tz;
  const int *tperm;         // cell order -> original target index
  const int *tinv;          // original target index -> cell order
  const double *wslot;      // Shepard weight per slot (0: not a member)
  const double *cx, *cy, *cz;
  int64_t M;
  int F;
  double rinv;
  const double *patchRinv;  // per-patch 1 / r (adaptive patches) or null
  double *out;              // F x M, original order
};
void launch_blend(const BlendArgs &a, hipStream_t s);
// out[f * M + tperm[t]] = sum of contrib over the member slots of t (the
// weights are inside the stored operator).
void launch_blend_sum(const BlendArgs &a, hipStream_t s);
// contribS[(srcOff[k] + e) * F + f] = sum over the member pairs q of patch k
// of op row q, column e, times a.gSorted[f][target of q] (transpose apply
// through the stored operator; fields <= 8 per launch).
void launch_op_transpose(const SolveArgs &a, int64_t K, const int64_t *opOff,
                         const OpStore &op, double *contribS, hipStream_t s);
// inv[perm[p]] = p for p < n.
void launch_invert_perm(const int *perm, int64_t n, int *inv, hipStream_t s);
// Shepard weight of every member slot (0 for the others).
void launch_slot_weights(const BlendArgs &a, double *wslot, hipStream_t s);

// Nearest-source transfer (FE NEAREST_NEAREST): out[t] = vals[nearest source
// of target t] (ties: lowest source index), status[t] = 0 (OUTSIDE);
// device pointers, points row-major [n x dim].
void fe_nearest_transfer(int dim, const double *dSrc, int64_t N,
                         const double *dVals, const double *dTgt, int64_t M,
                         double *dOut, int *dStatus, hipStream_t s);

void launch_gather_values(const double *vals, const int *perm, int64_t N,
                          int F, double *out, hipStream_t s);

// Gradient with respect to the target coordinates (rbf_grad.hip).
// Per member pair of patch k (one workgroup per patch), with alpha/coef the
// saved solution of the forward apply and gSorted the upstream values
// (F x M, target cell order): gslot[4 s + 0] = sum_f g_f t_kf(y),
// gslot[4 s + 1 + a] = sum_f g_f d t_kf / d y_a at the pair's slot s;
// accumulate adds to gslot instead of overwriting it.
void launch_target_grad_eval(const SolveArgs &a, int64_t K, int kind,
                             const double *alpha, const double *coef,
                             double *gslot, bool accumulate, hipStream_t s);
// Per target: d/dy of sum_k w_k(y) (slot values), with the Shepard weights'
// derivative; coordGrad[o * dim + a] in original target order.
void launch_target_grad_blend(const BlendArgs &a, const double *gslot,
                              int dim, double *coordGrad, hipStream_t s);

} // namespace pacman_amd
