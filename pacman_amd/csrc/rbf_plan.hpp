// RBF-PUM transfer plan: setup (search, clustering, membership) once, apply
// (per-patch solve + partition-of-unity blend) for any number of fields.
// Host-side orchestration of the kernels in rbf_geometry.hip / rbf_solve.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <vector>

#include "rbf_internal.hpp"

namespace pacman_amd {

struct PlanParams {
  int dim = 3;
  int kind = 1;            // Wendland C0..C8 -> 0..4
  int npc = 50;            // nodes per cluster (interpolator.hxx:39)
  double overlap = 0.15;   // relative overlap (interpolator.hxx:40)
  double support = 0.1;    // Wendland support radius; <= 0 means the patch radius
  // 0: FP64 throughout; 1: mixed - patches solved on the workspace path
  // store FP32 factors and are refined to FP64 accuracy (iterative
  // refinement, FP64 fallback per patch); 2: auto - per field count, the
  // precision of the workspace path that the size-bin table measured
  // faster (rbf_tune_table.hpp).
  int precision = 0;
  int maxIR = 6;           // refinement iteration limit (mixed)
  // Patches from the density-aware builder (rbf_adaptive.hip) instead of the
  // reference's FindRadius / CreateClusters (npc is then unused).
  bool adaptive = false;
  AdaptiveParams ap;
};

class RbfPumPlan {
public:
  RbfPumPlan(const PlanParams &p, const double *dSrc, int64_t N,
             const double *dTgt, int64_t M, hipStream_t s);

  // dVals: F x N (field-major, original point order), dOut: F x M.
  void apply(const double *dVals, int F, double *dOut, hipStream_t s);
  // The same as a HIP graph replay (captured on the first call for these
  // pointers and F; asynchronous on s). For repeated transfers.
  void applyGraph(const double *dVals, int F, double *dOut, hipStream_t s);
  // apply with the per-call scratch (sorted source values, (target, patch)
  // contributions, polynomial coefficients) carved from `scratch` (device,
  // `capacity` doubles, 256-byte aligned) instead of plan buffers: nothing
  // is allocated. Throws when the scratch is too small, when the plan is on
  // another device than the current one, or when the plan lacks the
  // per-field-count bins or (F >= 2) the target permutation and slot weights
  // that a first apply of that field count builds.
  void applyNoAlloc(const double *dVals, int F, double *dOut, hipStream_t s,
                    double *scratch, int64_t capacity);
  // Doubles of scratch applyNoAlloc needs for F fields.
  int64_t noAllocScratch(int F) const;
  int device() const { return device_; }
  // Transpose of apply (the gradient with respect to the source values):
  // dG: F x M (original target order), dGrad: F x N (original source order).
  void applyTranspose(const double *dG, int F, double *dGrad, hipStream_t s);
  // Solver state an apply saves for the target-coordinate gradient: per
  // chunk of fields solved together, the coefficients alpha of every patch
  // source entry, then the polynomial coefficients of every patch.
  int64_t stateSize(int F) const;
  // apply that also writes the state (device buffer of stateSize(F)).
  void applyState(const double *dVals, int F, double *dOut, double *dState,
                  hipStream_t s);
  // Gradient of sum_f <dG_f, u_f> with respect to the target coordinates at
  // the plan's targets, the patch structure held fixed: dCoordGrad is M x dim
  // (original order, row-major); dState from applyState with the same F.
  void targetCoordGrad(const double *dState, int F, const double *dG,
                       double *dCoordGrad, hipStream_t s);

  // Stores the weight-scaled local operator of every patch (sum over patches
  // of pairs x sources entries of HBM, FP64 or with fp32 FP32); later applies
  // use it instead of solving: one read of the operator per apply.
  // dropOperator frees it.
  void buildOperator(hipStream_t s, bool fp32 = false);
  void dropOperator();
  int64_t operatorBytes() const;

  const PlanParams &params() const { return prm_; }
  int64_t numSources() const { return N_; }
  int64_t numTargets() const { return M_; }
  double radius() const { return r_; }
  double support() const { return delta_; }
  int64_t numPatches() const { return K_; }
  int64_t numCandidates() const { return lat_.nb; }
  const int64_t *latticeShape() const { return lat_.shape; }
  int64_t numPairs() const { return pairs_.P; }
  int maxPatchSize() const { return maxN_; }
  // Patches above the register-resident solve's limit (solved on the
  // workspace path).
  int64_t numLargePatches() const { return (int64_t)wsHost_.size(); }
  void setPrecision(int precision, int maxIR);
  // Size-bin tuning: median time in ms of `reps` launches, after the value
  // gather and polynomial fit of dVals (F x N, device), of the
  // register-resident solve of bin R with `waves` waves per workgroup
  // (R >= 1), or of the workspace-path launches (R = 0) with FP64
  // (wsPrecision 0) or FP32 factors and FP64 refinement plus the FP64
  // fallback pass (1). -1 when the plan has no patch in that class or the
  // variant does not exist.
  double timeSolve(const double *dVals, int F, int R, int waves,
                   int wsPrecision, int reps, hipStream_t s);
  // Counters of the last apply (synchronises s): 0 numerically singular
  // patches, 1 refinement iterations (summed over patches), 2 patches that
  // fell back from FP32 refinement to FP64.
  void applyCounters(int *out, hipStream_t s) const;
  bool adaptive() const { return prm_.adaptive; }
  // Per-patch radii (host copy, K entries; adaptive plans: each patch's
  // own, otherwise the one radius).
  void patchRadii(double *r, hipStream_t s) const;

  // Host copies for inspection and parity tests (original point indices,
  // each list sorted ascending).
  void centers(double *xyz, int64_t *candidate, hipStream_t s) const;
  int64_t csrSize(int which) const;  // 0 sources, 1 member targets, 2 weights
  void csr(int which, int64_t *off, int64_t *idx, hipStream_t s) const;
  void patchInfo(int *pv, int *status, hipStream_t s) const;

  // Phase times in ms. Setup phases are host-timed (each ends with a
  // synchronisation): 0 grids, 1 radius, 2 centres, 3 membership, 4 setup
  // total. Apply phases come from HIP events of the last apply on its
  // stream: 5 value gather, 6 per-patch solve, 7 blend, 8 apply total,
  // 9 polynomial fit of the first field chunk, 10 first-pass workspace
  // launches (patches above tileMaxN_) of the first field chunk, 11 the
  // register-resident solve launches of the first field chunk, 12 from the
  // end of those to the end of the chunked evaluation of the first field
  // chunk (the evaluation of the register-resident patches runs beside the
  // workspace launches, so 10 lies inside 12).
  static constexpr int kPhases = 13;
  void phaseTimes(double *ms) const;
  // Work counters: 0 K, 1 sum n, 2 sum n^2, 3 sum n^3, 4 sum n m (member
  // targets), 5 pairs (targets within r of a centre), 6 max n; over the
  // patches above tileMaxN_: 7 count, 8 sum n^3, 9 sum n^2, 10 sum n m of
  // those evaluated by the workspace kernel (none), 11 sum n m of the
  // patches evaluated in chunks (all of them); over the patches of the
  // register-resident solve: 12 sum n^3, 13 sum n^2, 14 sum n m of those it
  // evaluates itself (none).
  static constexpr int kWork = 15;
  void work(double *w) const;
  ~RbfPumPlan();

private:
  PlanParams prm_;
  int64_t N_ = 0, M_ = 0, K_ = 0;
  double lo_[3], hi_[3], tlo_[3], thi_[3];
  double r_ = 0.0, delta_ = 0.0;
  Lattice lat_{};
  DeviceGrid gs_, gt_;
  DevBuf<double> cx_, cy_, cz_;
  DevBuf<int64_t> cand_;
  DevBuf<int64_t> srcOff_;
  DevBuf<int> srcIdx_;
  PairSet pairs_;
  // Register-resident solve: patches of up to tileMaxN_ sources, binned by
  // tile blocks R (waves per workgroup from the size-bin table).
  int tileMaxN_ = kTileMaxN;
  // Patches binned by tile blocks R, one binning per fields-per-launch
  // count F (index F - 1).
  struct Bins {
    bool built = false;
    std::vector<int> R, count;
    std::vector<int64_t> first;  // bin b: patches dev[first[b] ..]
    std::vector<int> host;       // kept alive for the async copy
    DevBuf<int> dev;
  };
  std::vector<Bins> bins_;
  // Workspace path: patches above tileMaxN_ sources, largest first, in two
  // queues (bordered matrices that fit the LDS variant, larger ones); the
  // FP64 fallback pass of mixed precision takes any of them.
  struct WsQueue {
    std::vector<int> host;
    DevBuf<int> dev;
    int nbMax = 0, slots = 0;
  };
  WsQueue wsQ_[3];            // 0: LDS class, 1: larger, 2: fallback shape
  std::vector<int> wsHost_;   // all workspace patches
  DevBuf<int> wsFallback_, counters_;
  DevBuf<double> wsWork_, wsVec_;  // wsWork_ also holds FP32 blocks
  void plan_workspace(hipStream_t s);
  void setup_reference(const double *dSrc, const double *dTgt, hipStream_t s,
                       std::chrono::steady_clock::time_point t0);
  void setup_adaptive(const double *dSrc, const double *dTgt, hipStream_t s);
  DevBuf<double> patchR_, patchRinv_, patchDinv_;  // adaptive plans only
  // Every patch is evaluated apart from its solve, in chunks of pairs (one
  // wave each) by launches of their own.
  void plan_eval(hipStream_t s);
  DevBuf<double> alphaHeavy_;
  // chunks of the wave evaluation with 1, 2 and 3 pairs per lane, [0] of
  // the patches of the register-resident solve, [1] of the workspace ones:
  // nEvalChunks_[g][i] chunks from evalChunks_ + evalFirst_[g][i]
  DevBuf<int2> evalChunks_;
  int64_t evalFirst_[2][3] = {};
  int nEvalChunks_[2][3] = {};
  hipEvent_t efork_ = nullptr, ejoin_ = nullptr;
  DevBuf<double> slotXyz_;  // 4 per member slot, for the wave evaluation
  void require_register_patches(const char *what) const;
  void launch_workspace(const SolveArgs &a, const double *coef, hipStream_t s,
                        bool timed, bool fp32);
  bool workspace_fp32(int F) const;
  std::vector<int> patchN_;  // sources per patch (host)
  const Bins &bins_for(int F, hipStream_t s);
  // launch(R, args, stream) once per size bin, args.patches/count set
  template <class Launch>
  void launch_bins(const Bins &B, SolveArgs a, hipStream_t s, Launch &&launch);
  void prepare(int F, hipStream_t s);
  // per-call buffers of an apply: F x N sorted source values, pairs x Fc
  // contributions, K x Fs x 4 polynomial coefficients
  struct ApplyBufs {
    double *vals, *contrib, *coef;
  };
  void enqueue(const double *dVals, int F, double *dOut, hipStream_t s,
               bool events, double *dState = nullptr,
               const ApplyBufs *bufs = nullptr);
  int device_ = 0;
  void releaseGraph();
  // Grows a buffer that enqueue() reads or writes; a reallocation bumps the
  // generation that a captured graph was recorded against, so the next graph
  // apply captures again instead of replaying stale pointers. Returns true
  // when it reallocated.
  template <typename T> bool grow(DevBuf<T> &b, size_t n) {
    if (b.n >= n)
      return false;
    b.alloc(n);
    ++bufGen_;
    return true;
  }
  void ensure_wslot(hipStream_t s);
  uint64_t bufGen_ = 0, gGen_ = 0;
  SolveArgs solveArgs(int F, const double *vals) const;
  BlendArgs blendArgs(int F) const;
  bool adjReady_ = false;
  DevBuf<double> wslot_, gSorted_, gamma_, contribS_, gslot_;
  DevBuf<int64_t> srcTOff_;
  DevBuf<int> srcTEnt_;
  hipStream_t gstream_ = nullptr;
  hipEvent_t gev_[2] = {nullptr, nullptr};
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
  const double *gVals_ = nullptr;
  double *gOut_ = nullptr;
  int gF_ = 0;
  int maxN_ = 0;
  int64_t totalSrc_ = 0;  // sum over patches of the source counts
  DevBuf<int> pv_, status_;
  bool opBuilt_ = false;
  DevBuf<int64_t> opOff_;
  DevBuf<double> op_;
  DevBuf<float> op32_;  // set instead of op_ for an FP32 operator
  OpStore opStore() const;
  void ensure_tinv(int F, hipStream_t s);
  DevBuf<int> tinv_;  // original target index -> cell order (operator apply)
  DevBuf<double> valsSorted_, contrib_, coef_;
  double setupMs_[5] = {0, 0, 0, 0, 0};
  hipEvent_t ev_[10] = {};
  bool wsTimed_ = false;  // ev_[5], ev_[6] recorded by the last apply
  // more streams of the size-bin launches, forked from and joined to the
  // apply's stream
  static constexpr int kBinStreams = 3;
  hipStream_t bstream_[kBinStreams - 1] = {};
  hipEvent_t bfork_ = nullptr, bjoin_[kBinStreams - 1] = {};
  bool applied_ = false;
};

} // namespace pacman_amd
