// Host orchestration of the RBF-PUM plan. Follows the constructor sequence
// of src/rbf_pum/interpolator.hpp:58-94: FindRadius, SetRadiusInv on the RBF
// (1 / support) and on the PU weight (1 / r), CreateClusters, Interpolate.
// Compiled with -ffp-contract=off so every host-side formula (probe points,
// lattice spacing) rounds exactly as the CPU restatement does.

#include "rbf_plan.hpp"
#include "rbf_tune_table.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>

namespace pacman_amd {

TraceRange::TraceRange(const char *name) { roctxRangePushA(name); }
TraceRange::~TraceRange() { roctxRangePop(); }

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now() - t0)
      .count();
}

} // namespace

RbfPumPlan::RbfPumPlan(const PlanParams &p, const double *dSrc, int64_t N,
                       const double *dTgt, int64_t M, hipStream_t s)
    : prm_(p), N_(N), M_(M) {
  TraceRange tr("pacman::plan setup");
  const int D = p.dim;
  if (D < 1 || D > 3)
    throw std::runtime_error("The dimension of the points can only be: 1, 2 or 3.");
  if (p.kind < 0 || p.kind > 4)
    throw std::invalid_argument("Unsupported RBF function selector");
  if (N <= 0)
    throw std::invalid_argument("RBF-PUM needs at least one source point");
  if (M < 0)
    throw std::invalid_argument("negative number of target points");
  if (p.npc <= 0 && !p.adaptive)
    throw std::invalid_argument("nodesPerCluster must be positive");
  if (!(p.overlap > 0.0 && p.overlap < 1.0))
    throw std::invalid_argument("relativeOverlap must lie in (0, 1)");
  if (N > INT_MAX / 2 || M > INT_MAX / 2)
    throw std::invalid_argument("point counts above 2^30 are not supported");

  const auto tSetup = std::chrono::steady_clock::now();
  if (p.adaptive)
    setup_adaptive(dSrc, dTgt, s);
  else
    setup_reference(dSrc, dTgt, s, tSetup);
  auto tPhase = std::chrono::steady_clock::now();

  // Size bins by register tile: per-patch source counts on the host.
  TraceRange trb("pacman::plan bookkeeping");
  std::vector<int64_t> off(K_ + 1);
  PACMAN_HIP_CHECK(hipMemcpyAsync(off.data(), srcOff_.get(), 8 * (K_ + 1),
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  maxN_ = 0;
  for (int64_t kk = 0; kk < K_; ++kk)
    maxN_ = std::max(maxN_, (int)(off[kk + 1] - off[kk]));
  patchN_.resize(K_);
  for (int64_t kk = 0; kk < K_; ++kk)
    patchN_[kk] = (int)(off[kk + 1] - off[kk]);
  totalSrc_ = off[K_];
  PACMAN_HIP_CHECK(hipGetDevice(&device_));
  counters_.alloc(kWsCounters);
  launch_ws_reset(counters_.get(), true, s);
  plan_workspace(s);
  plan_eval(s);
  bins_.resize(solve_fields_per_launch());
  bins_for(1, s);
  pv_.alloc(std::max<int64_t>(K_, 1));
  status_.alloc(std::max<int64_t>(K_, 1));
  for (auto &e : ev_)
    PACMAN_HIP_CHECK(hipEventCreate(&e));
  PACMAN_HIP_CHECK(hipEventCreateWithFlags(&bfork_, hipEventDisableTiming));
  PACMAN_HIP_CHECK(hipEventCreateWithFlags(&efork_, hipEventDisableTiming));
  PACMAN_HIP_CHECK(hipEventCreateWithFlags(&ejoin_, hipEventDisableTiming));
  for (int k = 0; k + 1 < kBinStreams; ++k) {
    PACMAN_HIP_CHECK(
        hipStreamCreateWithFlags(&bstream_[k], hipStreamNonBlocking));
    PACMAN_HIP_CHECK(
        hipEventCreateWithFlags(&bjoin_[k], hipEventDisableTiming));
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  setupMs_[3] += ms_since(tPhase);
  setupMs_[4] = ms_since(tSetup);
}

// Density-aware patches (rbf_adaptive.hip): phases 0 (point sets), 2
// (patches) and 3 (membership) share one host-timed interval, reported as
// phase 2; radius and lattice phases are zero.
void RbfPumPlan::setup_adaptive(const double *dSrc, const double *dTgt,
                                hipStream_t s) {
  TraceRange tr("pacman::setup density-aware patches");
  const auto t0 = std::chrono::steady_clock::now();
  AdaptivePatches ad;
  build_adaptive(prm_.ap, prm_.dim, dSrc, N_, dTgt, M_, ad, s);
  for (int a = 0; a < 3; ++a) {
    lo_[a] = ad.lo[a];
    hi_[a] = ad.hi[a];
    tlo_[a] = ad.lo[a];
    thi_[a] = ad.hi[a];
  }
  gs_ = std::move(ad.src);
  gt_ = std::move(ad.tgt);
  cx_ = std::move(ad.cx);
  cy_ = std::move(ad.cy);
  cz_ = std::move(ad.cz);
  cand_ = std::move(ad.key);
  patchR_ = std::move(ad.r);
  patchRinv_ = std::move(ad.rinv);
  patchDinv_ = std::move(ad.dinv);
  srcOff_ = std::move(ad.srcOff);
  srcIdx_ = std::move(ad.srcIdx);
  pairs_ = std::move(ad.pairs);
  K_ = ad.K;
  lat_ = Lattice{};
  lat_.dim = prm_.dim;
  lat_.nb = K_;
  for (int a = 0; a < 3; ++a)
    lat_.shape[a] = 0;
  // reported radius and support: the largest patch's
  std::vector<double> r(std::max<int64_t>(K_, 1), 0.0);
  if (K_ > 0)
    PACMAN_HIP_CHECK(hipMemcpyAsync(r.data(), patchR_.get(), 8 * K_,
                                    hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  r_ = *std::max_element(r.begin(), r.end());
  delta_ = prm_.ap.supportScale * r_;
  setupMs_[2] = ms_since(t0);
}

// The reference's patch construction: FindRadius, CreateClusters and the
// membership queries on the uniform cell grids.
void RbfPumPlan::setup_reference(const double *dSrc, const double *dTgt,
                                 hipStream_t s,
                                 std::chrono::steady_clock::time_point t0) {
  TraceRange tr("pacman::setup reference patches");
  const PlanParams &p = prm_;
  const int D = p.dim;
  const int64_t N = N_, M = M_;
  auto tPhase = t0;
  // Bounding boxes (the source box is the BVH bounds of the reference).
  bbox_device(dSrc, N, D, lo_, hi_, s);
  bbox_device(dTgt, M, D, tlo_, thi_, s);
  build_grid(dSrc, N, D, lo_, hi_, 4.0, gs_, s);
  build_grid(dTgt, M, D, tlo_, thi_, 4.0, gt_, s);
  setupMs_[0] = ms_since(tPhase);
  tPhase = std::chrono::steady_clock::now();

  // FindRadius (clusters_radius.hpp:22-125)
  double center[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a)
    center[a] = (lo_[a] + hi_[a]) / 2.;
  const int nprobe = 2 * D + 1;
  double probes[7][3];
  for (int a = 0; a < 3; ++a)
    probes[0][a] = center[a];
  for (int ax = 0; ax < D; ++ax) {
    double smp[3] = {center[0], center[1], center[2]};
    const double len = std::fabs(hi_[ax] - lo_[ax]);
    smp[ax] -= len * 0.25;
    for (int a = 0; a < 3; ++a)
      probes[ax + 1][a] = smp[a];
    smp[ax] += len * 0.5;
    for (int a = 0; a < 3; ++a)
      probes[D + ax + 1][a] = smp[a];
  }
  ProbeResult pr{};
  const int k = (int)std::min<int64_t>(p.npc, N);
  probe_radius(gs_.view, &probes[0][0], nprobe, k, lo_, hi_, pr, s);
  std::vector<double> m(pr.maxd2, pr.maxd2 + nprobe);
  std::sort(m.begin(), m.end());
  r_ = std::sqrt(m[D]);
  if (!(r_ > 0.0))
    throw std::runtime_error(
        "patch radius is zero (all probe neighbourhoods are single points)");
  delta_ = p.support > 0.0 ? p.support : r_;
  setupMs_[1] = ms_since(tPhase);
  tPhase = std::chrono::steady_clock::now();

  // CreateClusters lattice (clustering.hpp:137-204)
  lat_.dim = D;
  const double spacing = std::sqrt(4.0 / D) * r_ * (1.0 - p.overlap);
  double minDistance = DBL_MAX;
  int64_t nb = 1;
  for (int a = 0; a < 3; ++a) {
    lat_.shape[a] = 1;
    lat_.lower[a] = 0.0;
    lat_.dist[a] = 0.0;
  }
  for (int a = 0; a < D; ++a) {
    double edge = hi_[a] - lo_[a];
    edge = (edge < 0) ? (-edge) : (edge);
    const int shp = (int)std::ceil(std::max(1.0, edge / spacing));
    lat_.dist[a] = edge / shp;
    if (lat_.dist[a] < minDistance)
      minDistance = lat_.dist[a];
    lat_.shape[a] = shp + 1;
    lat_.lower[a] = lo_[a];
    nb *= lat_.shape[a];
  }
  lat_.nb = nb;
  lat_.r = r_;
  lat_.r2 = r_ * r_;
  double thr = 0.4 * minDistance;
  thr *= thr;
  lat_.thr = thr;
  // ZCurveNeighborOffsets<D> (clustering.hpp:34-116): the 3^D - 1 row-major
  // unit offsets. Only the positive half is needed by the parallel rule.
  const int total = D == 1 ? 3 : D == 2 ? 9 : 27;
  for (int t = 0; t < total; ++t) {
    int o[3] = {0, 0, 0}, tt = t;
    bool zero = true;
    for (int d = 0; d < D; ++d) {
      o[d] = tt % 3 - 1;
      tt /= 3;
      zero = zero && o[d] == 0;
    }
    if (zero)
      continue;
    int64_t idx = 0, stride = 1;
    for (int d = D - 1; d >= 0; --d) {
      idx += o[d] * stride;
      stride *= lat_.shape[d];
    }
    if (idx > 0)
      lat_.posOffsets.push_back(idx);
  }
  K_ = make_centers(lat_, gs_.view, gt_.view, cx_, cy_, cz_, cand_, s);
  setupMs_[2] = ms_since(tPhase);
  tPhase = std::chrono::steady_clock::now();

  // Interpolate(): membership (solver.hpp:44-67) and weights (:128-169)
  patch_sources(gs_.view, cx_.get(), cy_.get(), cz_.get(), K_, r_, srcOff_,
                srcIdx_, s);
  patch_targets(gt_.view, cx_.get(), cy_.get(), cz_.get(), K_, r_, pairs_, s);
  setupMs_[3] = ms_since(tPhase);
}

// Patches grouped by tile blocks R = solve_bin_rows(n, F), one launch each.
const RbfPumPlan::Bins &RbfPumPlan::bins_for(int F, hipStream_t s) {
  Bins &B = bins_[F - 1];
  if (B.built)
    return B;
  TraceRange tr("pacman::size bins");
  // every bin's patch list in one buffer (ascending patch within a bin, bins
  // by R), one copy: count per bin, then place
  const int nR = solve_bin_rows(tileMaxN_, F) + 1;
  std::vector<int64_t> at(nR, 0);
  for (int64_t kk = 0; kk < K_; ++kk)
    if (patchN_[kk] <= tileMaxN_)  // larger ones: workspace path
      ++at[solve_bin_rows(patchN_[kk], F)];
  int64_t total = 0;
  for (int R = 0; R < nR; ++R) {
    const int64_t c = at[R];
    at[R] = total;
    if (c == 0)
      continue;
    B.first.push_back(total);
    B.count.push_back((int)c);
    B.R.push_back(R);
    total += c;
  }
  B.host.resize(total);
  for (int64_t kk = 0; kk < K_; ++kk)
    if (patchN_[kk] <= tileMaxN_)
      B.host[at[solve_bin_rows(patchN_[kk], F)]++] = (int)kk;
  if (!B.host.empty()) {
    B.dev.alloc(B.host.size());
    PACMAN_HIP_CHECK(hipMemcpyAsync(B.dev.get(), B.host.data(),
                                    4 * B.host.size(), hipMemcpyHostToDevice,
                                    s));
  }
  B.built = true;
  return B;
}

// The size bins go round-robin over s and kBinStreams - 1 more streams
// (their patches and slots are disjoint), so one bin's last workgroups run
// beside the next bins' first; s waits for all of them at the end.
template <class Launch>
void RbfPumPlan::launch_bins(const Bins &B, SolveArgs a, hipStream_t s,
                             Launch &&launch) {
  const int ns = (int)std::min<size_t>(kBinStreams, B.R.size());
  if (ns > 1) {
    PACMAN_HIP_CHECK(hipEventRecord(bfork_, s));
    for (int k = 1; k < ns; ++k)
      PACMAN_HIP_CHECK(hipStreamWaitEvent(bstream_[k - 1], bfork_, 0));
  }
  for (size_t b = 0; b < B.R.size(); ++b) {
    a.patches = B.dev.get() + B.first[b];
    a.count = B.count[b];
    const int k = (int)(b % (size_t)ns);
    launch(B.R[b], a, k == 0 ? s : bstream_[k - 1]);
  }
  for (int k = 1; k < ns; ++k) {
    PACMAN_HIP_CHECK(hipEventRecord(bjoin_[k - 1], bstream_[k - 1]));
    PACMAN_HIP_CHECK(hipStreamWaitEvent(s, bjoin_[k - 1], 0));
  }
}

RbfPumPlan::~RbfPumPlan() {
  releaseGraph();
  if (gstream_)
    (void)hipStreamDestroy(gstream_);
  for (auto &e : gev_)
    if (e)
      (void)hipEventDestroy(e);
  for (auto &e : ev_)
    if (e)
      (void)hipEventDestroy(e);
  if (bfork_)
    (void)hipEventDestroy(bfork_);
  if (efork_)
    (void)hipEventDestroy(efork_);
  if (ejoin_)
    (void)hipEventDestroy(ejoin_);
  for (int k = 0; k + 1 < kBinStreams; ++k) {
    if (bjoin_[k])
      (void)hipEventDestroy(bjoin_[k]);
    if (bstream_[k])
      (void)hipStreamDestroy(bstream_[k]);
  }
}

void RbfPumPlan::phaseTimes(double *ms) const {
  for (int i = 0; i < 5; ++i)
    ms[i] = setupMs_[i];
  for (int i = 5; i < kPhases; ++i)
    ms[i] = 0.0;
  if (!applied_)
    return;
  PACMAN_HIP_CHECK(hipEventSynchronize(ev_[3]));
  float t = 0.f;
  PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[0], ev_[1]));
  ms[5] = t;
  PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[1], ev_[2]));
  ms[6] = t;
  PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[2], ev_[3]));
  ms[7] = t;
  PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[0], ev_[3]));
  ms[8] = t;
  PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[1], ev_[4]));
  ms[9] = t;
  if (wsTimed_) {
    PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[5], ev_[6]));
    ms[10] = t;
  }
  if (!opBuilt_) {
    PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[4], ev_[7]));
    ms[11] = t;
    PACMAN_HIP_CHECK(hipEventElapsedTime(&t, ev_[8], ev_[9]));
    ms[12] = t;
  }
}

void RbfPumPlan::work(double *w) const {
  std::vector<int64_t> off(K_ + 1), po(K_ + 1);
  PACMAN_HIP_CHECK(hipMemcpy(off.data(), srcOff_.get(), 8 * (K_ + 1),
                             hipMemcpyDeviceToHost));
  PACMAN_HIP_CHECK(hipMemcpy(po.data(), pairs_.off.get(), 8 * (K_ + 1),
                             hipMemcpyDeviceToHost));
  std::vector<int> t(pairs_.P);
  if (pairs_.P)
    PACMAN_HIP_CHECK(hipMemcpy(t.data(), pairs_.tgt.get(), 4 * pairs_.P,
                               hipMemcpyDeviceToHost));
  double sn = 0, sn2 = 0, sn3 = 0, snm = 0;
  double wk = 0, wn3 = 0, wn2 = 0, wnm = 0, hnm = 0;
  double tn3 = 0, tn2 = 0, tnm = 0;
  for (int64_t k = 0; k < K_; ++k) {
    const double n = (double)(off[k + 1] - off[k]);
    int64_t m = 0;
    for (int64_t e = po[k]; e < po[k + 1]; ++e)
      m += ((unsigned)t[e] >> 31) ? 1 : 0;
    sn += n;
    sn2 += n * n;
    sn3 += n * n * n;
    snm += n * (double)m;
    const bool heavy = true;  // every patch is evaluated in chunks
    if (heavy)
      hnm += n * (double)m;
    if (n > tileMaxN_) {
      wk += 1;
      wn3 += n * n * n;
      wn2 += n * n;
      if (!heavy)
        wnm += n * (double)m;
    } else {
      tn3 += n * n * n;
      tn2 += n * n;
      if (!heavy)
        tnm += n * (double)m;
    }
  }
  w[0] = (double)K_;
  w[1] = sn;
  w[2] = sn2;
  w[3] = sn3;
  w[4] = snm;
  w[5] = (double)pairs_.P;
  w[6] = (double)maxN_;
  w[7] = wk;
  w[8] = wn3;
  w[9] = wn2;
  w[10] = wnm;
  w[11] = hnm;
  w[12] = tn3;
  w[13] = tn2;
  w[14] = tnm;
}

// Buffers and size bins an apply of F fields needs (allocations happen here,
// never inside enqueue, so enqueue can be captured into a graph).
void RbfPumPlan::prepare(int F, hipStream_t s) {
  if (F < 1)
    throw std::invalid_argument("apply needs at least one field");
  TraceRange tr("pacman::apply buffers and size bins");
  const int64_t P = std::max<int64_t>(pairs_.P, 1);
  grow(valsSorted_, (size_t)F * N_);
  const int perSolve = solve_fields_per_launch();
  // fields per pass: up to kOpFMax over a stored operator, else per solve
  const int Fc = std::min(F, opBuilt_ ? kOpFMax : perSolve);
  grow(contrib_, (size_t)P * Fc);
  const int Fs = std::min(F, perSolve);
  grow(coef_, (size_t)std::max<int64_t>(K_, 1) * Fs * 4);
  for (int f0 = 0; f0 < F; f0 += perSolve)
    bins_for(std::min(perSolve, F - f0), s);
  ensure_tinv(F, s);
}

// Original target index -> cell order and the slot weights, for the blends
// that write several fields in original order (built on first use by such an
// apply).
void RbfPumPlan::ensure_tinv(int F, hipStream_t s) {
  if (F < 2)
    return;
  if (grow(tinv_, std::max<int64_t>(M_, 1)))
    launch_invert_perm(gt_.view.perm, M_, tinv_.get(), s);
  ensure_wslot(s);
}

// Shepard weight of every member slot (built once; read by the multi-field
// blends, the transpose chain and the operator build).
void RbfPumPlan::ensure_wslot(hipStream_t s) {
  if (grow(wslot_, std::max<int64_t>(pairs_.P, 1)))
    launch_slot_weights(blendArgs(1), wslot_.get(), s);
}

void RbfPumPlan::apply(const double *dVals, int F, double *dOut,
                       hipStream_t s) {
  TraceRange tr("pacman::apply");
  if (M_ == 0)
    return;
  prepare(F, s);
  enqueue(dVals, F, dOut, s, true);
  applied_ = true;
}

namespace {
int64_t up256(int64_t doubles) { return (doubles + 31) / 32 * 32; }
} // namespace

int64_t RbfPumPlan::noAllocScratch(int F) const {
  if (F < 1)
    throw std::invalid_argument("apply needs at least one field");
  const int perSolve = solve_fields_per_launch();
  const int Fc = std::min(F, opBuilt_ ? kOpFMax : perSolve);
  const int Fs = std::min(F, perSolve);
  return up256((int64_t)F * N_) +
         up256(std::max<int64_t>(pairs_.P, 1) * Fc) +
         up256(std::max<int64_t>(K_, 1) * Fs * 4);
}

// The buffers prepare() would grow come from the caller's scratch; the
// per-field-count state that prepare() builds must already exist.
void RbfPumPlan::applyNoAlloc(const double *dVals, int F, double *dOut,
                              hipStream_t s, double *scratch,
                              int64_t capacity) {
  TraceRange tr("pacman::apply (no allocation)");
  const int64_t need = noAllocScratch(F);
  int dev = -1;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  if (dev != device_)
    throw std::invalid_argument(
        "the plan was built on device " + std::to_string(device_) +
        " but device " + std::to_string(dev) + " is current");
  if (need > capacity)
    throw std::length_error(
        std::to_string(F) + " fields over this plan need " +
        std::to_string((need * 8 + 1048575) / 1048576) +
        " MiB of scratch; the allocation-free apply holds " +
        std::to_string(capacity * 8 / 1048576) + " MiB");
  if (M_ == 0)
    return;
  if (!opBuilt_) {
    const int perSolve = solve_fields_per_launch();
    for (int f0 = 0; f0 < F; f0 += perSolve) {
      const int nf = std::min(perSolve, F - f0);
      if (!bins_[nf - 1].built)
        throw std::logic_error(
            "the plan has no size bins for " + std::to_string(nf) +
            " fields per solve; one allocating apply of " +
            std::to_string(F) + " fields builds them");
    }
  }
  if (F >= 2 && (tinv_.n == 0 || wslot_.n == 0))
    throw std::logic_error(
        "the multi-field blend needs the target permutation and slot "
        "weights; one allocating apply of two or more fields builds them");
  const int64_t nv = up256((int64_t)F * N_);
  const int perSolve = solve_fields_per_launch();
  const int Fc = std::min(F, opBuilt_ ? kOpFMax : perSolve);
  const int64_t nc = up256(std::max<int64_t>(pairs_.P, 1) * Fc);
  const ApplyBufs bu{scratch, scratch + nv, scratch + nv + nc};
  enqueue(dVals, F, dOut, s, true, nullptr, &bu);
  applied_ = true;
}

// The apply sequence as a HIP graph, captured on first use for these
// pointers and field count and replayed afterwards on a private stream
// ordered after `s` (before and after) through events. No allocation,
// host synchronisation or memset happens inside the captured sequence.
void RbfPumPlan::applyGraph(const double *dVals, int F, double *dOut,
                            hipStream_t s) {
  TraceRange tr("pacman::apply (graph)");
  if (M_ == 0)
    return;
  prepare(F, s);
  if (!gstream_) {
    PACMAN_HIP_CHECK(hipStreamCreateWithFlags(&gstream_, hipStreamNonBlocking));
    for (auto &e : gev_)
      PACMAN_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (!gexec_ || gVals_ != dVals || gOut_ != dOut || gF_ != F ||
      gGen_ != bufGen_) {
    releaseGraph();
    PACMAN_HIP_CHECK(
        hipStreamBeginCapture(gstream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(dVals, F, dOut, gstream_, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(gstream_, &g);
      if (g)
        (void)hipGraphDestroy(g);
      throw;
    }
    PACMAN_HIP_CHECK(hipStreamEndCapture(gstream_, &graph_));
    PACMAN_HIP_CHECK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
    gVals_ = dVals;
    gOut_ = dOut;
    gF_ = F;
    gGen_ = bufGen_;
  }
  PACMAN_HIP_CHECK(hipEventRecord(gev_[0], s));
  PACMAN_HIP_CHECK(hipStreamWaitEvent(gstream_, gev_[0], 0));
  PACMAN_HIP_CHECK(hipGraphLaunch(gexec_, gstream_));
  PACMAN_HIP_CHECK(hipEventRecord(gev_[1], gstream_));
  PACMAN_HIP_CHECK(hipStreamWaitEvent(s, gev_[1], 0));
}

void RbfPumPlan::releaseGraph() {
  if (gexec_)
    (void)hipGraphExecDestroy(gexec_);
  if (graph_)
    (void)hipGraphDestroy(graph_);
  gexec_ = nullptr;
  graph_ = nullptr;
  gVals_ = nullptr;
  gOut_ = nullptr;
  gF_ = 0;
}

SolveArgs RbfPumPlan::solveArgs(int F, const double *vals) const {
  SolveArgs a{};
  a.srcOff = srcOff_.get();
  a.srcIdx = srcIdx_.get();
  a.sx = gs_.view.x;
  a.sy = gs_.view.y;
  a.sz = gs_.view.z;
  a.vals = vals;
  a.N = N_;
  a.F = F;
  a.pairOff = pairs_.off.get();
  a.pairTgt = pairs_.tgt.get();
  a.pairSlot = pairs_.slot.get();
  a.tx = gt_.view.x;
  a.ty = gt_.view.y;
  a.tz = gt_.view.z;
  a.contrib = contrib_.get();
  a.cx = cx_.get();
  a.cy = cy_.get();
  a.cz = cz_.get();
  a.rinv = 1.0 / r_;
  a.dinv = 1.0 / delta_;
  a.dim = prm_.dim;
  a.pvOut = pv_.get();
  a.status = status_.get();
  a.M = M_;
  a.opRow0 = -1;
  a.phi0 = kPhi0[prm_.kind];
  a.singular = counters_.get() + kWsSingular;
  a.patchDinv = patchDinv_.n ? patchDinv_.get() : nullptr;
  a.patchRinv = patchRinv_.n ? patchRinv_.get() : nullptr;
  a.alphaHeavy = alphaHeavy_.get();
  a.slotXyz = slotXyz_.n ? slotXyz_.get() : nullptr;
  return a;
}

BlendArgs RbfPumPlan::blendArgs(int F) const {
  BlendArgs bl{};
  bl.toff = pairs_.toff.get();
  bl.slotPatch = pairs_.slotPatch.get();
  bl.slotMember = pairs_.slotMember.get();
  bl.contrib = contrib_.get();
  bl.tx = gt_.view.x;
  bl.ty = gt_.view.y;
  bl.tz = gt_.view.z;
  bl.tperm = gt_.view.perm;
  bl.tinv = tinv_.get();
  bl.wslot = wslot_.n ? wslot_.get() : nullptr;
  bl.cx = cx_.get();
  bl.cy = cy_.get();
  bl.cz = cz_.get();
  bl.M = M_;
  bl.F = F;
  bl.rinv = 1.0 / r_;  // interpolator.hpp:91
  bl.patchRinv = patchRinv_.n ? patchRinv_.get() : nullptr;
  return bl;
}

// Transpose of apply: grad_f = sum_k S_k^T [gamma_k + Q_k G_k^-1 (V_k^T h_k -
// Q_k^T gamma_k)], gamma_k = A_k^-1 E_k^T h_k, h_k = w_k g at the member
// targets of patch k (the slot weights of the blend). Deterministic: every
// source sums its patch entries in ascending patch order.
void RbfPumPlan::applyTranspose(const double *dG, int F, double *dGrad,
                                hipStream_t s) {
  TraceRange tr("pacman::apply transpose");
  require_register_patches("the transpose apply");
  if (F < 1)
    throw std::invalid_argument("transpose apply needs at least one field");
  if (N_ == 0)
    return;
  const int64_t total = std::max<int64_t>(csrSize(0), 1);
  const int perLaunch = solve_fields_per_launch();
  const int Fc = std::min(F, perLaunch);
  if (!adjReady_) {
    ensure_wslot(s);
    srcTOff_.alloc(N_ + 1);
    srcTEnt_.alloc(total);
    build_source_transpose(srcIdx_.get(), csrSize(0), N_, srcTOff_.get(),
                           srcTEnt_.get(), s);
    adjReady_ = true;
  }
  if (gSorted_.n < (size_t)(Fc * std::max<int64_t>(M_, 1)))
    gSorted_.alloc((size_t)Fc * std::max<int64_t>(M_, 1));
  if (gamma_.n < (size_t)(total * Fc))
    gamma_.alloc((size_t)total * Fc);
  if (contribS_.n < (size_t)(total * Fc))
    contribS_.alloc((size_t)total * Fc);
  for (int f0 = 0; f0 < F; f0 += perLaunch) {
    const int nf = std::min(perLaunch, F - f0);
    const Bins &B = bins_for(nf, s);
    launch_gather_values(dG + (int64_t)f0 * M_, gt_.view.perm, M_, nf,
                         gSorted_.get(), s);
    SolveArgs a = solveArgs(nf, nullptr);
    a.wslot = wslot_.get();
    a.gSorted = gSorted_.get();
    a.gammaOut = gamma_.get();
    if (opBuilt_) {
      // the stored rows already hold the weights: a column sum per source
      launch_op_transpose(a, K_, opOff_.get(), opStore(), contribS_.get(), s);
    } else {
      launch_bins(B, a, s, [&](int R, const SolveArgs &ab, hipStream_t sb) {
        launch_solve_f64_adjoint(R, prm_.kind, ab, sb);
      });
      launch_adjoint_poly(a, K_, gamma_.get(), contribS_.get(), s);
    }
    launch_source_reduce(srcTOff_.get(), srcTEnt_.get(), N_, nf,
                         contribS_.get(), gs_.view.perm,
                         dGrad + (int64_t)f0 * N_, s);
  }
}

int64_t RbfPumPlan::stateSize(int F) const {
  return (int64_t)std::max(F, 0) * (totalSrc_ + 4 * K_);
}

void RbfPumPlan::applyState(const double *dVals, int F, double *dOut,
                            double *dState, hipStream_t s) {
  if (M_ == 0)
    return;
  if (dState == nullptr && stateSize(F) > 0)
    throw std::invalid_argument("null state buffer");
  prepare(F, s);
  enqueue(dVals, F, dOut, s, true, dState);
  applied_ = true;
}

// Per chunk of nf fields (as the apply solved them): slot sums of g times the
// local value and of g times its gradient, accumulated over the chunks, then
// one blend with the derivative of the Shepard weights.
void RbfPumPlan::targetCoordGrad(const double *dState, int F, const double *dG,
                                 double *dCoordGrad, hipStream_t s) {
  TraceRange tr("pacman::target coordinate gradient");
  require_register_patches("the target-coordinate gradient");
  if (prm_.adaptive)
    throw std::runtime_error("the target-coordinate gradient is available "
                             "for plans with one patch radius (not adaptive)");
  if (F < 1)
    throw std::invalid_argument("target gradient needs at least one field");
  if (M_ == 0)
    return;
  const int64_t total = totalSrc_;
  const int perLaunch = solve_fields_per_launch();
  const int Fc = std::min(F, perLaunch);
  if (gSorted_.n < (size_t)(Fc * M_))
    gSorted_.alloc((size_t)Fc * M_);
  if (gslot_.n < (size_t)(4 * std::max<int64_t>(pairs_.P, 1)))
    gslot_.alloc((size_t)4 * std::max<int64_t>(pairs_.P, 1));
  for (int f0 = 0; f0 < F; f0 += perLaunch) {
    const int nf = std::min(perLaunch, F - f0);
    launch_gather_values(dG + (int64_t)f0 * M_, gt_.view.perm, M_, nf,
                         gSorted_.get(), s);
    SolveArgs a = solveArgs(nf, nullptr);
    a.gSorted = gSorted_.get();
    launch_target_grad_eval(a, K_, prm_.kind, dState + (int64_t)f0 * total,
                            dState + (int64_t)F * total + (int64_t)f0 * K_ * 4,
                            gslot_.get(), f0 > 0, s);
  }
  launch_target_grad_blend(blendArgs(1), gslot_.get(), prm_.dim, dCoordGrad,
                           s);
}

// The local operator of every patch, scaled by the Shepard weights, kept in
// HBM: row q of patch k maps the patch's source values to w * t at its pair
// q. Built with the transpose chain (the same factorisation) on unit
// upstream values, eight rows per launch; stored in FP64, or rounded to FP32
// (halves its memory and the bytes each apply reads; applied with FP64
// accumulation).
void RbfPumPlan::buildOperator(hipStream_t s, bool fp32) {
  TraceRange tr("pacman::build operator");
  require_register_patches("the stored operator");
  releaseGraph();
  op_.release();
  op32_.release();
  opBuilt_ = false;
  if (M_ == 0)
    return;
  std::vector<int64_t> poff(K_ + 1), ooff(K_ + 1, 0);
  PACMAN_HIP_CHECK(hipMemcpyAsync(poff.data(), pairs_.off.get(),
                                  8 * (K_ + 1), hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  int64_t maxRows = 0;
  for (int64_t k = 0; k < K_; ++k) {
    const int64_t m = poff[k + 1] - poff[k];
    maxRows = std::max(maxRows, m);
    ooff[k + 1] = ooff[k] + m * patchN_[k];
  }
  opOff_.alloc(K_ + 1);
  PACMAN_HIP_CHECK(hipMemcpyAsync(opOff_.get(), ooff.data(), 8 * (K_ + 1),
                                  hipMemcpyHostToDevice, s));
  if (fp32)
    op32_.alloc(std::max<int64_t>(ooff[K_], 1));
  else
    op_.alloc(std::max<int64_t>(ooff[K_], 1));
  ensure_tinv(2, s);
  const int Fb = solve_fields_per_launch();
  const int64_t total = std::max<int64_t>(totalSrc_, 1);
  ensure_wslot(s);
  if (gamma_.n < (size_t)(total * Fb))
    gamma_.alloc((size_t)total * Fb);
  if (contribS_.n < (size_t)(total * Fb))
    contribS_.alloc((size_t)total * Fb);
  const Bins &B = bins_for(Fb, s);
  SolveArgs a = solveArgs(Fb, nullptr);
  a.wslot = wslot_.get();
  a.gammaOut = gamma_.get();
  for (int64_t r0 = 0; r0 < maxRows; r0 += Fb) {
    a.opRow0 = (int)r0;
    launch_bins(B, a, s, [&](int R, const SolveArgs &ab, hipStream_t sb) {
      launch_solve_f64_adjoint(R, prm_.kind, ab, sb);
    });
    launch_adjoint_poly(a, K_, gamma_.get(), contribS_.get(), s);
    launch_op_store(a, K_, contribS_.get(), opOff_.get(), opStore(), s);
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  opBuilt_ = true;
}

void RbfPumPlan::dropOperator() {
  releaseGraph();
  op_.release();
  op32_.release();
  opOff_.release();
  opBuilt_ = false;
}

int64_t RbfPumPlan::operatorBytes() const {
  return opBuilt_ ? (int64_t)(op_.n * sizeof(double) + op32_.n * sizeof(float))
                  : 0;
}

OpStore RbfPumPlan::opStore() const {
  OpStore o;
  if (op32_.n)
    o.f32 = op32_.get();
  else
    o.f64 = op_.get();
  return o;
}

void RbfPumPlan::enqueue(const double *dVals, int F, double *dOut,
                         hipStream_t s, bool events, double *dState,
                         const ApplyBufs *bufs) {
  const ApplyBufs bu =
      bufs ? *bufs : ApplyBufs{valsSorted_.get(), contrib_.get(), coef_.get()};
  const bool useOp = opBuilt_ && dState == nullptr;
  // the stored operator takes up to kOpFMax fields per pass over it
  const int perLaunch = useOp ? kOpFMax : solve_fields_per_launch();
  const int64_t total = totalSrc_;
  if (events)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[0], s));
  launch_ws_reset(counters_.get(), true, s);
  launch_gather_values(dVals, gs_.view.perm, N_, F, bu.vals, s);
  if (events)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[1], s));
  // a single chunk blends after all solves (phases stay separable)
  BlendArgs pending{};
  bool havePending = false;
  for (int f0 = 0; f0 < F; f0 += perLaunch) {
    const int nf = std::min(perLaunch, F - f0);
    SolveArgs a = solveArgs(nf, bu.vals + (int64_t)f0 * N_);
    a.contrib = bu.contrib;
    BlendArgs bl = blendArgs(nf);
    bl.contrib = bu.contrib;
    bl.out = dOut + (int64_t)f0 * M_;
    if (useOp) {
      if (f0 == 0 && events)
        PACMAN_HIP_CHECK(hipEventRecord(ev_[4], s));
      launch_op_apply(a, K_, opOff_.get(), opStore(), s);
      if (F <= perLaunch) {
        pending = bl;
        havePending = true;
      } else {
        launch_blend_sum(bl, s);
      }
      continue;
    }
    launch_poly(a, K_, bu.coef, s);
    if (f0 == 0 && events)
      PACMAN_HIP_CHECK(hipEventRecord(ev_[4], s));
    if (dState) {
      a.alphaOut = dState + (int64_t)f0 * total;
      if (K_ > 0)
        PACMAN_HIP_CHECK(hipMemcpyAsync(
            dState + (int64_t)F * total + (int64_t)f0 * K_ * 4, bu.coef,
            sizeof(double) * (size_t)K_ * nf * 4, hipMemcpyDeviceToDevice, s));
    }
    launch_bins(bins_[nf - 1], a, s, [&](int R, const SolveArgs &ab,
                                         hipStream_t sb) {
      launch_solve_f64(R, kTune.waves[nf - 1][R], prm_.kind, ab, bu.coef, sb);
    });
    if (f0 == 0 && events)
      PACMAN_HIP_CHECK(hipEventRecord(ev_[7], s));
    if (f0 == 0 && events)
      PACMAN_HIP_CHECK(hipEventRecord(ev_[8], s));
    // the register-resident patches' evaluation on a second stream, beside
    // the workspace kernel (one workgroup per CU leaves room for its waves)
    const bool side = !wsHost_.empty();
    hipStream_t es = side ? bstream_[0] : s;
    if (side) {
      PACMAN_HIP_CHECK(hipEventRecord(efork_, s));
      PACMAN_HIP_CHECK(hipStreamWaitEvent(es, efork_, 0));
    }
    for (int i = 2; i >= 0; --i)
      launch_wave_eval(prm_.kind, a, bu.coef,
                       evalChunks_.get() + evalFirst_[0][i],
                       nEvalChunks_[0][i], i + 1, es);
    if (!wsHost_.empty()) {
      if (f0 > 0)
        launch_ws_reset(counters_.get(), false, s);
      launch_workspace(a, bu.coef, s, events && f0 == 0,
                       workspace_fp32(nf));
    }
    for (int i = 2; i >= 0; --i)
      launch_wave_eval(prm_.kind, a, bu.coef,
                       evalChunks_.get() + evalFirst_[1][i],
                       nEvalChunks_[1][i], i + 1, s);
    if (side) {
      PACMAN_HIP_CHECK(hipEventRecord(ejoin_, es));
      PACMAN_HIP_CHECK(hipStreamWaitEvent(s, ejoin_, 0));
    }
    if (f0 == 0 && events)
      PACMAN_HIP_CHECK(hipEventRecord(ev_[9], s));
    if (F <= perLaunch) {
      pending = bl;
      havePending = true;
    } else {
      launch_blend(bl, s);
    }
  }
  if (events)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[2], s));
  if (havePending) {
    if (useOp)
      launch_blend_sum(pending, s);
    else
      launch_blend(pending, s);
  }
  if (events)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[3], s));
}

void RbfPumPlan::centers(double *xyz, int64_t *candidate, hipStream_t s) const {
  std::vector<double> x(K_), y(K_), z(K_);
  if (K_ > 0) {
    PACMAN_HIP_CHECK(hipMemcpyAsync(x.data(), cx_.get(), 8 * K_,
                                    hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipMemcpyAsync(y.data(), cy_.get(), 8 * K_,
                                    hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipMemcpyAsync(z.data(), cz_.get(), 8 * K_,
                                    hipMemcpyDeviceToHost, s));
    if (candidate)
      PACMAN_HIP_CHECK(hipMemcpyAsync(candidate, cand_.get(), 8 * K_,
                                      hipMemcpyDeviceToHost, s));
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  for (int64_t k = 0; k < K_; ++k) {
    xyz[3 * k] = x[k];
    xyz[3 * k + 1] = y[k];
    xyz[3 * k + 2] = z[k];
  }
}

int64_t RbfPumPlan::csrSize(int which) const {
  if (which == 0) {
    int64_t tot = 0;
    PACMAN_HIP_CHECK(hipMemcpy(&tot, srcOff_.get() + K_, 8,
                               hipMemcpyDeviceToHost));
    return tot;
  }
  if (which == 2)
    return pairs_.P;
  // member pairs only
  std::vector<int> t(pairs_.P);
  if (pairs_.P)
    PACMAN_HIP_CHECK(hipMemcpy(t.data(), pairs_.tgt.get(), 4 * pairs_.P,
                               hipMemcpyDeviceToHost));
  int64_t c = 0;
  for (int v : t)
    c += ((unsigned)v >> 31) ? 1 : 0;
  return c;
}

void RbfPumPlan::csr(int which, int64_t *off, int64_t *idx,
                     hipStream_t s) const {
  if (which == 0) {
    std::vector<int64_t> o(K_ + 1);
    PACMAN_HIP_CHECK(hipMemcpyAsync(o.data(), srcOff_.get(), 8 * (K_ + 1),
                                    hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int> pos(o[K_]), perm(N_);
    if (o[K_])
      PACMAN_HIP_CHECK(hipMemcpy(pos.data(), srcIdx_.get(), 4 * o[K_],
                                 hipMemcpyDeviceToHost));
    PACMAN_HIP_CHECK(hipMemcpy(perm.data(), gs_.view.perm, 4 * N_,
                               hipMemcpyDeviceToHost));
    for (int64_t k = 0; k <= K_; ++k)
      off[k] = o[k];
    for (int64_t k = 0; k < K_; ++k) {
      for (int64_t e = o[k]; e < o[k + 1]; ++e)
        idx[e] = perm[pos[e]];
      std::sort(idx + o[k], idx + o[k + 1]);
    }
    return;
  }
  std::vector<int64_t> po(K_ + 1);
  std::vector<int> t(pairs_.P), tperm(M_);
  PACMAN_HIP_CHECK(hipMemcpy(po.data(), pairs_.off.get(), 8 * (K_ + 1),
                             hipMemcpyDeviceToHost));
  if (pairs_.P)
    PACMAN_HIP_CHECK(hipMemcpy(t.data(), pairs_.tgt.get(), 4 * pairs_.P,
                               hipMemcpyDeviceToHost));
  if (M_)
    PACMAN_HIP_CHECK(hipMemcpy(tperm.data(), gt_.view.perm, 4 * M_,
                               hipMemcpyDeviceToHost));
  if (which == 1) {
    int64_t w = 0;
    off[0] = 0;
    for (int64_t k = 0; k < K_; ++k) {
      const int64_t b = w;
      for (int64_t e = po[k]; e < po[k + 1]; ++e)
        if ((unsigned)t[e] >> 31)
          idx[w++] = tperm[t[e] & 0x7fffffff];
      std::sort(idx + b, idx + w);
      off[k + 1] = w;
    }
    return;
  }
  // which == 2: per original target, the patches within r (ascending)
  std::vector<std::vector<int64_t>> lists(M_);
  for (int64_t k = 0; k < K_; ++k)
    for (int64_t e = po[k]; e < po[k + 1]; ++e)
      lists[tperm[t[e] & 0x7fffffff]].push_back(k);
  off[0] = 0;
  for (int64_t i = 0; i < M_; ++i) {
    std::sort(lists[i].begin(), lists[i].end());
    std::copy(lists[i].begin(), lists[i].end(), idx + off[i]);
    off[i + 1] = off[i] + (int64_t)lists[i].size();
  }
}

void RbfPumPlan::patchInfo(int *pv, int *status, hipStream_t s) const {
  if (K_ == 0)
    return;
  if (pv)
    PACMAN_HIP_CHECK(hipMemcpyAsync(pv, pv_.get(), 4 * K_,
                                    hipMemcpyDeviceToHost, s));
  if (status)
    PACMAN_HIP_CHECK(hipMemcpyAsync(status, status_.get(), 4 * K_,
                                    hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

// Patches above tileMaxN_ sources, largest first (ties: lower patch index),
// and the per-slot workspaces of their persistent launches: a slot holds the
// lower block triangle of the largest patch of its queue plus its Linv
// blocks (FP64 size; FP32 storage uses the first half) and 4 + 4 F vectors.
// Queue 0: bordered matrices (n + 8 rows) within the LDS variant's limit;
// queue 1: larger; shape 2 (the fallback pass) covers both. Slots: one per
// CU (the kernel's register budget allows one workgroup per CU), fewer if
// the workspaces would exceed the budget
// (PACMAN_AMD_WS_BUDGET_MB, default 8192). Launches are stream-ordered, so
// the queues share one buffer.
void RbfPumPlan::plan_workspace(hipStream_t s) {
  TraceRange tr("pacman::plan workspace queues");
  wsHost_.clear();
  for (auto &q : wsQ_) {
    q.host.clear();
    q.nbMax = q.slots = 0;
  }
  for (int64_t k = 0; k < K_; ++k)
    if (patchN_[k] > tileMaxN_)
      wsHost_.push_back((int)k);
  if (wsHost_.empty())
    return;
  std::stable_sort(wsHost_.begin(), wsHost_.end(), [&](int x, int y) {
    return patchN_[x] > patchN_[y];
  });
  const int fmax = solve_fields_per_launch();
  auto rows = [&](int k) { return (patchN_[k] + fmax + kWsBlock - 1) / kWsBlock; };
  for (int k : wsHost_) {
    WsQueue &q = wsQ_[rows(k) <= ws_lds_block_rows() ? 0 : 1];
    q.host.push_back(k);
    q.nbMax = std::max(q.nbMax, rows(k));
  }
  wsQ_[2].nbMax = std::max(wsQ_[0].nbMax, wsQ_[1].nbMax);
  int cu = 0;
  int dev = 0;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  PACMAN_HIP_CHECK(
      hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  double budgetMb = 8192.0;
  if (const char *e = std::getenv("PACMAN_AMD_WS_BUDGET_MB"))
    budgetMb = std::max(1.0, std::atof(e));
  int64_t workMax = 0, vecMax = 0;
  for (int i = 0; i < 3; ++i) {
    WsQueue &q = wsQ_[i];
    const int64_t count = i == 2 ? (int64_t)wsHost_.size()
                                 : (int64_t)q.host.size();
    if (count == 0 || q.nbMax == 0)
      continue;
    const int64_t slotBytes =
        8 * (ws_slot_blocks(q.nbMax) + ws_slot_vec(q.nbMax));
    const int64_t bySize = std::max<int64_t>(
        1, (int64_t)(budgetMb * 1048576.0) / std::max<int64_t>(slotBytes, 1));
    q.slots = (int)std::min<int64_t>(
        {count, kWsSlotsPerCu * (int64_t)std::max(cu, 1), bySize});
    workMax = std::max(workMax, q.slots * ws_slot_blocks(q.nbMax));
    vecMax = std::max(vecMax, q.slots * ws_slot_vec(q.nbMax));
    if (i < 2 && !q.host.empty()) {
      q.dev.alloc(q.host.size());
      PACMAN_HIP_CHECK(hipMemcpyAsync(q.dev.get(), q.host.data(),
                                      4 * q.host.size(),
                                      hipMemcpyHostToDevice, s));
    }
  }
  wsFallback_.alloc(wsHost_.size());
  wsWork_.alloc((size_t)workMax);
  wsVec_.alloc((size_t)vecMax);
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));  // queues copied
}

// The solve kernels store alpha (alphaHeavy_, n x 8 doubles reserved per
// patch) and leave the evaluation of every patch to the wave evaluation:
//  * inside the register-resident solve the evaluation ran at the solve's
//    occupancy (two workgroups per CU, 8 or 16 waves each), where its
//    dependent pair, target and slot loads left it latency-bound (about a
//    quarter of the solve's time);
//  * inside the workspace kernel (one workgroup per CU) it lengthened the
//    kernel by half (3.5 -> 2.1 ms without it, uneven cloud);
//  * patches with more than kHeavyPairs (target, patch) pairs (large cells
//    of sparse regions next to dense ones: up to ~1e5 pairs) would keep one
//    workgroup evaluating for milliseconds.
// They are evaluated apart, one wave per chunk of u units of 64 pairs
// (u pairs per lane): a patch of one unit in one chunk, the others in
// chunks of two units and, for an odd count, one last chunk of three. The
// chunks of the register-resident patches run beside the workspace kernel
// (their alpha is ready once the size bins are done), those of workspace
// patches after it.
void RbfPumPlan::plan_eval(hipStream_t s) {
  TraceRange tr("pacman::plan evaluation chunks");
  // chunks per patch: units u of 64 pairs -> one chunk of one unit, or
  // chunks of two units and, for an odd u, a last one of three
  plan_eval_chunks(srcOff_.get(), pairs_.off.get(), K_, tileMaxN_,
                   evalChunks_, evalFirst_, nEvalChunks_, s);
  if (K_ == 0)
    return;
  // alpha records: 8 doubles reserved per member slot (patch k's at
  // 8 srcOff[k]), plus one spare record
  alphaHeavy_.alloc((size_t)8 * totalSrc_ + solve_fields_per_launch());
  slotXyz_.alloc((size_t)4 * (totalSrc_ + 1));  // one spare record
  launch_slot_xyz(srcIdx_.get(), totalSrc_, gs_.view.x, gs_.view.y,
                  gs_.view.z, slotXyz_.get(), s);
}

void RbfPumPlan::launch_workspace(const SolveArgs &a, const double *coef,
                                  hipStream_t s, bool timed, bool fp32) {
  const bool mixed = fp32;
  auto args = [&](const WsQueue &q) {
    WsArgs w{};
    w.a = a;
    w.coef = coef;
    w.nbMax = q.nbMax;
    w.work = wsWork_.get();
    w.slotBlocks = ws_slot_blocks(q.nbMax);
    w.linvOffset = ws_linv_offset(q.nbMax);
    w.vec = wsVec_.get();
    w.slotVec = ws_slot_vec(q.nbMax);
    w.counters = counters_.get();
    w.fallback = wsFallback_.get();
    w.countPtr = nullptr;
    w.queueIdx = kWsQueue;
    w.maxIR = prm_.maxIR;
    w.phi0 = kPhi0[prm_.kind];
    return w;
  };
  if (timed)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[5], s));
  for (int i = 0; i < 2; ++i) {
    const WsQueue &q = wsQ_[i];
    if (q.host.empty())
      continue;
    if (i == 1)  // the queue counter starts again at zero
      launch_ws_reset(counters_.get(), false, s, /*keepFallbacks=*/true);
    WsArgs w = args(q);
    w.a.patches = q.dev.get();
    w.a.count = (int)q.host.size();
    launch_ws_solve(mixed, prm_.kind, w, q.slots, s);
  }
  if (timed)
    PACMAN_HIP_CHECK(hipEventRecord(ev_[6], s));
  wsTimed_ = wsTimed_ || timed;
  if (mixed) {
    // FP64 pass over the patches the refinement handed back (the list
    // length is read on the device; an empty list ends every workgroup at
    // once)
    WsArgs w = args(wsQ_[2]);
    w.a.patches = wsFallback_.get();
    w.countPtr = counters_.get() + kWsFallbacks;
    w.queueIdx = kWsQueue2;
    launch_ws_solve(false, prm_.kind, w, wsQ_[2].slots, s);
  }
}

bool RbfPumPlan::workspace_fp32(int F) const {
  return prm_.precision == 1 || (prm_.precision == 2 && kTune.wsFp32[F - 1]);
}

double RbfPumPlan::timeSolve(const double *dVals, int F, int R, int waves,
                             int wsPrecision, int reps, hipStream_t s) {
  if (F < 1 || F > solve_fields_per_launch())
    throw std::invalid_argument("tuning: 1 <= fields <= 8");
  if (reps < 1 || reps > 1000)
    throw std::invalid_argument("tuning: 1 <= repetitions <= 1000");
  if (wsPrecision != 0 && wsPrecision != 1)
    throw std::invalid_argument("tuning: workspace precision 0 or 1");
  if (M_ == 0)
    return -1.0;
  if (R > 0 && !solve_tile_variant(R, waves))
    return -1.0;
  if (R == 0 && wsHost_.empty())
    return -1.0;
  prepare(F, s);
  const Bins &B = bins_for(F, s);
  int b = -1;
  for (size_t i = 0; i < B.R.size(); ++i)
    if (B.R[i] == R)
      b = (int)i;
  if (R > 0 && b < 0)
    return -1.0;
  SolveArgs a = solveArgs(F, valsSorted_.get());
  a.contrib = contrib_.get();
  launch_gather_values(dVals, gs_.view.perm, N_, F, valsSorted_.get(), s);
  launch_poly(a, K_, coef_.get(), s);
  hipEvent_t e[2];
  PACMAN_HIP_CHECK(hipEventCreate(&e[0]));
  PACMAN_HIP_CHECK(hipEventCreate(&e[1]));
  std::vector<float> ms;
  try {
    for (int r = 0; r < reps; ++r) {
      launch_ws_reset(counters_.get(), true, s);
      PACMAN_HIP_CHECK(hipEventRecord(e[0], s));
      if (R > 0) {
        a.patches = B.dev.get() + B.first[b];
        a.count = B.count[b];
        launch_solve_f64(R, waves, prm_.kind, a, coef_.get(), s);
      } else {
        launch_workspace(a, coef_.get(), s, false, wsPrecision == 1);
      }
      PACMAN_HIP_CHECK(hipEventRecord(e[1], s));
      PACMAN_HIP_CHECK(hipEventSynchronize(e[1]));
      float t = 0.f;
      PACMAN_HIP_CHECK(hipEventElapsedTime(&t, e[0], e[1]));
      ms.push_back(t);
    }
  } catch (...) {
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    throw;
  }
  PACMAN_HIP_CHECK(hipEventDestroy(e[0]));
  PACMAN_HIP_CHECK(hipEventDestroy(e[1]));
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

void RbfPumPlan::setPrecision(int precision, int maxIR) {
  if (precision < 0 || precision > 2)
    throw std::invalid_argument(
        "precision must be 0 (FP64), 1 (mixed) or 2 (auto)");
  if (maxIR < 0 || maxIR > 100)
    throw std::invalid_argument("refinement iteration limit out of range");
  releaseGraph();
  prm_.precision = precision;
  prm_.maxIR = maxIR;
}

void RbfPumPlan::applyCounters(int *out, hipStream_t s) const {
  int c[kWsCounters];
  PACMAN_HIP_CHECK(hipMemcpyAsync(c, counters_.get(), sizeof(c),
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  out[0] = c[kWsSingular];
  out[1] = c[kWsIrIters];
  out[2] = c[kWsFallbackTotal];
}

void RbfPumPlan::patchRadii(double *r, hipStream_t s) const {
  if (K_ == 0)
    return;
  if (!patchR_.n) {
    for (int64_t k = 0; k < K_; ++k)
      r[k] = r_;
    return;
  }
  PACMAN_HIP_CHECK(hipMemcpyAsync(r, patchR_.get(), 8 * K_,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

void RbfPumPlan::require_register_patches(const char *what) const {
  if (maxN_ > kMaxPatch) {
    int64_t larger = 0;
    for (int n : patchN_)
      larger += n > kMaxPatch ? 1 : 0;
    throw std::runtime_error(
        std::string(what) + " is available for patches of up to " +
        std::to_string(kMaxPatch) + " source points; this plan has " +
        std::to_string(larger) + " larger patches (largest " +
        std::to_string(maxN_) + ")");
  }
}

} // namespace pacman_amd
