// Density-aware ("adaptive") patches on the GPU. Not in the reference: its
// single patch radius from FindRadius (src/rbf_pum/clustering/
// clusters_radius.hpp:22-125) gives, on strongly clustered clouds, patches
// of one to ~1e5 points and leaves sparse regions uncovered. The steps and
// every floating-point expression follow the oracle's runAdaptiveSetup
// (compiled here with -ffp-contract=off, so centres, radii and member lists
// are bit-identical):
//
//  1. box of sources and targets; 16-bit-per-axis Morton codes (quantised
//     coordinate q = min(65535, trunc((x - lo) / ext * 65536)), bits
//     interleaved from the most significant, axis 0 first); a stable LSD
//     radix sort (8-bit digits) orders each point set by code, which also
//     gives the Morton-ordered SoA coordinates the solve reads;
//  2. a binary tree over the code bits, one launch per depth: a cell with
//     more than `leaf` sources is halved (the split position is a binary
//     search in the sorted codes), cells with neither sources nor targets
//     are dropped, leaves with targets become patches;
//  3. per patch (one wave): centre = cell centre, rc = (1 + overlap) * half
//     the cell diagonal; when fewer than nmin sources lie within rc the
//     radius is the distance to the nmin-th nearest source (bisection on
//     the bits of the squared distances over an exact candidate set);
//  4. patches holding more than nmax sources are halved again;
//  5. patches sorted by cell key; members: sources within r_k, targets
//     within r_k (weights) and r_k - 1e-14 (members); the few target balls
//     of sparse regions expected to hold more than 16384 targets are scanned
//     by a workgroup each instead of a wave.
//
// Ball queries read the sorted codes: the 3^D cells around the centre at
// the deepest depth whose cells are at least (1 + 1e-9) r wide on every
// axis hold every point within r; each cell is one contiguous code range
// (two binary searches, one lane each), scanned by the wave.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <map>
#include <mutex>
#include <cmath>
#include <vector>

#include "rbf_internal.hpp"

namespace pacman_amd {
namespace {

constexpr int kBits = 16;
constexpr int kThreads = 256;

struct Frame {
  double lo[3], ext[3];
  int D, DB;
  double hmin[3 * kBits + 1];  // smallest cell edge at each depth
};

// A point set in code order. tab[c] = the first position whose code is at
// least c << tabShift, for the 2^tabBits coarse cells of the code's leading
// bits and c = 2^tabBits (= n): a lower bound in the whole set starts inside
// the coarse cell of its key (an L2-resident table read instead of ~22
// dependent loads over all n codes).
struct SortedSet {
  const uint64_t *code;
  const double *x, *y, *z;
  int n;
  const int *tab;
  int tabShift, tabBits;
};

__host__ __device__ inline uint64_t quantize(double x, double lo, double ext) {
  const double t = ((x - lo) / ext) * 65536.0;
  if (!(t > 0.0))
    return 0;
  if (t >= 65535.0)
    return 65535;
  return (uint64_t)t;
}

__host__ __device__ inline int axis_bits(int d, int D, int a) {
  return (d + D - 1 - a) / D;
}

// Bit k of a 16-bit axis coordinate moves to bit D k (and back): the code
// holds axis a's bit k at position D k + (D - 1 - a), axis 0 first from the
// most significant bit.
__host__ __device__ inline uint64_t spread_bits(uint64_t x, int D) {
  x &= 0xFFFFull;
  if (D == 3) {
    x = (x | (x << 16)) & 0x0000FF0000FFull;
    x = (x | (x << 8)) & 0x00F00F00F00Full;
    x = (x | (x << 4)) & 0x0C30C30C30C3ull;
    x = (x | (x << 2)) & 0x249249249249ull;
  } else if (D == 2) {
    x = (x | (x << 8)) & 0x00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0Full;
    x = (x | (x << 2)) & 0x33333333ull;
    x = (x | (x << 1)) & 0x55555555ull;
  }
  return x;
}

__host__ __device__ inline uint64_t compact_bits(uint64_t x, int D) {
  if (D == 3) {
    x &= 0x249249249249ull;
    x = (x ^ (x >> 2)) & 0x0C30C30C30C3ull;
    x = (x ^ (x >> 4)) & 0x00F00F00F00Full;
    x = (x ^ (x >> 8)) & 0x0000FF0000FFull;
    x = (x ^ (x >> 16)) & 0xFFFFull;
  } else if (D == 2) {
    x &= 0x55555555ull;
    x = (x ^ (x >> 1)) & 0x33333333ull;
    x = (x ^ (x >> 2)) & 0x0F0F0F0Full;
    x = (x ^ (x >> 4)) & 0x00FF00FFull;
    x = (x ^ (x >> 8)) & 0xFFFFull;
  } else {
    x &= 0xFFFFull;
  }
  return x;
}

// Prefix (the leading d code bits) of the cell at depth d whose axis-a
// index is j[a] (axis_bits(d, D, a) bits): j[a] as the leading bits of each
// axis coordinate, interleaved, cut to d bits.
__host__ __device__ inline uint64_t cell_prefix(const uint64_t *j, int D,
                                                int d) {
  uint64_t c = 0;
  for (int a = 0; a < D; ++a) {
    const int ba = axis_bits(d, D, a);
    c |= spread_bits(j[a] << (kBits - ba), D) << (D - 1 - a);
  }
  return c >> (kBits * D - d);
}

// Axis-a index of the cell with prefix p at depth d (cell_prefix inverted).
__host__ __device__ inline uint64_t cell_axis_index(uint64_t p, int d, int D,
                                                    int a) {
  const int ba = axis_bits(d, D, a);
  if (ba == 0)
    return 0;
  const uint64_t c = p << (kBits * D - d);
  return compact_bits(c >> (D - 1 - a), D) >> (kBits - ba);
}

__device__ inline int lower_bound_u64(const uint64_t *a, int lo, int hi,
                                      uint64_t key) {
  while (lo < hi) {
    const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
    if (a[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First position in the whole of S whose code is >= key (as
// lower_bound_u64(S.code, 0, S.n, key)).
__device__ inline int code_lower_bound(const SortedSet &S, uint64_t key) {
  const uint64_t c = key >> S.tabShift;
  if (c >= ((uint64_t)1 << S.tabBits))
    return S.n;
  return lower_bound_u64(S.code, S.tab[c], S.tab[c + 1], key);
}

// ---- codes and the radix sort ---------------------------------------------

__global__ void morton_codes_kernel(const double *pts, int64_t n, int dim,
                                    Frame f, uint64_t *key, int *val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  uint64_t q[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a)
    q[a] = quantize(pts[i * dim + a], f.lo[a], f.ext[a]);
  key[i] = cell_prefix(q, dim, f.DB);
  val[i] = (int)i;
}

constexpr int kRounds = 16;
constexpr int kTile = kThreads * kRounds;

__global__ __launch_bounds__(kThreads) void rs_hist_kernel(const uint64_t *key,
                                                           int64_t n, int shift,
                                                           int nblk, int *hist) {
  __shared__ int h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int r = 0; r < kRounds; ++r) {
    const int64_t i = base + (int64_t)r * kThreads + tid;
    if (i < n)
      atomicAdd(&h[(key[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(int64_t)tid * nblk + blockIdx.x] = h[tid];
}

// Stable scatter: within a round the order is wave by wave, lane by lane
// (elements in input order); rounds in order; blocks through the digit-major
// scan of the histograms.
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(
    const uint64_t *kin, const int *vin, int64_t n, int shift, int nblk,
    const int64_t *hoff, uint64_t *kout, int *vout) {
  __shared__ int64_t base[256];
  __shared__ int wcnt[kThreads / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base[tid] = hoff[(int64_t)tid * nblk + blockIdx.x];
  for (int q = tid; q < (kThreads / 64) * 256; q += kThreads)
    wcnt[q >> 8][q & 255] = 0;
  __syncthreads();
  const int64_t b0 = (int64_t)blockIdx.x * kTile;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kRounds; ++r) {
    const int64_t i = b0 + (int64_t)r * kThreads + tid;
    const bool valid = i < n;
    const uint64_t k = valid ? kin[i] : 0;
    const int v = valid ? vin[i] : 0;
    const int dig = valid ? (int)((k >> shift) & 255u) : 0;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool on = (dig >> b) & 1;
      const uint64_t bb = __ballot(valid && on);
      m &= on ? bb : ~bb;
    }
    const int rank = __popcll(m & below);
    const int tot = __popcll(m);
    if (valid && rank == tot - 1)
      wcnt[wave][dig] = tot;
    __syncthreads();
    if (valid) {
      int pre = 0;
      for (int w = 0; w < wave; ++w)
        pre += wcnt[w][dig];
      const int64_t pos = base[dig] + pre + rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
      add += wcnt[w][tid];
      wcnt[w][tid] = 0;
    }
    base[tid] += add;
    __syncthreads();
  }
}

// Sorts (key, val) by the low `bits` bits of key (a multiple of 16), stably;
// the result is back in (k0, v0).
void radix_sort_pairs(uint64_t *k0, int *v0, uint64_t *k1, int *v1, int64_t n,
                      int bits, hipStream_t s) {
  if (n <= 1)
    return;
  const int nblk = (int)((n + kTile - 1) / kTile);
  DevBuf<int> hist((size_t)256 * nblk);
  DevBuf<int64_t> hoff((size_t)256 * nblk + 1);
  uint64_t *ka = k0, *kb = k1;
  int *va = v0, *vb = v1;
  for (int shift = 0; shift < bits; shift += 8) {
    rs_hist_kernel<<<nblk, kThreads, 0, s>>>(ka, n, shift, nblk, hist.get());
    PACMAN_HIP_CHECK(hipGetLastError());
    exclusive_scan_i32_to_i64_async(hist.get(), hoff.get(), (int64_t)256 * nblk,
                                    s);
    rs_scatter_kernel<<<nblk, kThreads, 0, s>>>(ka, va, n, shift, nblk,
                                                hoff.get(), kb, vb);
    PACMAN_HIP_CHECK(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != k0) {  // odd number of passes
    PACMAN_HIP_CHECK(hipMemcpyAsync(k0, ka, 8 * n, hipMemcpyDeviceToDevice, s));
    PACMAN_HIP_CHECK(hipMemcpyAsync(v0, va, 4 * n, hipMemcpyDeviceToDevice, s));
  }
}

// tab[c] = lower_bound(code, c << shift) for c in [0, 2^bits]; tab[2^bits] = n.
__global__ void code_table_kernel(const uint64_t *code, int n, int shift,
                                  int bits, int *tab) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)1 << bits;
  if (c > nc)
    return;
  tab[c] = c == nc ? n : lower_bound_u64(code, 0, n, (uint64_t)c << shift);
}

__global__ void gather_sorted_kernel(const double *pts, int64_t n, int dim,
                                     const int *order, double *x, double *y,
                                     double *z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int64_t o = order[i];
  x[i] = pts[o * dim];
  y[i] = dim > 1 ? pts[o * dim + 1] : 0.0;
  z[i] = dim > 2 ? pts[o * dim + 2] : 0.0;
}

// ---- the tree --------------------------------------------------------------

struct TNode {
  uint64_t p;
  int d, slo, shi, tlo, thi, pad;
};

// One depth of the tree: the *inCount nodes of `in` (read on the device, so
// the depths are launched back to back without a host round trip; a
// grid-stride loop covers any count), children appended to `out`, leaves
// with targets to the leaf list. Appends past maxOut / maxLeaf are counted
// but not stored (the host checks the counts once at the end).
__global__ void tree_level_kernel(const TNode *in, const int *inCount,
                                  TNode *out, int *outCount, int maxOut,
                                  int *leafCount, uint64_t *leafP, int *leafD,
                                  int maxLeaf, SortedSet S, SortedSet T,
                                  int leaf, int DB) {
  const int nIn = min(*inCount, maxOut);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nIn;
       i += gridDim.x * blockDim.x) {
    const TNode nd = in[i];
    if (nd.shi - nd.slo > leaf && nd.d < DB) {
      const uint64_t key = (2 * nd.p + 1) << (DB - nd.d - 1);
      const int ms = lower_bound_u64(S.code, nd.slo, nd.shi, key);
      const int mt = lower_bound_u64(T.code, nd.tlo, nd.thi, key);
      const TNode c[2] = {{2 * nd.p, nd.d + 1, nd.slo, ms, nd.tlo, mt, 0},
                          {2 * nd.p + 1, nd.d + 1, ms, nd.shi, mt, nd.thi, 0}};
      for (int h = 0; h < 2; ++h)
        if (c[h].shi > c[h].slo || c[h].thi > c[h].tlo) {
          const int o = atomicAdd(outCount, 1);
          if (o < maxOut)
            out[o] = c[h];
        }
    } else if (nd.thi > nd.tlo) {
      const int o = atomicAdd(leafCount, 1);
      if (o < maxLeaf) {
        leafP[o] = nd.p;
        leafD[o] = nd.d;
      }
    }
  }
}

// ---- ball queries ------------------------------------------------------------

__device__ inline void cell_geometry(uint64_t p, int d, const Frame &f,
                                     double overlap, double *cen, double *rc) {
  double s = 0.0;
  cen[0] = cen[1] = cen[2] = 0.0;
  for (int a = 0; a < f.D; ++a) {
    const int ba = axis_bits(d, f.D, a);
    const uint64_t ia = cell_axis_index(p, d, f.D, a);
    const double h = f.ext[a] * ldexp(1.0, -ba);
    cen[a] = f.lo[a] + ((double)ia + 0.5) * h;
    s = s + h * h;
  }
  *rc = (1.0 + overlap) * (0.5 * sqrt(s));
}

__device__ inline int depth_for(const Frame &f, double r) {
  int d = 0;
  for (int dd = 1; dd <= f.DB; ++dd) {
    if (f.hmin[dd] >= r * (1.0 + 1e-9))
      d = dd;
    else
      break;
  }
  return d;
}

__device__ inline double dist3(double ax, double ay, double az, double bx,
                               double by, double bz) {
  return sqrt(sqdiff3(ax, ay, az, bx, by, bz));
}

// First position in [lo, hi) whose code is >= key, by the whole wave: 64
// probes per round narrow the range 64-fold (a few rounds for millions of
// points instead of ~22 dependent loads).
__device__ inline int wave_lower_bound(const uint64_t *code, int lo, int hi,
                                       uint64_t key) {
  const int lane = threadIdx.x & 63;
  while (hi - lo > 64) {
    const int step = (hi - lo + 63) / 64;
    const int pos = lo + lane * step;
    const bool less = pos < hi && code[pos] < key;
    const int cnt = __popcll(__ballot(less));
    if (cnt == 0)
      return lo;
    const int nlo = lo + (cnt - 1) * step + 1;
    const int nhi = min(hi, lo + cnt * step);
    lo = nlo;
    hi = nhi;
  }
  const int pos = lo + lane;
  const bool less = pos < hi && code[pos] < key;
  return lo + __popcll(__ballot(less));
}

// Squared distance from c to the box of the cell (prefix p, depth d), the
// box widened by 1e-12 of the frame on each side (points are assigned to
// cells through rounded quantisation).
__device__ inline double cell_box_dist2(uint64_t p, int d, const Frame &f,
                                        const double *c) {
  double s = 0.0;
  for (int a = 0; a < f.D; ++a) {
    const int ba = axis_bits(d, f.D, a);
    const uint64_t ia = cell_axis_index(p, d, f.D, a);
    const double h = f.ext[a] * ldexp(1.0, -ba);
    const double pad = 1e-12 * (f.ext[a] + fabs(f.lo[a]));
    const double blo = f.lo[a] + (double)ia * h - pad;
    const double bhi = f.lo[a] + (double)(ia + 1) * h + pad;
    const double e = fmax(fmax(blo - c[a], c[a] - bhi), 0.0);
    s += e * e;
  }
  return s;
}

// Squared distance from c to the farthest point of the same widened box.
__device__ inline double cell_box_maxdist2(uint64_t p, int d, const Frame &f,
                                           const double *c) {
  double s = 0.0;
  for (int a = 0; a < f.D; ++a) {
    const int ba = axis_bits(d, f.D, a);
    const uint64_t ia = cell_axis_index(p, d, f.D, a);
    const double h = f.ext[a] * ldexp(1.0, -ba);
    const double pad = 1e-12 * (f.ext[a] + fabs(f.lo[a]));
    const double blo = f.lo[a] + (double)ia * h - pad;
    const double bhi = f.lo[a] + (double)(ia + 1) * h + pad;
    const double e = fmax(fabs(c[a] - blo), fabs(bhi - c[a]));
    s += e * e;
  }
  return s;
}

constexpr int kRefine = 4096;    // longer code ranges are halved first
constexpr int kStack = 128;      // pending ranges per wave
struct ScanNode {
  uint64_t p;
  int d, rs, re;
};

// A chunk of at most 64 consecutive positions of one scanned range.
struct ScanChunk {
  int start, count, inside;
};
constexpr int kChunks = 64;  // chunks queued per wave before a flush
constexpr int kChunkGroup = 4;  // chunks whose loads are issued together

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The scans' default stop test: never stop early.
struct NoStop {
  __device__ constexpr bool operator()() const { return false; }
};

// fn over the queued chunks, in queue order, one 64-lane pass per chunk
// (lane i: position start + i, valid below count); the coordinates of
// kChunkGroup chunks are loaded before the first of them is used, so their
// memory latencies overlap. Ends early once the wave-uniform stop() holds.
template <class Fn, class Stop = NoStop>
__device__ inline void flush_chunks(const SortedSet &S, const ScanChunk *pend,
                                    int np, Fn &&fn, Stop stop = Stop{}) {
  const int lane = threadIdx.x & 63;
  for (int g = 0; g < np && !stop(); g += kChunkGroup) {
    double x[kChunkGroup], y[kChunkGroup], z[kChunkGroup];
    int q[kChunkGroup];
    bool v[kChunkGroup], in[kChunkGroup];
#pragma unroll
    for (int u = 0; u < kChunkGroup; ++u) {
      x[u] = y[u] = z[u] = 0.0;
      q[u] = 0;
      v[u] = in[u] = false;
      if (g + u < np) {
        const ScanChunk ch = pend[g + u];
        q[u] = ch.start + lane;
        v[u] = lane < ch.count;
        in[u] = ch.inside != 0;
        const int qq = v[u] ? q[u] : ch.start;
        x[u] = S.x[qq];
        y[u] = S.y[qq];
        z[u] = S.z[qq];
      }
    }
#pragma unroll
    for (int u = 0; u < kChunkGroup; ++u)
      if (g + u < np)
        fn(q[u], v[u], in[u], x[u], y[u], z[u]);
  }
}

// fn(position, valid, inside, x, y, z) over a superset of the points of S
// within r of c (each candidate once, with its coordinates; ranges in a
// fixed order, so the order is deterministic): the 3^D cells around c at the
// deepest depth whose cells are >= r wide; a range longer than kRefine is
// halved along the next code bit and the halves whose box lies farther than
// r from c are dropped. A range whose (widened) box lies within
// r sqrt(1 - 1e-9) of c holds only points within r whatever the rounding of
// a distance (relative error ~1e-15): such a range of more than 64 points is
// not halved, whole(rs, re) may take it at once (returning true), else it is
// scanned with inside = true (INSIDE = false: no such test, for balls that
// never cover whole ranges). onlyCell >= 0 restricts the scan to that one of
// the 3^D starting cells (a workgroup shares a large ball out that way).
// The ranges to scan are cut into chunks of 64 positions and queued in
// `pend` (kChunks per wave); fn sees them in the order the ranges are
// reached, so the positions of the hits are the same as scanning each range
// when it is reached. A wave-uniform stop() that holds ends the scan early
// (between ranges and between groups of chunks): fn has then seen a prefix
// of the candidates.
template <bool INSIDE = true, class Whole, class Fn, class Stop = NoStop>
__device__ inline void ball_scan(const SortedSet &S, const Frame &f,
                                 const double *c, double r, ScanNode *stk,
                                 ScanChunk *pend, Whole &&whole, Fn &&fn,
                                 int onlyCell = -1, Stop stop = Stop{}) {
  const int lane = threadIdx.x & 63;
  const int d0 = depth_for(f, r);
  const int nr = f.D == 1 ? 3 : f.D == 2 ? 9 : 27;
  const double r2 = r * r * (1.0 + 1e-9) + 1e-300;
  const double rin2 = r * r * (1.0 - 1e-9);
  // the neighbour cells, one lane each
  int rs = 0, re = 0;
  uint64_t pref = 0;
  if (lane < nr) {
    uint64_t j[3] = {0, 0, 0};
    bool ok = true;
    int t = lane;
    for (int a = 0; a < f.D; ++a) {
      const int o = t % 3 - 1;
      t /= 3;
      const int ba = axis_bits(d0, f.D, a);
      const int64_t ia =
          (int64_t)(quantize(c[a], f.lo[a], f.ext[a]) >> (kBits - ba)) + o;
      if (ia < 0 || ia >= ((int64_t)1 << ba))
        ok = false;
      j[a] = (uint64_t)(ia < 0 ? 0 : ia);
    }
    if (ok) {
      pref = cell_prefix(j, f.D, d0);
      const int sh = f.DB - d0;
      rs = code_lower_bound(S, pref << sh);
      re = code_lower_bound(S, (pref + 1) << sh);
    }
  }
  // push the non-empty ones in reverse lane order (popped in lane order)
  const bool keep =
      lane < nr && re > rs && (onlyCell < 0 || lane == onlyCell);
  const uint64_t km = __ballot(keep);
  const int nk = __popcll(km);
  if (keep) {
    const int above = __popcll(km >> lane) - 1;  // kept lanes above this one
    stk[above] = ScanNode{pref, d0, rs, re};
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int top = nk;
  int np = 0;
  while (top > 0 && !stop()) {
    const ScanNode nd = stk[top - 1];
    --top;
    const bool inside = INSIDE && nd.re - nd.rs > 64 &&
                        cell_box_maxdist2(nd.p, nd.d, f, c) <= rin2;
    if (inside && whole(nd.rs, nd.re))
      continue;
    if (!inside && nd.re - nd.rs > kRefine && nd.d < f.DB &&
        top + 2 <= kStack) {
      const uint64_t key = (2 * nd.p + 1) << (f.DB - nd.d - 1);
      const int mid = wave_lower_bound(S.code, nd.rs, nd.re, key);
      const ScanNode ch[2] = {{2 * nd.p, nd.d + 1, nd.rs, mid},
                              {2 * nd.p + 1, nd.d + 1, mid, nd.re}};
      bool use[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
        use[h] = ch[h].re > ch[h].rs &&
                 cell_box_dist2(ch[h].p, ch[h].d, f, c) <= r2;
      // every lane has read the popped entry before lane 0 reuses its slot
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        int t = top;
        if (use[1])  // child 0 on top: popped first
          stk[t++] = ch[1];
        if (use[0])
          stk[t++] = ch[0];
      }
      top += (use[0] ? 1 : 0) + (use[1] ? 1 : 0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      continue;
    }
    for (int b = nd.rs; b < nd.re;) {
      if (np == kChunks) {
        flush_chunks(S, pend, np, fn, stop);
        np = 0;
        wave_sync_lds();  // every lane has read the queue
      }
      const int take = min((nd.re - b + 63) / 64, kChunks - np);
      if (lane < take) {
        const int st = b + 64 * lane;
        pend[np + lane] = ScanChunk{st, min(64, nd.re - st), inside ? 1 : 0};
      }
      np += take;
      b += 64 * take;
    }
    wave_sync_lds();  // queue entries visible; popped entry read by all
  }
  flush_chunks(S, pend, np, fn, stop);
  wave_sync_lds();  // the queue may be reused by the next scan
}

// #{p : dist(c, p) <= r} when that is at most cap, else some count above
// cap (the scan stops once its count passes cap).
__device__ inline int ball_count(const SortedSet &S, const Frame &f,
                                 const double *c, double r, int cap,
                                 ScanNode *stk, ScanChunk *pend) {
  int cnt = 0;  // wave-uniform: sums of range lengths and ballot counts
  ball_scan(
      S, f, c, r, stk, pend,
      [&](int rs, int re) {
        cnt += re - rs;
        return true;
      },
      [&](int, bool valid, bool inside, double x, double y, double z) {
    const bool hit =
        valid && (inside || dist3(c[0], c[1], c[2], x, y, z) <= r);
    cnt += __popcll(__ballot(hit));
  }, -1, [&] { return cnt > cap; });
  return cnt;
}

// #{p : sqdiff(c, p) <= t} (candidates within sqrt(t)) when that is at most
// cap, else some count above cap: the scan stops as soon as its count passes
// cap, so a ball over a dense cluster costs a few chunks, not the cluster.
__device__ inline int count_d2(const SortedSet &S, const Frame &f,
                               const double *c, double t, int cap,
                               ScanNode *stk, ScanChunk *pend) {
  int cnt = 0;  // wave-uniform: sums of range lengths and ballot counts
  ball_scan(
      S, f, c, sqrt(t), stk, pend,
      [&](int rs, int re) {
        cnt += re - rs;
        return true;
      },
      [&](int, bool valid, bool inside, double x, double y, double z) {
    const bool hit =
        valid && (inside || sqdiff3(c[0], c[1], c[2], x, y, z) <= t);
    cnt += __popcll(__ballot(hit));
  }, -1, [&] { return cnt > cap; });
  return cnt;
}

constexpr int kKnnCap = 512;  // candidates a wave selects from in LDS

__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t u = __shfl_xor(v, o);
    v = u < v ? u : v;
  }
  return v;
}

// Distance to the kq-th nearest point of S from c (kq <= S.n): the smallest
// t with #{sqdiff <= t} >= kq, i.e. the kq-th smallest squared distance
// (with multiplicity). The upper end starts at r0^2 and doubles until it
// holds kq points; while more than kKnnCap candidates lie within it, it is
// bisected on the bits of t (non-negative doubles order as their bit
// patterns), each count over the pruned ball of radius sqrt(t). The counts
// are capped at max(kq, kKnnCap): each test below (< kq, >= kq, >
// kKnnCap) reads the same from a capped count as from the exact one. Then the
// candidates within it are gathered into the wave's LDS buffer (`buf`,
// kKnnCap doubles) and the kq-th smallest is taken exactly: kq rounds of
// wave minimum, each removing one instance.
__device__ inline double knn_distance(const SortedSet &S, const Frame &f,
                                      const double *c, int kq, double r0,
                                      ScanNode *stk, ScanChunk *pend,
                                      double *buf) {
  double R = r0 > 0.0 ? r0 : 1e-300;
  const int cap = kq > kKnnCap ? kq : kKnnCap;
  int chi;
  while ((chi = count_d2(S, f, c, R * R, cap, stk, pend)) < kq)
    R *= 2.0;
  uint64_t lo = 0, hi = (uint64_t)__double_as_longlong(R * R);
  while (lo < hi && chi > kKnnCap) {
    const uint64_t mid = lo + (hi - lo) / 2;
    const int cm = count_d2(S, f, c, __longlong_as_double((long long)mid),
                            cap, stk, pend);
    if (cm >= kq) {
      hi = mid;
      chi = cm;
    } else {
      lo = mid + 1;
    }
  }
  if (lo >= hi)
    return sqrt(__longlong_as_double((long long)hi));
  const int lane = threadIdx.x & 63;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const double t = __longlong_as_double((long long)hi);
  int w = 0;
  ball_scan(
      S, f, c, sqrt(t), stk, pend, [](int, int) { return false; },
      [&](int, bool valid, bool inside, double x, double y, double z) {
    const double q = valid ? sqdiff3(c[0], c[1], c[2], x, y, z) : 0.0;
    const bool hit = valid && (inside || q <= t);
    const uint64_t bal = __ballot(hit);
    const int o = w + __popcll(bal & below);
    if (hit && o < kKnnCap)
      buf[o] = q;
    w += __popcll(bal);
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int nv = min(w, kKnnCap);  // == chi
  constexpr int kPer = kKnnCap / 64;
  uint64_t v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int idx = j * 64 + lane;
    v[j] = idx < nv ? (uint64_t)__double_as_longlong(buf[idx]) : ~0ull;
  }
  // the buffer is read before the next cell's gather writes it
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  uint64_t res = hi;
  for (int round = 0; round < kq; ++round) {
    uint64_t lm = v[0];
#pragma unroll
    for (int j = 1; j < kPer; ++j)
      lm = v[j] < lm ? v[j] : lm;
    const uint64_t m = wave_min_u64(lm);
    res = m;
    const uint64_t holders = __ballot(lm == m);
    if (lane == __ffsll((long long)holders) - 1) {
      bool done = false;
#pragma unroll
      for (int j = 0; j < kPer; ++j)
        if (!done && v[j] == m) {
          v[j] = ~0ull;
          done = true;
        }
    }
  }
  return sqrt(__longlong_as_double((long long)res));
}

// One wave per cell: radius, source count; cells within nmax are appended
// to the done list, larger ones are halved (children with targets go to the
// next work list).
struct CellEvalArgs {
  const uint64_t *cellP;
  const int *cellD;
  int count;
  SortedSet S, T;
  Frame f;
  double overlap;
  int kq, nmax;
  uint64_t *nextP;
  int *nextD;
  int maxNext;
  uint64_t *doneKey;
  int *doneD, *doneN;     // doneN: sources within the final radius
  double *doneC, *doneR;  // doneC: 3 per entry
  int maxDone;
  int *counters;          // [0] next, [1] done
};

__global__ __launch_bounds__(kThreads) void cell_eval_kernel(CellEvalArgs e) {
  const int i = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (i >= e.count)
    return;
  const int lane = threadIdx.x & 63;
  __shared__ ScanNode stacks[kThreads / 64][kStack];
  __shared__ ScanChunk pends[kThreads / 64][kChunks];
  __shared__ double knnBuf[kThreads / 64][kKnnCap];
  ScanNode *stk = stacks[threadIdx.x >> 6];
  ScanChunk *pend = pends[threadIdx.x >> 6];
  const uint64_t p = e.cellP[i];
  const int d = e.cellD[i];
  double c[3], rc;
  cell_geometry(p, d, e.f, e.overlap, c, &rc);
  // A cell above nmax is halved (below the deepest level) and its count is
  // not kept, so the counts may stop past max(nmax, kq): each test below
  // reads the same from such a capped count as from the exact one.
  const int cap = d < e.f.DB ? max(e.nmax, e.kq) : INT_MAX;
  int n = ball_count(e.S, e.f, c, rc, cap, stk, pend);
  double r = rc;
  if (n < e.kq) {
    r = knn_distance(e.S, e.f, c, e.kq, rc, stk, pend,
                     knnBuf[threadIdx.x >> 6]);
    if (r < rc)  // max(rc, dk) exactly
      r = rc;
    n = ball_count(e.S, e.f, c, r, cap, stk, pend);
  }
  if (lane != 0)
    return;
  if (n > e.nmax && d < e.f.DB) {
    for (int h = 0; h < 2; ++h) {
      const uint64_t cp = 2 * p + (uint64_t)h;
      const int sh = e.f.DB - d - 1;
      const int a = code_lower_bound(e.T, cp << sh);
      const int b = code_lower_bound(e.T, (cp + 1) << sh);
      if (b > a) {
        const int o = atomicAdd(&e.counters[0], 1);
        if (o < e.maxNext) {
          e.nextP[o] = cp;
          e.nextD[o] = d + 1;
        }
      }
    }
  } else {
    const int o = atomicAdd(&e.counters[1], 1);
    if (o < e.maxDone) {
      e.doneKey[o] = p << (e.f.DB - d);
      e.doneD[o] = d;
      e.doneN[o] = n;
      e.doneC[3 * o] = c[0];
      e.doneC[3 * o + 1] = c[1];
      e.doneC[3 * o + 2] = c[2];
      e.doneR[o] = r;
    }
  }
}

// Final patch arrays in key order; srcCnt[k]: the patch's source count (the
// ball count of its final radius from cell_eval_kernel, the same predicate
// as the membership scan).
__global__ void patch_arrays_kernel(const int *order, int K, const int *doneD,
                                    const int *doneN, const double *doneC,
                                    const double *doneR, double supportScale,
                                    double *cx, double *cy, double *cz,
                                    double *r, double *rinv, double *dinv,
                                    int *depth, int *srcCnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K)
    return;
  const int o = order[k];
  srcCnt[k] = doneN[o];
  cx[k] = doneC[3 * o];
  cy[k] = doneC[3 * o + 1];
  cz[k] = doneC[3 * o + 2];
  const double rr = doneR[o];
  r[k] = rr;
  rinv[k] = 1.0 / rr;
  dinv[k] = 1.0 / (supportScale * rr);
  depth[k] = doneD[o];
}

// Membership: count pass (fill = false) and fill pass. Sources: positions
// within r_k. Targets: positions within r_k with the member flag (within
// r_k - 1e-14) in the top bit, and the patch of every pair.
struct MemberArgs {
  SortedSet P;
  Frame f;
  const double *cx, *cy, *cz, *r;
  int K;
  bool targets;
  double rHeavy;  // patches with a larger radius: member_heavy_kernel
  int *cnt;
  const int64_t *off;
  int *out, *pairPatch;
  int *mismatch;  // fill pass: set when a list's fill differs from its count
};

// Members of patch k found from the starting cell onlyCell (-1: all), the
// fill pass writing from position w; returns the position after the last
// (the count pass: w = 0 and the count). The count pass takes ranges inside
// the ball at once; the fill pass scans them with every point a hit, so both
// passes agree by construction.
template <bool INSIDE>
__device__ inline int64_t member_scan(const MemberArgs &m, int k, bool fill,
                                      int64_t w, ScanNode *stk,
                                      ScanChunk *pend, int onlyCell) {
  const int lane = threadIdx.x & 63;
  const double c[3] = {m.cx[k], m.cy[k], m.cz[k]};
  const double rad = m.r[k], radMember = rad - kEps;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the fill never writes past the patch's list (its count came from
  // another pass; a disagreement is reported, not written out of range)
  const int64_t end = fill ? m.off[k + 1] : 0;
  ball_scan<INSIDE>(
      m.P, m.f, c, rad, stk, pend,
      [&](int rs, int re) {
        if (fill)
          return false;
        w += re - rs;
        return true;
      },
      [&](int p, bool valid, bool inside, double x, double y, double z) {
        double dd = 0.0;
        if (valid)
          dd = dist3(c[0], c[1], c[2], x, y, z);
        const bool hit = valid && (inside || dd <= rad);
        const uint64_t bal = __ballot(hit);
        const int64_t o = w + __popcll(bal & below);
        if (fill && hit && o < end) {
          if (m.targets) {
            const unsigned member = dd <= radMember ? 1u : 0u;
            m.out[o] = (int)((unsigned)p | (member << 31));
            m.pairPatch[o] = k;
          } else {
            m.out[o] = p;
          }
        }
        w += __popcll(bal);
      },
      onlyCell);
  return w;
}

// One wave per patch (the patches above rHeavy are left to
// member_heavy_kernel). INSIDE: the targets (large balls in sparse regions).
template <bool INSIDE>
__global__ __launch_bounds__(kThreads) void member_kernel(MemberArgs m,
                                                          bool fill) {
  const int k = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (k >= m.K || m.r[k] > m.rHeavy)
    return;
  __shared__ ScanNode stacks[kThreads / 64][kStack];
  __shared__ ScanChunk pends[kThreads / 64][kChunks];
  ScanNode *stk = stacks[threadIdx.x >> 6];
  const int64_t w = member_scan<INSIDE>(m, k, fill, fill ? m.off[k] : 0, stk,
                                        pends[threadIdx.x >> 6], -1);
  if (!fill && (threadIdx.x & 63) == 0)
    m.cnt[k] = (int)w;
  if (fill && (threadIdx.x & 63) == 0 && w != m.off[k + 1])
    atomicOr(m.mismatch, 1);
}

// Patches with r > rHeavy, listed by heavy_select_kernel (any order: each
// list entry is handled on its own).
__global__ void heavy_select_kernel(const double *r, int K, double rHeavy,
                                    int *list, int *count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K && r[k] > rHeavy)
    list[atomicAdd(count, 1)] = k;
}

// Large balls (sparse regions: up to ~2e5 targets): one workgroup per
// listed patch, its 3^D starting cells dealt over the waves; the count pass
// keeps a count per starting cell, so the fill pass writes every cell's
// members where the one-wave scan (cells in order) would.
constexpr int kHeavyWaves = 16;
constexpr int kCells = 27;
__global__ __launch_bounds__(64 * kHeavyWaves) void member_heavy_kernel(
    MemberArgs m, const int *list, const int *count, int *cellCnt, bool fill) {
  __shared__ ScanNode stacks[kHeavyWaves][kStack];
  __shared__ ScanChunk pends[kHeavyWaves][kChunks];
  const int wave = threadIdx.x >> 6;
  ScanNode *stk = stacks[wave];
  const int nc = m.f.D == 1 ? 3 : m.f.D == 2 ? 9 : 27;
  const int hc = *count;
  for (int h = blockIdx.x; h < hc; h += gridDim.x) {
    const int k = list[h];
    int *cc = cellCnt + (int64_t)h * kCells;
    for (int j = wave; j < nc; j += kHeavyWaves) {
      int64_t w = 0;
      if (fill) {
        w = m.off[k];
        for (int jj = 0; jj < j; ++jj)
          w += cc[jj];
      }
      const int64_t w0 = w;
      w = member_scan<true>(m, k, fill, w, stk, pends[wave], j);
      if (!fill && (threadIdx.x & 63) == 0)
        cc[j] = (int)w;
      if (fill && (threadIdx.x & 63) == 0 && w != w0 + cc[j])
        atomicOr(m.mismatch, 1);
    }
    __syncthreads();
    if (!fill && threadIdx.x == 0) {
      int64_t t = 0;
      for (int j = 0; j < nc; ++j)
        t += cc[j];
      m.cnt[k] = (int)t;
    }
    __syncthreads();
  }
}

unsigned blocks(int64_t n, int per) {
  return (unsigned)std::max<int64_t>(1, (n + per - 1) / per);
}

// A second stream per device (default priority), kept for the process: the
// workgroup scans of the large target balls run beside the one-wave scans
// of the others. Each build forks and joins it with events of its own.
hipStream_t side_stream() {
  static std::mutex mu;
  static std::map<int, hipStream_t> per;
  int dev = 0;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  hipStream_t &st = per[dev];
  if (!st)
    PACMAN_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return st;
}

struct ForkJoin {
  hipStream_t side;
  hipEvent_t fork = nullptr, join = nullptr;
  ForkJoin() : side(side_stream()) {
    PACMAN_HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    PACMAN_HIP_CHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
  }
  ~ForkJoin() {
    (void)hipEventDestroy(fork);
    (void)hipEventDestroy(join);
  }
  ForkJoin(const ForkJoin &) = delete;
  ForkJoin &operator=(const ForkJoin &) = delete;
};

} // namespace

void build_adaptive(const AdaptiveParams &ap, int dim, const double *dSrc,
                    int64_t N, const double *dTgt, int64_t M,
                    AdaptivePatches &out, hipStream_t s) {
  if (ap.leaf < 1 || ap.nmin < 1 || ap.nmax < ap.nmin)
    throw std::invalid_argument(
        "adaptive patches need leafPoints >= 1 and 1 <= minPoints <= maxPoints");
  if (!(ap.overlap > 0.0 && ap.overlap < 1.0))
    throw std::invalid_argument("relativeOverlap must lie in (0, 1)");
  if (!(ap.supportScale > 0.0))
    throw std::invalid_argument("supportScale must be positive");
  // 1. frame: box of both point sets
  double slo[3], shi[3], tlo[3], thi[3];
  bbox_device(dSrc, N, dim, slo, shi, s);
  bbox_device(dTgt, M, dim, tlo, thi, s);
  Frame f{};
  f.D = dim;
  f.DB = kBits * dim;
  for (int a = 0; a < 3; ++a) {
    f.lo[a] = 0.0;
    f.ext[a] = 1.0;
  }
  for (int a = 0; a < dim; ++a) {
    const double lo = M > 0 ? std::min(slo[a], tlo[a]) : slo[a];
    const double hi = M > 0 ? std::max(shi[a], thi[a]) : shi[a];
    f.lo[a] = lo;
    f.ext[a] = hi - lo;
    if (!(f.ext[a] > 0.0))
      f.ext[a] = 1.0;
    out.lo[a] = lo;
    out.hi[a] = hi;
  }
  for (int d = 0; d <= f.DB; ++d) {
    double h = DBL_MAX;
    for (int a = 0; a < dim; ++a)
      h = std::min(h, f.ext[a] * std::ldexp(1.0, -axis_bits(d, dim, a)));
    f.hmin[d] = h;
  }
  const int bits = f.DB;  // 16, 32 or 48: an even number of 8-bit passes
  // sorted codes and Morton-ordered SoA coordinates of both sets
  auto sort_set = [&](const double *pts, int64_t n, DevBuf<uint64_t> &code,
                      DeviceGrid &g) {
    const int64_t n1 = std::max<int64_t>(n, 1);
    code.alloc(n1);
    DevBuf<uint64_t> k1(n1);
    DevBuf<int> v1(n1);
    g.perm.alloc(n1);
    g.x.alloc(n1);
    g.y.alloc(n1);
    g.z.alloc(n1);
    if (n > 0) {
      morton_codes_kernel<<<blocks(n, kThreads), kThreads, 0, s>>>(
          pts, n, dim, f, code.get(), g.perm.get());
      PACMAN_HIP_CHECK(hipGetLastError());
      radix_sort_pairs(code.get(), g.perm.get(), k1.get(), v1.get(), n, bits,
                       s);
      gather_sorted_kernel<<<blocks(n, kThreads), kThreads, 0, s>>>(
          pts, n, dim, g.perm.get(), g.x.get(), g.y.get(), g.z.get());
      PACMAN_HIP_CHECK(hipGetLastError());
    }
    g.view = GridView{};
    g.view.n = (int)n;
    g.view.x = g.x.get();
    g.view.y = g.y.get();
    g.view.z = g.z.get();
    g.view.perm = g.perm.get();
    g.view.cell_start = nullptr;
    g.ncell = 0;
  };
  DevBuf<uint64_t> scode, tcode;
  sort_set(dSrc, N, scode, out.src);
  sort_set(dTgt, M, tcode, out.tgt);
  const int tabBits = std::min(f.DB, 18);
  const int64_t nTab = ((int64_t)1 << tabBits) + 1;
  DevBuf<int> stab(nTab), ttab(nTab);
  code_table_kernel<<<blocks(nTab, kThreads), kThreads, 0, s>>>(
      scode.get(), (int)N, f.DB - tabBits, tabBits, stab.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  code_table_kernel<<<blocks(nTab, kThreads), kThreads, 0, s>>>(
      tcode.get(), (int)M, f.DB - tabBits, tabBits, ttab.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  const SortedSet S{scode.get(),     out.src.x.get(), out.src.y.get(),
                    out.src.z.get(), (int)N,          stab.get(),
                    f.DB - tabBits,  tabBits};
  const SortedSet T{tcode.get(),     out.tgt.x.get(), out.tgt.y.get(),
                    out.tgt.z.get(), (int)M,          ttab.get(),
                    f.DB - tabBits,  tabBits};

  // 2. the tree, one launch per depth. The nodes of a depth below the root
  // are children of split cells, which are disjoint and hold more than
  // `leaf` sources each: at most 2 N / (leaf + 1) of them.
  const int64_t maxLeaf = std::max<int64_t>(M, 1);
  const int maxNodes =
      (int)std::min<int64_t>(INT_MAX / 2, 2 * (N / (ap.leaf + 1)) + 2);
  DevBuf<uint64_t> leafP(maxLeaf);
  DevBuf<int> leafD(maxLeaf);
  // lv[d]: nodes at depth d (lv[0] = the root), lv[DB + 2]: leaves
  std::vector<int> lvHost(f.DB + 3, 0);
  lvHost[0] = 1;
  DevBuf<int> lv(f.DB + 3), ctr(2);
  DevBuf<TNode> levels[2] = {DevBuf<TNode>(maxNodes), DevBuf<TNode>(maxNodes)};
  const TNode root{0, 0, 0, (int)N, 0, (int)M, 0};
  PACMAN_HIP_CHECK(hipMemcpyAsync(levels[0].get(), &root, sizeof(TNode),
                                  hipMemcpyHostToDevice, s));
  PACMAN_HIP_CHECK(hipMemcpyAsync(lv.get(), lvHost.data(),
                                  sizeof(int) * lvHost.size(),
                                  hipMemcpyHostToDevice, s));
  const unsigned tgrid =
      std::min<unsigned>(blocks(maxNodes, kThreads), 2048u);
  for (int d = 0; d <= f.DB; ++d) {
    tree_level_kernel<<<tgrid, kThreads, 0, s>>>(
        levels[d & 1].get(), lv.get() + d, levels[(d + 1) & 1].get(),
        lv.get() + d + 1, maxNodes, lv.get() + f.DB + 2, leafP.get(),
        leafD.get(), (int)maxLeaf, S, T, ap.leaf, f.DB);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  PACMAN_HIP_CHECK(hipMemcpyAsync(lvHost.data(), lv.get(),
                                  sizeof(int) * lvHost.size(),
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  for (int d = 0; d <= f.DB + 1; ++d)
    if (lvHost[d] > maxNodes)
      throw std::runtime_error("adaptive patches: tree level overflow");
  const int nLeaves = lvHost[f.DB + 2];
  if (nLeaves > maxLeaf)
    throw std::runtime_error("adaptive patches: leaf list overflow");

  // 3-4. radii; halve patches above nmax until none is
  const int kq = (int)std::min<int64_t>(ap.nmin, N);
  const int64_t maxDone = std::max<int64_t>(M, 1);
  DevBuf<uint64_t> doneKey(maxDone);
  DevBuf<int> doneD(maxDone), doneN(maxDone);
  DevBuf<double> doneC(3 * maxDone), doneR(maxDone);
  DevBuf<uint64_t> curP = std::move(leafP);
  DevBuf<int> curD = std::move(leafD);
  int nCur = nLeaves, nDone = 0;
  while (nCur > 0) {
    const int maxNext = 2 * nCur;
    DevBuf<uint64_t> nextP(maxNext);
    DevBuf<int> nextD(maxNext);
    int h[2] = {0, nDone};
    PACMAN_HIP_CHECK(
        hipMemcpyAsync(ctr.get(), h, 8, hipMemcpyHostToDevice, s));
    CellEvalArgs e{};
    e.cellP = curP.get();
    e.cellD = curD.get();
    e.count = nCur;
    e.S = S;
    e.T = T;
    e.f = f;
    e.overlap = ap.overlap;
    e.kq = kq;
    e.nmax = ap.nmax;
    e.nextP = nextP.get();
    e.nextD = nextD.get();
    e.maxNext = maxNext;
    e.doneKey = doneKey.get();
    e.doneD = doneD.get();
    e.doneN = doneN.get();
    e.doneC = doneC.get();
    e.doneR = doneR.get();
    e.maxDone = (int)maxDone;
    e.counters = ctr.get();
    cell_eval_kernel<<<blocks(nCur, kThreads / 64), kThreads, 0, s>>>(e);
    PACMAN_HIP_CHECK(hipGetLastError());
    PACMAN_HIP_CHECK(
        hipMemcpyAsync(h, ctr.get(), 8, hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    if (h[1] > maxDone)
      throw std::runtime_error("adaptive patches: patch list overflow");
    nCur = h[0];
    nDone = h[1];
    curP = std::move(nextP);
    curD = std::move(nextD);
  }

  // 5. patches in key order
  const int K = nDone;
  out.K = K;
  const int64_t K1 = std::max(K, 1);
  DevBuf<int> order(K1), ordTmp(K1);
  DevBuf<uint64_t> keyTmp(K1);
  {
    std::vector<int> iota(K1);
    for (int k = 0; k < K; ++k)
      iota[k] = k;
    PACMAN_HIP_CHECK(hipMemcpyAsync(order.get(), iota.data(), 4 * K1,
                                    hipMemcpyHostToDevice, s));
    radix_sort_pairs(doneKey.get(), order.get(), keyTmp.get(), ordTmp.get(),
                     K, bits, s);
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));  // iota copied
  }
  out.cx.alloc(K1);
  out.cy.alloc(K1);
  out.cz.alloc(K1);
  out.r.alloc(K1);
  out.rinv.alloc(K1);
  out.dinv.alloc(K1);
  out.key.alloc(K1);
  out.depth.alloc(K1);
  DevBuf<int> cnt(K1), mismatch(1);
  PACMAN_HIP_CHECK(hipMemsetAsync(mismatch.get(), 0, sizeof(int), s));
  // doneKey is now sorted; the other entries are in append order, reached
  // through `order`
  if (K > 0) {
    patch_arrays_kernel<<<blocks(K, kThreads), kThreads, 0, s>>>(
        order.get(), K, doneD.get(), doneN.get(), doneC.get(), doneR.get(),
        ap.supportScale, out.cx.get(), out.cy.get(), out.cz.get(),
        out.r.get(), out.rinv.get(), out.dinv.get(), out.depth.get(),
        cnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
    PACMAN_HIP_CHECK(hipMemcpyAsync(out.key.get(), doneKey.get(), 8 * K,
                                    hipMemcpyDeviceToDevice, s));
  }

  // membership
  MemberArgs m{};
  m.f = f;
  m.cx = out.cx.get();
  m.cy = out.cy.get();
  m.cz = out.cz.get();
  m.r = out.r.get();
  m.K = K;
  m.mismatch = mismatch.get();
  const unsigned wb = blocks(K, kThreads / 64);
  // sources (at most nmax per patch: never a large ball); their counts come
  // from the radius pass (patch_arrays_kernel)
  m.P = S;
  m.targets = false;
  m.rHeavy = DBL_MAX;
  m.cnt = cnt.get();
  out.srcOff.alloc(K + 1);
  const int64_t tot = exclusive_scan_i32_to_i64(cnt.get(), out.srcOff.get(), K, s);
  if (tot > INT_MAX)
    throw std::runtime_error("patch source lists exceed 2^31 entries");
  out.srcIdx.alloc(std::max<int64_t>(tot, 1));
  m.off = out.srcOff.get();
  m.out = out.srcIdx.get();
  if (K > 0) {
    member_kernel<false><<<wb, kThreads, 0, s>>>(m, true);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  // targets; balls expected to hold more than kHeavyPoints targets (from
  // the mean target density of the frame) go to member_heavy_kernel
  m.P = T;
  m.targets = true;
  {
    constexpr double kHeavyPoints = 16384.0;
    double vol = 1.0;
    for (int a = 0; a < dim; ++a)
      vol *= f.ext[a];
    const double cD = dim == 1 ? 2.0 : dim == 2 ? M_PI : 4.0 * M_PI / 3.0;
    m.rHeavy = M > 0 ? std::pow(kHeavyPoints * vol / (cD * (double)M),
                                1.0 / dim)
                     : DBL_MAX;
  }
  DevBuf<int> hList(K1), hCount(1), cellCnt((size_t)K1 * kCells);
  ForkJoin fj;
  const unsigned hb = (unsigned)std::min<int64_t>(K1, 1024);
  PairSet &ps = out.pairs;
  ps.off.alloc(K + 1);
  if (K > 0) {
    PACMAN_HIP_CHECK(hipMemsetAsync(hCount.get(), 0, sizeof(int), s));
    heavy_select_kernel<<<blocks(K, kThreads), kThreads, 0, s>>>(
        out.r.get(), K, m.rHeavy, hList.get(), hCount.get());
    PACMAN_HIP_CHECK(hipGetLastError());
    PACMAN_HIP_CHECK(hipEventRecord(fj.fork, s));
    PACMAN_HIP_CHECK(hipStreamWaitEvent(fj.side, fj.fork, 0));
    member_heavy_kernel<<<hb, 64 * kHeavyWaves, 0, fj.side>>>(
        m, hList.get(), hCount.get(), cellCnt.get(), false);
    PACMAN_HIP_CHECK(hipGetLastError());
    member_kernel<true><<<wb, kThreads, 0, s>>>(m, false);
    PACMAN_HIP_CHECK(hipGetLastError());
    PACMAN_HIP_CHECK(hipEventRecord(fj.join, fj.side));
    PACMAN_HIP_CHECK(hipStreamWaitEvent(s, fj.join, 0));
  }
  const int64_t P = exclusive_scan_i32_to_i64(cnt.get(), ps.off.get(), K, s);
  if (P > INT_MAX)
    throw std::runtime_error("patch/target pairs exceed 2^31 entries");
  ps.P = P;
  ps.tgt.alloc(std::max<int64_t>(P, 1));
  DevBuf<int> pairPatch(std::max<int64_t>(P, 1));
  m.off = ps.off.get();
  m.out = ps.tgt.get();
  m.pairPatch = pairPatch.get();
  if (K > 0) {
    PACMAN_HIP_CHECK(hipEventRecord(fj.fork, s));
    PACMAN_HIP_CHECK(hipStreamWaitEvent(fj.side, fj.fork, 0));
    member_heavy_kernel<<<hb, 64 * kHeavyWaves, 0, fj.side>>>(
        m, hList.get(), hCount.get(), cellCnt.get(), true);
    PACMAN_HIP_CHECK(hipGetLastError());
    member_kernel<true><<<wb, kThreads, 0, s>>>(m, true);
    PACMAN_HIP_CHECK(hipGetLastError());
    PACMAN_HIP_CHECK(hipEventRecord(fj.join, fj.side));
    PACMAN_HIP_CHECK(hipStreamWaitEvent(s, fj.join, 0));
  }
  build_target_slots(ps, pairPatch.get(), M, s);
  int bad = 0;
  PACMAN_HIP_CHECK(hipMemcpyAsync(&bad, mismatch.get(), sizeof(int),
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));  // temporaries in use above
  if (bad)
    throw std::runtime_error(
        "adaptive patches: a membership list's fill differs from its count");
}

} // namespace pacman_amd
