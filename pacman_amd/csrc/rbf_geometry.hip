// Geometry kernels of the RBF-PUM path: bounding boxes, cell-sorted grids,
// the patch-radius probes, the patch-centre lattice (tag, snap, prune,
// re-tag, compact), patch membership and the per-target transpose.
//
// Compiled with -ffp-contract=off: every squared distance, lattice
// coordinate and comparison is evaluated exactly as written, so membership
// decisions are bit-identical to the CPU restatement of the reference
// (src/rbf_pum/clustering/*.hpp, src/rbf_pum/solver/solver.hpp:44-67,
// 128-169).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

constexpr int kBlock = 256;

__global__ void fill_i32_kernel(int *p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

void fill_i32(int *p, int64_t n, int v, hipStream_t s) {
  if (n <= 0)
    return;
  fill_i32_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, n,
                                                                           v);
  PACMAN_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Bounding box (per-block partial min/max, finished on the host).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(
    const double *pts, int64_t n, int dim, double *part) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    for (int a = 0; a < dim; ++a) {
      const double v = pts[i * dim + a];
      mn[a] = fmin(mn[a], v);
      mx[a] = fmax(mx[a], v);
    }
  }
  __shared__ double smn[3][kBlock / kWave], smx[3][kBlock / kWave];
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int a = 0; a < 3; ++a) {
    const double lo = wave_min(mn[a]), hi = wave_max(mx[a]);
    if (lane == 0) {
      smn[a][w] = lo;
      smx[a][w] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      double lo = smn[a][0], hi = smx[a][0];
      for (int k = 1; k < kBlock / kWave; ++k) {
        lo = fmin(lo, smn[a][k]);
        hi = fmax(hi, smx[a][k]);
      }
      part[blockIdx.x * 6 + a] = lo;
      part[blockIdx.x * 6 + 3 + a] = hi;
    }
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan (block-local scan, recursive scan of block sums, add).
// ---------------------------------------------------------------------------
constexpr int kScanPer = 4;
constexpr int kScanTile = kBlock * kScanPer;

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kBlock) void scan_local_kernel(const Tin *in,
                                                            Tout *out,
                                                            int64_t n,
                                                            Tout *blockSums) {
  const int64_t base =
      (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  Tout v[kScanPer];
  Tout s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t idx = base + k;
    const Tout x = idx < n ? (Tout)in[idx] : (Tout)0;
    v[k] = s;
    s += x;
  }
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  Tout incl = s;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const Tout t = __shfl_up(incl, o, kWave);
    if (lane >= o)
      incl += t;
  }
  __shared__ Tout wsum[kBlock / kWave];
  if (lane == kWave - 1)
    wsum[w] = incl;
  __syncthreads();
  Tout woff = 0;
  for (int k = 0; k < w; ++k)
    woff += wsum[k];
  const Tout excl = woff + incl - s;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t idx = base + k;
    if (idx < n)
      out[idx] = excl + v[k];
  }
  if (threadIdx.x == kBlock - 1)
    blockSums[blockIdx.x] = excl + s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scan_add_kernel(T *out, int64_t n,
                                                          const T *offs) {
  const T o = offs[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int k = threadIdx.x; k < kScanTile; k += kBlock) {
    const int64_t idx = base + k;
    if (idx < n)
      out[idx] += o;
  }
}

template <typename T>
__global__ void set_last_kernel(T *out, int64_t n, const T *src) {
  out[n] = *src;
}

// Exclusive scan of at most kSmallScan entries in one launch: one workgroup,
// thread t scanning entries [kSmallPer t, kSmallPer (t + 1)) in registers,
// then the thread sums across the workgroup; out[n] = total.
constexpr int kSmallThreads = 1024;
constexpr int kSmallPer = 16;
constexpr int64_t kSmallScan = (int64_t)kSmallThreads * kSmallPer;

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kSmallThreads) void scan_small_kernel(
    const Tin *in, Tout *out, int64_t n) {
  const int t = threadIdx.x;
  const int64_t base = (int64_t)t * kSmallPer;
  Tout v[kSmallPer];
  Tout sum = 0;
#pragma unroll
  for (int k = 0; k < kSmallPer; ++k) {
    const int64_t idx = base + k;
    const Tout x = idx < n ? (Tout)in[idx] : (Tout)0;
    v[k] = sum;
    sum += x;
  }
  const int lane = t % kWave, w = t / kWave;
  Tout incl = sum;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const Tout u = __shfl_up(incl, o, kWave);
    if (lane >= o)
      incl += u;
  }
  __shared__ Tout wsum[kSmallThreads / kWave];
  if (lane == kWave - 1)
    wsum[w] = incl;
  __syncthreads();
  Tout woff = 0;
  for (int k = 0; k < w; ++k)
    woff += wsum[k];
  const Tout excl = woff + incl - sum;
#pragma unroll
  for (int k = 0; k < kSmallPer; ++k) {
    const int64_t idx = base + k;
    if (idx < n)
      out[idx] = excl + v[k];
  }
  if (t == kSmallThreads - 1)
    out[n] = excl + sum;
}

// out has n + 1 entries; out[n] = total. Returns the total (synchronises).
template <typename Tin, typename Tout>
void scan_device(const Tin *in, Tout *out, int64_t n, hipStream_t s,
                 bool fetchTotal, Tout *totalHost) {
  if (n <= kSmallScan) {
    scan_small_kernel<Tin, Tout><<<1, kSmallThreads, 0, s>>>(in, out,
                                                             std::max<int64_t>(n, 0));
    PACMAN_HIP_CHECK(hipGetLastError());
    if (fetchTotal) {
      PACMAN_HIP_CHECK(hipMemcpyAsync(totalHost, out + std::max<int64_t>(n, 0),
                                      sizeof(Tout), hipMemcpyDeviceToHost, s));
      PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    }
    return;
  }
  const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  DevBuf<Tout> sums(nb + 1);
  scan_local_kernel<Tin, Tout><<<(unsigned)nb, kBlock, 0, s>>>(in, out, n,
                                                               sums.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  if (nb == 1) {
    set_last_kernel<Tout><<<1, 1, 0, s>>>(out, n, sums.get());
  } else {
    DevBuf<Tout> offs(nb + 1);
    scan_device<Tout, Tout>(sums.get(), offs.get(), nb, s, false, nullptr);
    scan_add_kernel<Tout><<<(unsigned)nb, kBlock, 0, s>>>(out, n, offs.get());
    set_last_kernel<Tout><<<1, 1, 0, s>>>(out, n, offs.get() + nb);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  PACMAN_HIP_CHECK(hipGetLastError());
  // Temporaries (offs, sums) may be released while the kernels still run:
  // the device cache reuses a released block only after a device
  // synchronisation, and without the cache hipFree synchronises.
  if (fetchTotal) {
    PACMAN_HIP_CHECK(hipMemcpyAsync(totalHost, out + n, sizeof(Tout),
                                    hipMemcpyDeviceToHost, s));
    PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  }
}

// ---------------------------------------------------------------------------
// Cell-sorted grid
// ---------------------------------------------------------------------------
__device__ inline int cell_of(const GridView &g, double x, double y,
                              double z) {
  return (grid_axis_cell(g, 0, x) * g.dims[1] + grid_axis_cell(g, 1, y)) *
             g.dims[2] +
         grid_axis_cell(g, 2, z);
}

__device__ inline double coord(const double *pts, int64_t i, int dim, int a) {
  return a < dim ? pts[i * dim + a] : 0.0;
}

__global__ void grid_count_kernel(GridView g, const double *pts, int64_t n,
                                  int dim, int *key, int *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = cell_of(g, coord(pts, i, dim, 0), coord(pts, i, dim, 1),
                        coord(pts, i, dim, 2));
  key[i] = c;
  atomicAdd(&cnt[c], 1);
}

__global__ void grid_scatter_kernel(const int *key, int64_t n,
                                    const int *start, int *fill, int *perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = key[i];
  perm[start[c] + atomicAdd(&fill[c], 1)] = (int)i;
}

// Restores ascending original index inside each cell (deterministic order).
__global__ void grid_sort_cells_kernel(const int *start, int64_t ncell,
                                       int *perm) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell)
    return;
  const int s = start[c], e = start[c + 1];
  for (int i = s + 1; i < e; ++i) {
    const int v = perm[i];
    int j = i - 1;
    while (j >= s && perm[j] > v) {
      perm[j + 1] = perm[j];
      --j;
    }
    perm[j + 1] = v;
  }
}

__global__ void grid_gather_kernel(const double *pts, int64_t n, int dim,
                                   const int *perm, double *x, double *y,
                                   double *z) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const int64_t i = perm[p];
  x[p] = coord(pts, i, dim, 0);
  y[p] = coord(pts, i, dim, 1);
  z[p] = coord(pts, i, dim, 2);
}

// ---------------------------------------------------------------------------
// Wave-level scan over the column runs of a query box.
// ---------------------------------------------------------------------------
template <class F>
__device__ inline void wave_box_scan(const GridView &g, const BoxRange &b,
                                     F &&f) {
  // Consecutive column runs share one pass of the wave while they fit in its
  // 64 lanes (lanes in column order, points in order within a column), so a
  // box of short runs takes a few passes instead of one per column; a run
  // longer than the wave takes passes of its own, as before.
  const int lane = threadIdx.x % kWave;
  const int ny = b.y1 - b.y0 + 1;
  const int ncol = (b.x1 - b.x0 + 1) * ny;
  int col = 0, cs = 0, ce = 0;  // run of column `col` (uniform)
  auto load_run = [&](int c) {
    const int ix = b.x0 + c / ny, iy = b.y0 + c % ny;
    const int base = (ix * g.dims[1] + iy) * g.dims[2];
    cs = g.cell_start[base + b.z0];
    ce = g.cell_start[base + b.z1 + 1];
  };
  if (ncol > 0)
    load_run(0);
  while (col < ncol) {
    if (ce - cs > kWave) {  // a long run: whole-wave passes, then the rest
      for (; ce - cs > kWave; cs += kWave)
        f(cs + lane, true);
      continue;
    }
    // pack this run and the following ones while they fit
    int used = 0, p = -1;
    while (col < ncol && used + (ce - cs) <= kWave) {
      const int len = ce - cs;
      if (lane >= used && lane < used + len)
        p = cs + (lane - used);
      used += len;
      if (++col < ncol)
        load_run(col);
    }
    f(p < 0 ? 0 : p, p >= 0);
  }
}

// Lexicographic (d2, original index) minimum across the wave.
__device__ inline void wave_argmin(double &d2, int &orig, int &pos) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double od = __shfl_xor(d2, o, kWave);
    const int oo = __shfl_xor(orig, o, kWave);
    const int op = __shfl_xor(pos, o, kWave);
    if (od < d2 || (od == d2 && oo < orig)) {
      d2 = od;
      orig = oo;
      pos = op;
    }
  }
}

// ---------------------------------------------------------------------------
// FindRadius probes (src/rbf_pum/clustering/clusters_radius.hpp:22-125)
// ---------------------------------------------------------------------------
constexpr int kKnnCap = 8192;

__device__ inline void block_box_cells(const BoxRange &b, int64_t c, int &ix,
                                       int &iy, int &iz) {
  const int ny = b.y1 - b.y0 + 1, nz = b.z1 - b.z0 + 1;
  iz = b.z0 + (int)(c % nz);
  iy = b.y0 + (int)((c / nz) % ny);
  ix = b.x0 + (int)(c / ((int64_t)nz * ny));
}

// Nearest source to each probe point; ties go to the lowest original index.
__global__ __launch_bounds__(kBlock) void probe_nearest_kernel(
    GridView g, const double *probes, double rad0, double radMax, int *outPos) {
  const double px = probes[3 * blockIdx.x], py = probes[3 * blockIdx.x + 1],
               pz = probes[3 * blockIdx.x + 2];
  __shared__ double sd[kBlock / kWave];
  __shared__ int so[kBlock / kWave], sp[kBlock / kWave];
  __shared__ int done;
  for (double rad = rad0;; rad *= 2.0) {
    const BoxRange b = grid_box(g, px, py, pz, rad);
    const int64_t nc = (int64_t)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) *
                       (b.z1 - b.z0 + 1);
    const double r2 = rad * rad;
    const bool all = rad > radMax;
    double best = DBL_MAX;
    int bo = INT_MAX, bp = -1;
    for (int64_t c = threadIdx.x; c < nc; c += kBlock) {
      int ix, iy, iz;
      block_box_cells(b, c, ix, iy, iz);
      const int cell = (ix * g.dims[1] + iy) * g.dims[2] + iz;
      for (int p = g.cell_start[cell]; p < g.cell_start[cell + 1]; ++p) {
        const double d2 = sqdiff3(px, py, pz, g.x[p], g.y[p], g.z[p]);
        if (!(d2 <= r2) && !all)
          continue;
        const int o = g.perm[p];
        if (d2 < best || (d2 == best && o < bo)) {
          best = d2;
          bo = o;
          bp = p;
        }
      }
    }
    wave_argmin(best, bo, bp);
    const int w = threadIdx.x / kWave;
    if (threadIdx.x % kWave == 0) {
      sd[w] = best;
      so[w] = bo;
      sp[w] = bp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bd = sd[0];
      int oo = so[0], pp = sp[0];
      for (int k = 1; k < kBlock / kWave; ++k)
        if (sd[k] < bd || (sd[k] == bd && so[k] < oo)) {
          bd = sd[k];
          oo = so[k];
          pp = sp[k];
        }
      done = (pp >= 0) ? 1 : 0;
      if (pp >= 0)
        outPos[blockIdx.x] = pp;
    }
    __syncthreads();
    if (done)
      return;
    __syncthreads();
  }
}

__device__ int64_t block_count_within(const GridView &g, const BoxRange &b,
                                      double qx, double qy, double qz,
                                      double r2, int64_t *red) {
  const int64_t nc = (int64_t)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) *
                     (b.z1 - b.z0 + 1);
  int cnt = 0;
  for (int64_t c = threadIdx.x; c < nc; c += kBlock) {
    int ix, iy, iz;
    block_box_cells(b, c, ix, iy, iz);
    const int cell = (ix * g.dims[1] + iy) * g.dims[2] + iz;
    for (int p = g.cell_start[cell]; p < g.cell_start[cell + 1]; ++p)
      cnt += sqdiff3(qx, qy, qz, g.x[p], g.y[p], g.z[p]) <= r2 ? 1 : 0;
  }
  cnt = wave_sum_int(cnt);
  __syncthreads();
  if (threadIdx.x % kWave == 0)
    red[threadIdx.x / kWave] = cnt;
  __syncthreads();
  int64_t tot = 0;
  for (int k = 0; k < kBlock / kWave; ++k)
    tot += red[k];
  __syncthreads();
  return tot;
}

// Largest squared distance among the k nearest sources of each projected
// probe (DistanceToKNearestCallback, callbacks.hpp:21-34, then the max at
// clusters_radius.hpp:107-119). That value is the k-th smallest squared
// distance, which ties cannot change.
__global__ __launch_bounds__(kBlock) void probe_knn_kernel(
    GridView g, const int *probePos, int k, double rad0, double radMax,
    double *outMax) {
  const int pos = probePos[blockIdx.x];
  const double qx = g.x[pos], qy = g.y[pos], qz = g.z[pos];
  __shared__ int64_t red[kBlock / kWave];
  __shared__ double vals[kKnnCap];
  __shared__ int nvals;
  double rad = rad0, lo = 0.0;
  int64_t cnt = 0;
  for (;;) {
    const BoxRange b = grid_box(g, qx, qy, qz, rad);
    cnt = block_count_within(g, b, qx, qy, qz, rad * rad, red);
    if (cnt >= k || rad > radMax)
      break;
    lo = rad;
    rad *= 1.5;
  }
  // Shrink the ball until the candidates fit in LDS.
  for (int it = 0; it < 200 && cnt > kKnnCap; ++it) {
    const double mid = 0.5 * (lo + rad);
    const BoxRange b = grid_box(g, qx, qy, qz, mid);
    const int64_t c = block_count_within(g, b, qx, qy, qz, mid * mid, red);
    if (c >= k) {
      rad = mid;
      cnt = c;
    } else {
      lo = mid;
    }
  }
  if (cnt > kKnnCap) {
    if (threadIdx.x == 0)
      outMax[blockIdx.x] = __builtin_nan("");
    return;
  }
  if (threadIdx.x == 0)
    nvals = 0;
  __syncthreads();
  const BoxRange b = grid_box(g, qx, qy, qz, rad);
  const double r2 = rad * rad;
  const int64_t nc = (int64_t)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) *
                     (b.z1 - b.z0 + 1);
  for (int64_t c = threadIdx.x; c < nc; c += kBlock) {
    int ix, iy, iz;
    block_box_cells(b, c, ix, iy, iz);
    const int cell = (ix * g.dims[1] + iy) * g.dims[2] + iz;
    for (int p = g.cell_start[cell]; p < g.cell_start[cell + 1]; ++p) {
      const double d2 = sqdiff3(qx, qy, qz, g.x[p], g.y[p], g.z[p]);
      if (d2 <= r2) {
        const int slot = atomicAdd(&nvals, 1);
        if (slot < kKnnCap)
          vals[slot] = d2;
      }
    }
  }
  __syncthreads();
  const int nv = min(nvals, kKnnCap);
  const int kk = (int)min<int64_t>(k, nv);
  for (int e = threadIdx.x; e < nv; e += kBlock) {
    const double v = vals[e];
    int less = 0, leq = 0;
    for (int f = 0; f < nv; ++f) {
      less += vals[f] < v ? 1 : 0;
      leq += vals[f] <= v ? 1 : 0;
    }
    if (less < kk && kk <= leq)
      outMax[blockIdx.x] = v;  // every writer holds the same value
  }
}

// ---------------------------------------------------------------------------
// Patch-centre lattice (clustering.hpp:127-271)
// ---------------------------------------------------------------------------
struct LatticeDev {
  int dim;
  int64_t shape[3];
  double lower[3];
  double dist[3];
  int64_t nb;
  double r, r2, thr;
  int nOff;
  int64_t off[13];
};

__device__ inline void lattice_position(const LatticeDev &L, int64_t idx,
                                        double *c) {
  int64_t rem = idx;
  c[0] = c[1] = c[2] = 0.0;
  for (int a = L.dim - 1; a >= 0; --a) {
    const int64_t co = rem % L.shape[a];
    rem /= L.shape[a];
    // clustering.hpp:327-330: Point{lower}, then += distances * coords
    c[a] = L.lower[a] + L.dist[a] * (double)co;
  }
}

// One wave per candidate: steps 6-8 (tag empty of sources, tag empty of
// targets, snap to the nearest source).
__global__ __launch_bounds__(kBlock) void lattice_tag_snap_kernel(
    LatticeDev L, GridView src, GridView tgt, double *cx, double *cy,
    double *cz, int *flag) {
  const int64_t idx =
      (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (idx >= L.nb)
    return;
  double c[3];
  lattice_position(L, idx, c);
  // nearest source within r (SquaredDifference(center, value) <= r*r)
  double best = DBL_MAX;
  int bo = INT_MAX, bp = -1;
  const BoxRange bs = grid_box(src, c[0], c[1], c[2], L.r);
  wave_box_scan(src, bs, [&](int p, bool valid) {
    if (!valid)
      return;
    const double d2 = sqdiff3(c[0], c[1], c[2], src.x[p], src.y[p], src.z[p]);
    if (d2 <= L.r2) {
      const int o = src.perm[p];
      if (d2 < best || (d2 == best && o < bo)) {
        best = d2;
        bo = o;
        bp = p;
      }
    }
  });
  wave_argmin(best, bo, bp);
  const bool lane0 = (threadIdx.x % kWave) == 0;
  if (bp < 0) {
    if (lane0)
      flag[idx] = 0;
    return;
  }
  // any target within r
  bool found = false;
  const BoxRange bt = grid_box(tgt, c[0], c[1], c[2], L.r);
  wave_box_scan(tgt, bt, [&](int p, bool valid) {
    if (found)
      return;  // uniform: found is set from a ballot below
    const bool hit =
        valid &&
        sqdiff3(c[0], c[1], c[2], tgt.x[p], tgt.y[p], tgt.z[p]) <= L.r2;
    if (__ballot(hit) != 0)
      found = true;
  });
  if (lane0) {
    flag[idx] = found ? 1 : 0;
    cx[idx] = src.x[bp];
    cy[idx] = src.y[bp];
    cz[idx] = src.z[bp];
  }
}

// Step 9: a live centre dies iff a live later lattice neighbour is closer
// than 0.4 * min spacing. This equals the reference's sequential host loop
// (clustering.hpp:230-254): the neighbour set is symmetric, and a live
// earlier neighbour that is close would already have died when it saw this
// one.
__global__ void lattice_prune_kernel(LatticeDev L, const double *cx,
                                     const double *cy, const double *cz,
                                     const int *flag, int *flag2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.nb)
    return;
  if (!flag[i]) {
    flag2[i] = 0;
    return;
  }
  int alive = 1;
  for (int k = 0; k < L.nOff; ++k) {
    const int64_t j = i + L.off[k];
    if (j < L.nb && flag[j] &&
        sqdiff3(cx[i], cy[i], cz[i], cx[j], cy[j], cz[j]) < L.thr) {
      alive = 0;
      break;
    }
  }
  flag2[i] = alive;
}

// Step 10: re-tag against targets from the snapped position.
__global__ __launch_bounds__(kBlock) void lattice_retag_kernel(
    LatticeDev L, GridView tgt, const double *cx, const double *cy,
    const double *cz, const int *flag2, int *flag3) {
  const int64_t idx =
      (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (idx >= L.nb)
    return;
  const bool lane0 = (threadIdx.x % kWave) == 0;
  if (!flag2[idx]) {
    if (lane0)
      flag3[idx] = 0;
    return;
  }
  const double x = cx[idx], y = cy[idx], z = cz[idx];
  bool found = false;
  const BoxRange bt = grid_box(tgt, x, y, z, L.r);
  wave_box_scan(tgt, bt, [&](int p, bool valid) {
    if (found)
      return;
    const bool hit =
        valid && sqdiff3(x, y, z, tgt.x[p], tgt.y[p], tgt.z[p]) <= L.r2;
    if (__ballot(hit) != 0)
      found = true;
  });
  if (lane0)
    flag3[idx] = found ? 1 : 0;
}

// Step 11: stable compaction (remove_if keeps candidate order).
__global__ void lattice_compact_kernel(int64_t nb, const int *flag3,
                                       const int *pos, const double *cx,
                                       const double *cy, const double *cz,
                                       double *ox, double *oy, double *oz,
                                       int64_t *ocand) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb || !flag3[i])
    return;
  const int k = pos[i];
  ox[k] = cx[i];
  oy[k] = cy[i];
  oz[k] = cz[i];
  ocand[k] = i;
}

// ---------------------------------------------------------------------------
// Patch membership (solver.hpp:44-67): ArborX intersects(Sphere(c, r)) keeps
// points whose distance to c is <= r (targets use r - 1e-14).
// ---------------------------------------------------------------------------
__device__ inline double dist3(double ax, double ay, double az, double bx,
                               double by, double bz) {
  return sqrt(sqdiff3(ax, ay, az, bx, by, bz));
}

// Membership in one geometric pass: the members of patch k are written to
// its fixed-capacity row scratch[k * cap ...] (first cap members, in scan
// order) and counted; member_compact_kernel then moves them to their CSR
// place. `value(p, d)` gives the stored entry of point p at distance d.
template <class V>
__device__ inline void capped_fill(const GridView &g, double x, double y,
                                   double z, double rad, int cap, int *row,
                                   int *count, V &&value) {
  int w = 0;
  wave_box_scan(g, grid_box(g, x, y, z, rad), [&](int p, bool valid) {
    double d = 0.0;
    if (valid)
      d = dist3(x, y, z, g.x[p], g.y[p], g.z[p]);
    const bool hit = valid && d <= rad;
    const uint64_t m = __ballot(hit);
    const int o = w + lane_prefix(m);
    if (hit && o < cap)
      row[o] = value(p, d);
    w += __popcll(m);
  });
  if (threadIdx.x % kWave == 0)
    *count = w;
}

__global__ __launch_bounds__(kBlock) void patch_src_capped_kernel(
    GridView g, const double *cx, const double *cy, const double *cz,
    int64_t K, double rad, int cap, int *scratch, int *cnt) {
  const int64_t k =
      (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (k >= K)
    return;
  capped_fill(g, cx[k], cy[k], cz[k], rad, cap, scratch + k * cap, cnt + k,
              [](int p, double) { return p; });
}

__global__ __launch_bounds__(kBlock) void patch_tgt_capped_kernel(
    GridView g, const double *cx, const double *cy, const double *cz,
    int64_t K, double rad, double radMember, int cap, int *scratch, int *cnt) {
  const int64_t k =
      (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (k >= K)
    return;
  capped_fill(g, cx[k], cy[k], cz[k], rad, cap, scratch + k * cap, cnt + k,
              [&](int p, double d) {
                const unsigned member = d <= radMember ? 1u : 0u;
                return (int)((unsigned)p | (member << 31));
              });
}

// CSR from the capped rows: a wave per patch copies its row; a patch with
// more members than the row holds is listed again from the geometry (the
// same scan order, so the same list). pairPatch (optional) gets k.
__global__ __launch_bounds__(kBlock) void member_compact_kernel(
    GridView g, const double *cx, const double *cy, const double *cz,
    int64_t K, double rad, double radMember, bool targets, int cap,
    const int *scratch, const int *cnt, const int64_t *off, int *out,
    int *pairPatch) {
  const int64_t k =
      (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (k >= K)
    return;
  const int lane = threadIdx.x % kWave;
  const int c = cnt[k];
  const int64_t w0 = off[k];
  if (c <= cap) {
    for (int i = lane; i < c; i += kWave) {
      out[w0 + i] = scratch[k * cap + i];
      if (pairPatch)
        pairPatch[w0 + i] = (int)k;
    }
    return;
  }
  const double x = cx[k], y = cy[k], z = cz[k];
  int64_t w = w0;
  wave_box_scan(g, grid_box(g, x, y, z, rad), [&](int p, bool valid) {
    double d = 0.0;
    if (valid)
      d = dist3(x, y, z, g.x[p], g.y[p], g.z[p]);
    const bool hit = valid && d <= rad;
    const uint64_t m = __ballot(hit);
    if (hit) {
      const int64_t o = w + lane_prefix(m);
      if (targets) {
        const unsigned member = d <= radMember ? 1u : 0u;
        out[o] = (int)((unsigned)p | (member << 31));
      } else {
        out[o] = p;
      }
      if (pairPatch)
        pairPatch[o] = (int)k;
    }
    w += __popcll(m);
  });
}

__global__ void pair_count_kernel(const int *pairTgt, int64_t P, int *tcnt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P)
    atomicAdd(&tcnt[pairTgt[p] & 0x7fffffff], 1);
}

// Lists every pair under its target: the pair id with the member flag in the
// top bit, and its patch beside it (read here in pair order, coalesced).
__global__ void pair_scatter_kernel(const int *pairTgt, const int *pairPatch,
                                    int64_t P, const int64_t *toff, int *tfill,
                                    int *tlist, int *tpatch) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P)
    return;
  const int pe = pairTgt[p];
  const int t = pe & 0x7fffffff;
  const int64_t at = toff[t] + atomicAdd(&tfill[t], 1);
  tlist[at] = (int)((unsigned)p | ((unsigned)pe & 0x80000000u));
  tpatch[at] = pairPatch[p];
}

// Per target: order its pairs by pair id (pair ids grow with the patch id, so
// this is patch order) and record each pair's slot, patch and member flag.
__global__ void target_slots_kernel(const int64_t *toff, int64_t M, int *tlist,
                                    int *tpatch, int *pairSlot, int *slotPatch,
                                    unsigned char *slotMember) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M)
    return;
  const int64_t s = toff[t], e = toff[t + 1];
  for (int64_t i = s + 1; i < e; ++i) {
    const int v = tlist[i], vp = tpatch[i];
    const int key = v & 0x7fffffff;
    int64_t j = i - 1;
    while (j >= s && (tlist[j] & 0x7fffffff) > key) {
      tlist[j + 1] = tlist[j];
      tpatch[j + 1] = tpatch[j];
      --j;
    }
    tlist[j + 1] = v;
    tpatch[j + 1] = vp;
  }
  for (int64_t i = s; i < e; ++i) {
    const int v = tlist[i];
    pairSlot[v & 0x7fffffff] = (int)i;
    slotPatch[i] = tpatch[i];
    slotMember[i] = (unsigned char)((unsigned)v >> 31);
  }
}

unsigned blocks_for(int64_t n, int per) {
  return (unsigned)std::max<int64_t>(1, (n + per - 1) / per);
}

} // namespace

// ===========================================================================
// Host entry points
// ===========================================================================
int64_t exclusive_scan_i32_to_i64(const int *in, int64_t *out, int64_t n,
                                  hipStream_t s) {
  int64_t tot = 0;
  scan_device<int, int64_t>(in, out, n, s, true, &tot);
  return tot;
}

void exclusive_scan_i32_to_i64_async(const int *in, int64_t *out, int64_t n,
                                     hipStream_t s) {
  scan_device<int, int64_t>(in, out, n, s, false, nullptr);
}

int exclusive_scan_i32(const int *in, int *out, int64_t n, hipStream_t s) {
  int tot = 0;
  scan_device<int, int>(in, out, n, s, true, &tot);
  return tot;
}

void bbox_device(const double *pts, int64_t n, int dim, double *lo,
                 double *hi, hipStream_t s) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = 0.0;
    hi[a] = 0.0;
  }
  if (n <= 0)
    return;
  const unsigned nb = (unsigned)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
  DevBuf<double> part((size_t)nb * 6);
  bbox_partial_kernel<<<nb, kBlock, 0, s>>>(pts, n, dim, part.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  std::vector<double> h((size_t)nb * 6);
  PACMAN_HIP_CHECK(hipMemcpyAsync(h.data(), part.get(), h.size() * 8,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  for (int a = 0; a < dim; ++a) {
    lo[a] = DBL_MAX;
    hi[a] = -DBL_MAX;
    for (unsigned b = 0; b < nb; ++b) {
      lo[a] = std::min(lo[a], h[b * 6 + a]);
      hi[a] = std::max(hi[a], h[b * 6 + 3 + a]);
    }
  }
}

void build_grid(const double *pts, int64_t n, int dim, const double *lo,
                const double *hi, double pointsPerCell, DeviceGrid &g,
                hipStream_t s) {
  // Cell size from the point density over the non-degenerate axes.
  double vol = 1.0;
  int dd = 0;
  double edges[3];
  for (int a = 0; a < 3; ++a) {
    edges[a] = (a < dim) ? hi[a] - lo[a] : 0.0;
    if (edges[a] > 0) {
      vol *= edges[a];
      ++dd;
    }
  }
  double h = 1.0;
  if (dd > 0 && n > 0)
    h = std::pow(vol * pointsPerCell / (double)n, 1.0 / dd);
  for (int a = 0; a < 3; ++a)
    if (edges[a] > 0)
      h = std::min(h, edges[a]);
  if (!(h > 0))
    h = 1.0;
  int64_t ncell = 1;
  int dims[3];
  for (int it = 0;; ++it) {
    ncell = 1;
    for (int a = 0; a < 3; ++a) {
      dims[a] = edges[a] > 0 ? (int)std::floor(edges[a] / h) + 1 : 1;
      ncell *= dims[a];
    }
    if (ncell <= ((int64_t)1 << 27) && ncell <= 4 * std::max<int64_t>(n, 1) + 64)
      break;
    h *= 1.25;
  }
  GridView &v = g.view;
  for (int a = 0; a < 3; ++a) {
    v.lo[a] = lo[a];
    v.dims[a] = dims[a];
  }
  v.inv_h = 1.0 / h;
  v.n = (int)n;
  g.ncell = ncell;
  g.cell_start.alloc(ncell + 1);
  g.perm.alloc(std::max<int64_t>(n, 1));
  g.x.alloc(std::max<int64_t>(n, 1));
  g.y.alloc(std::max<int64_t>(n, 1));
  g.z.alloc(std::max<int64_t>(n, 1));
  DevBuf<int> key(std::max<int64_t>(n, 1)), cnt(ncell), fill(ncell);
  fill_i32(cnt.get(), ncell, 0, s);
  fill_i32(fill.get(), ncell, 0, s);
  v.cell_start = g.cell_start.get();
  v.perm = g.perm.get();
  v.x = g.x.get();
  v.y = g.y.get();
  v.z = g.z.get();
  if (n > 0) {
    grid_count_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        v, pts, n, dim, key.get(), cnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  exclusive_scan_i32(cnt.get(), g.cell_start.get(), ncell, s);
  if (n > 0) {
    grid_scatter_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        key.get(), n, g.cell_start.get(), fill.get(), g.perm.get());
    grid_sort_cells_kernel<<<blocks_for(ncell, kBlock), kBlock, 0, s>>>(
        g.cell_start.get(), ncell, g.perm.get());
    grid_gather_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(
        pts, n, dim, g.perm.get(), g.x.get(), g.y.get(), g.z.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  // key, cnt and fill are released here while the kernels may still run
  // (see scan_device)
}

void probe_radius(const GridView &src, const double *probes, int nprobe,
                  int k, const double *lo, const double *hi, ProbeResult &res,
                  hipStream_t s) {
  double diag2 = 0.0;
  for (int a = 0; a < 3; ++a)
    diag2 += (hi[a] - lo[a]) * (hi[a] - lo[a]);
  const double radMax = std::sqrt(diag2) * 1.01 + 1e-300;
  const double rad0 = 1.0 / src.inv_h;
  DevBuf<double> dprobes(3 * nprobe), dmax(nprobe);
  DevBuf<int> dpos(nprobe);
  PACMAN_HIP_CHECK(hipMemcpyAsync(dprobes.get(), probes, 24 * nprobe,
                                  hipMemcpyHostToDevice, s));
  probe_nearest_kernel<<<nprobe, kBlock, 0, s>>>(src, dprobes.get(), rad0,
                                                 radMax, dpos.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  probe_knn_kernel<<<nprobe, kBlock, 0, s>>>(src, dpos.get(), k, rad0, radMax,
                                             dmax.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  PACMAN_HIP_CHECK(hipMemcpyAsync(res.maxd2, dmax.get(), 8 * nprobe,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipMemcpyAsync(res.nearest, dpos.get(), 4 * nprobe,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  for (int q = 0; q < nprobe; ++q)
    if (std::isnan(res.maxd2[q]))
      throw std::runtime_error(
          "patch radius probe: too many equidistant neighbours");
}

int64_t make_centers(const Lattice &Lh, const GridView &src,
                     const GridView &tgt, DevBuf<double> &ocx,
                     DevBuf<double> &ocy, DevBuf<double> &ocz,
                     DevBuf<int64_t> &ocand, hipStream_t s) {
  LatticeDev L{};
  L.dim = Lh.dim;
  for (int a = 0; a < 3; ++a) {
    L.shape[a] = Lh.shape[a];
    L.lower[a] = Lh.lower[a];
    L.dist[a] = Lh.dist[a];
  }
  L.nb = Lh.nb;
  L.r = Lh.r;
  L.r2 = Lh.r2;
  L.thr = Lh.thr;
  L.nOff = (int)Lh.posOffsets.size();
  for (int k = 0; k < L.nOff && k < 13; ++k)
    L.off[k] = Lh.posOffsets[k];
  const int64_t nb = L.nb;
  if (nb > INT_MAX)
    throw std::runtime_error("too many lattice candidates");
  DevBuf<double> cx(nb), cy(nb), cz(nb);
  DevBuf<int> f1(nb), f2(nb), f3(nb), pos(nb + 1);
  const unsigned wb = blocks_for(nb, kBlock / kWave);
  lattice_tag_snap_kernel<<<wb, kBlock, 0, s>>>(L, src, tgt, cx.get(), cy.get(),
                                                cz.get(), f1.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  lattice_prune_kernel<<<blocks_for(nb, kBlock), kBlock, 0, s>>>(
      L, cx.get(), cy.get(), cz.get(), f1.get(), f2.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  lattice_retag_kernel<<<wb, kBlock, 0, s>>>(L, tgt, cx.get(), cy.get(),
                                             cz.get(), f2.get(), f3.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  const int K = exclusive_scan_i32(f3.get(), pos.get(), nb, s);
  ocx.alloc(std::max(K, 1));
  ocy.alloc(std::max(K, 1));
  ocz.alloc(std::max(K, 1));
  ocand.alloc(std::max(K, 1));
  lattice_compact_kernel<<<blocks_for(nb, kBlock), kBlock, 0, s>>>(
      nb, f3.get(), pos.get(), cx.get(), cy.get(), cz.get(), ocx.get(),
      ocy.get(), ocz.get(), ocand.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  // temporaries are released while the kernels may still run (scan_device)
  return K;
}

// Rows of the one-pass membership scratch: wide enough for the patches the
// solve takes; larger patches are listed again from the geometry.
constexpr int kMemberCap = 256;

void patch_sources(const GridView &src, const double *cx, const double *cy,
                   const double *cz, int64_t K, double r,
                   DevBuf<int64_t> &off, DevBuf<int> &idx, hipStream_t s) {
  DevBuf<int> cnt(std::max<int64_t>(K, 1));
  DevBuf<int> scratch((size_t)std::max<int64_t>(K, 1) * kMemberCap);
  off.alloc(K + 1);
  const unsigned wb = blocks_for(K, kBlock / kWave);
  if (K > 0) {
    patch_src_capped_kernel<<<wb, kBlock, 0, s>>>(src, cx, cy, cz, K, r,
                                                  kMemberCap, scratch.get(),
                                                  cnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  const int64_t tot = exclusive_scan_i32_to_i64(cnt.get(), off.get(), K, s);
  if (tot > INT_MAX)
    throw std::runtime_error("patch source lists exceed 2^31 entries");
  idx.alloc(std::max<int64_t>(tot, 1));
  if (K > 0) {
    member_compact_kernel<<<wb, kBlock, 0, s>>>(
        src, cx, cy, cz, K, r, r, false, kMemberCap, scratch.get(), cnt.get(),
        off.get(), idx.get(), nullptr);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  // temporaries are released while the kernels may still run (scan_device)
}

void patch_targets(const GridView &tgt, const double *cx, const double *cy,
                   const double *cz, int64_t K, double r, PairSet &ps,
                   hipStream_t s) {
  const int64_t M = tgt.n;
  DevBuf<int> cnt(std::max<int64_t>(K, 1));
  DevBuf<int> scratch((size_t)std::max<int64_t>(K, 1) * kMemberCap);
  ps.off.alloc(K + 1);
  const unsigned wb = blocks_for(K, kBlock / kWave);
  if (K > 0) {
    // solver.hpp:62: targets use the radius r - epsilon
    patch_tgt_capped_kernel<<<wb, kBlock, 0, s>>>(tgt, cx, cy, cz, K, r,
                                                  r - kEps, kMemberCap,
                                                  scratch.get(), cnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  const int64_t P = exclusive_scan_i32_to_i64(cnt.get(), ps.off.get(), K, s);
  if (P > INT_MAX)
    throw std::runtime_error("patch/target pairs exceed 2^31 entries");
  ps.P = P;
  const int64_t P1 = std::max<int64_t>(P, 1);
  ps.tgt.alloc(P1);
  DevBuf<int> pairPatch(P1);
  if (K > 0) {
    member_compact_kernel<<<wb, kBlock, 0, s>>>(
        tgt, cx, cy, cz, K, r, r - kEps, true, kMemberCap, scratch.get(),
        cnt.get(), ps.off.get(), ps.tgt.get(), pairPatch.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  build_target_slots(ps, pairPatch.get(), M, s);
  // temporaries are released while the kernels may still run (scan_device)
}

// Per-target transpose of the pairs (ps.off, ps.tgt, ps.P set; pairPatch[q]
// = patch of pair q): slots in ascending patch order.
void build_target_slots(PairSet &ps, const int *pairPatch, int64_t M,
                        hipStream_t s) {
  const int64_t P = ps.P;
  const int64_t P1 = std::max<int64_t>(P, 1);
  ps.slot.alloc(P1);
  ps.slotPatch.alloc(P1);
  ps.slotMember.alloc(P1);
  DevBuf<int> tlist(P1), tpatch(P1);
  DevBuf<int> tcnt(std::max<int64_t>(M, 1)), tfill(std::max<int64_t>(M, 1));
  fill_i32(tcnt.get(), M, 0, s);
  fill_i32(tfill.get(), M, 0, s);
  ps.toff.alloc(M + 1);
  if (P > 0) {
    pair_count_kernel<<<blocks_for(P, kBlock), kBlock, 0, s>>>(
        ps.tgt.get(), P, tcnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  exclusive_scan_i32_to_i64_async(tcnt.get(), ps.toff.get(), M, s);
  if (P > 0) {
    pair_scatter_kernel<<<blocks_for(P, kBlock), kBlock, 0, s>>>(
        ps.tgt.get(), pairPatch, P, ps.toff.get(), tfill.get(),
        tlist.get(), tpatch.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  if (M > 0) {
    target_slots_kernel<<<blocks_for(M, kBlock), kBlock, 0, s>>>(
        ps.toff.get(), M, tlist.get(), tpatch.get(), ps.slot.get(),
        ps.slotPatch.get(), ps.slotMember.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  // temporaries are released while the kernels may still run (scan_device)
}

} // namespace pacman_amd
