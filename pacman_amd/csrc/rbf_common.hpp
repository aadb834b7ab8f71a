// Device-side building blocks shared by the RBF-PUM kernels (gfx950).
//
// Geometry is padded to three coordinates (zeros beyond the space
// dimension); the padding adds exact zeros to every squared distance, so the
// 1-D and 2-D cases evaluate the same expressions as the 3-D one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pacman_amd {

constexpr double kEps = 1.0e-14;     // src/common/types.hpp:98
constexpr double kCondMax = 1.0e5;   // src/rbf_pum/solver/utils/serial_ops.hpp:130
constexpr int kWave = 64;

// Squared distance sum_a (rhs_a - lhs_a)^2 in axis order, as
// SquaredDifference (src/rbf_pum/utils/operators.hpp:125-137). The geometry
// translation unit is compiled with -ffp-contract=off so this is evaluated
// exactly as written.
__host__ __device__ inline double sqdiff3(double lx, double ly, double lz,
                                          double rx, double ry, double rz) {
  const double dx = rx - lx, dy = ry - ly, dz = rz - lz;
  double acc = dx * dx;
  acc = acc + dy * dy;
  acc = acc + dz * dz;
  return acc;
}

// Wendland kernels of src/rbf_pum/solver/rbf_functions.hpp:22-150.
// KIND 0..4 = C0, C2, C4, C6, C8; p = r * radiusInv, zero outside [0, 1).
// NONNEG: the caller guarantees r >= 0 and rinv > 0 (a distance), so the
// mask reduces to p < 1 with the same result; it is applied as q = max(1 - p,
// 0), which leaves q unchanged inside the support and makes every kernel
// (a power of q times a polynomial positive on [0, 1]) +0 outside.
template <int KIND, typename T, bool NONNEG = false>
__host__ __device__ inline T wendland(T r, T rinv) {
  const T p = r * rinv;
  T q = T(1) - p;
  if constexpr (NONNEG)
    q = fmax(q, T(0));
  const T q2 = q * q;
  T v;
  if constexpr (KIND == 0) {
    v = q2;
  } else if constexpr (KIND == 1) {
    v = (q2 * q2) * (T(4) * p + T(1));
  } else if constexpr (KIND == 2) {
    const T q6 = q2 * q2 * q2;
    v = q6 * (T(35) * p * p + (T(18) * p + T(3)));
  } else if constexpr (KIND == 3) {
    const T q4 = q2 * q2;
    v = (q4 * q4) * (T(32) * p * p * p + T(25) * p * p + (T(8) * p + T(1)));
  } else {
    const T q4 = q2 * q2;
    const T q10 = q4 * q4 * q2;
    const T p2 = p * p;
    v = q10 * (T(1287) * p2 * p2 + T(1350) * p2 * p + T(630) * p2 +
               (T(150) * p + T(15)));
  }
  if constexpr (NONNEG)
    return v;
  else
    return (r >= T(0) && p < T(1) && p >= T(0)) ? v : T(0);
}

// A cell-sorted point set. Cells are a uniform grid over the set's bounding
// box; the cell key is row-major with the last axis fastest, so the cells of
// one (x, y) column are contiguous and a box query reads one contiguous run
// of points per column. Inside a cell, points are in ascending original index.
struct GridView {
  double lo[3];
  double inv_h;
  int dims[3];
  int n;
  const int *cell_start;  // ncell + 1
  const double *x, *y, *z; // coordinates in cell order
  const int *perm;         // cell-order position -> original index
};

__host__ __device__ inline int grid_axis_cell(const GridView &g, int a,
                                              double v) {
  double t = floor((v - g.lo[a]) * g.inv_h);
  if (t < 0.0)
    t = 0.0;
  const double hi = (double)(g.dims[a] - 1);
  if (t > hi)
    t = hi;
  return (int)t;
}

// Column runs of the box [c - rad, c + rad] (padded by a relative 1e-9 so
// points accepted by a rounded distance test are never missed).
struct BoxRange {
  int x0, x1, y0, y1, z0, z1;
};

__device__ inline BoxRange grid_box(const GridView &g, double cx, double cy,
                                    double cz, double rad) {
  const double pad = rad * 1e-9 + 1e-300;
  BoxRange b;
  b.x0 = grid_axis_cell(g, 0, cx - rad - pad);
  b.x1 = grid_axis_cell(g, 0, cx + rad + pad);
  b.y0 = grid_axis_cell(g, 1, cy - rad - pad);
  b.y1 = grid_axis_cell(g, 1, cy + rad + pad);
  b.z0 = grid_axis_cell(g, 2, cz - rad - pad);
  b.z1 = grid_axis_cell(g, 2, cz + rad + pad);
  return b;
}

// Wave-wide reductions (wave64, no LDS).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1)
    v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}

// Lane rank among the lanes whose predicate is set (ordered compaction).
__device__ inline int lane_prefix(uint64_t mask) {
  const uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
  return __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
}

// Blocks are dispatched round-robin over the 8 XCDs of MI355X, each with
// its own L2. Map block b so every XCD walks one contiguous range of work
// items: neighbouring patches share points, so their gathers hit the same L2.
__device__ inline int xcd_remap(int b, int count) {
  const int q = count >> 3, rem = count & 7;
  const int x = b & 7, j = b >> 3;
  return x * q + min(x, rem) + j;
}

} // namespace pacman_amd
