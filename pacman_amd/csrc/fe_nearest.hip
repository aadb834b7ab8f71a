// Nearest-source transfer (FE method NEAREST_NEAREST) on gfx950.
//
// The reference (src/finite_elements/FTNearest.hpp:28-52) builds an ArborX
// BVH over the sources and copies, for every target, the value of its
// nearest source (NearestExtractIndex,
// src/finite_elements/utils/ArborXCallbacks.hpp:172-182); the status stays
// OUTSIDE and targets keep 0 when there is no source
// (src/common/transfer.hxx:100-109).
//
// Here the sources are cell-sorted once (the grid of rbf_geometry.hip) and
// every target thread searches shells of cells around its own cell until
// the best squared distance is below the distance to the unsearched region.
// Targets are cell-sorted too and processed in that order, so the threads of
// a wave search neighbouring cells (coherent loops, L2 reuse) and scatter
// their results back to the original order.
// The minimum is over (squared distance, source index): ties go to the
// lowest index, so the answer is unique and equals brute force. Compiled
// with -ffp-contract=off: distances round exactly like the oracle's.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

constexpr int kNearestBlock = 256;

// Distance from coordinate v to the part of the line outside the cell range
// [c0, c1] of axis a; infinity when the range reaches both ends of the grid.
__device__ inline double outside_gap(const GridView &g, int a, double v,
                                     int c0, int c1, double h) {
  double gap = DBL_MAX;
  if (c0 > 0)
    gap = fmin(gap, v - (g.lo[a] + (double)c0 * h));
  if (c1 < g.dims[a] - 1)
    gap = fmin(gap, (g.lo[a] + (double)(c1 + 1) * h) - v);
  return gap;
}

__global__ __launch_bounds__(kNearestBlock) void nearest_value_kernel(
    GridView g, GridView tg, const double *vals, double *out, int *status) {
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p0 >= tg.n)
    return;
  const int64_t t = tg.perm[p0];  // original target index
  const double q[3] = {tg.x[p0], tg.y[p0], tg.z[p0]};
  status[t] = 0;  // TransferStatus::OUTSIDE
  if (g.n == 0) {
    out[t] = 0.0;
    return;
  }
  const double h = 1.0 / g.inv_h;
  int c[3];
  for (int a = 0; a < 3; ++a)
    c[a] = grid_axis_cell(g, a, q[a]);
  double best = DBL_MAX;
  int bestIdx = INT_MAX;
  for (int rho = 0;; ++rho) {
    const int x0 = max(c[0] - rho, 0), x1 = min(c[0] + rho, g.dims[0] - 1);
    const int y0 = max(c[1] - rho, 0), y1 = min(c[1] + rho, g.dims[1] - 1);
    const int z0 = max(c[2] - rho, 0), z1 = min(c[2] + rho, g.dims[2] - 1);
    for (int ix = x0; ix <= x1; ++ix) {
      const bool xin = ix > c[0] - rho && ix < c[0] + rho;
      for (int iy = y0; iy <= y1; ++iy) {
        const bool yin = iy > c[1] - rho && iy < c[1] + rho;
        // cells of this (x, y) column on the shell of Chebyshev radius rho
        const int zs[2] = {c[2] - rho, c[2] + rho};
        const bool inner = xin && yin;
        const int nseg = inner ? (rho > 0 ? 2 : 1) : 1;
        for (int sgi = 0; sgi < nseg; ++sgi) {
          int za, zb;
          if (inner) {
            za = zb = zs[sgi];
            if (za < z0 || za > z1)
              continue;
          } else {
            za = z0;
            zb = z1;
          }
          const int cell0 = (ix * g.dims[1] + iy) * g.dims[2];
          for (int p = g.cell_start[cell0 + za]; p < g.cell_start[cell0 + zb + 1];
               ++p) {
            const double d2 = sqdiff3(q[0], q[1], q[2], g.x[p], g.y[p], g.z[p]);
            const int o = g.perm[p];
            if (d2 < best || (d2 == best && o < bestIdx)) {
              best = d2;
              bestIdx = o;
            }
          }
        }
      }
    }
    // every point outside the searched cells is at least `gap` away
    double gap = DBL_MAX;
    gap = fmin(gap, outside_gap(g, 0, q[0], x0, x1, h));
    gap = fmin(gap, outside_gap(g, 1, q[1], y0, y1, h));
    gap = fmin(gap, outside_gap(g, 2, q[2], z0, z1, h));
    if (gap == DBL_MAX)
      break;  // the whole grid has been searched
    // conservative margin for the rounding of the cell assignment
    const double lb = gap - 1e-9 * (h + fabs(q[0]) + fabs(q[1]) + fabs(q[2]));
    if (lb > 0.0 && best < lb * lb)
      break;
  }
  // no comparable source (non-finite coordinates): keep the initial 0
  out[t] = bestIdx == INT_MAX ? 0.0 : vals[bestIdx];
}

} // namespace

void fe_nearest_transfer(int dim, const double *dSrc, int64_t N,
                         const double *dVals, const double *dTgt, int64_t M,
                         double *dOut, int *dStatus, hipStream_t s) {
  if (M <= 0)
    return;
  double lo[3], hi[3], tlo[3], thi[3];
  bbox_device(dSrc, N, dim, lo, hi, s);
  bbox_device(dTgt, M, dim, tlo, thi, s);
  DeviceGrid g, tg;
  build_grid(dSrc, N, dim, lo, hi, 2.0, g, s);
  build_grid(dTgt, M, dim, tlo, thi, 4.0, tg, s);
  const unsigned blocks = (unsigned)((M + kNearestBlock - 1) / kNearestBlock);
  nearest_value_kernel<<<blocks, kNearestBlock, 0, s>>>(g.view, tg.view,
                                                         dVals, dOut, dStatus);
  PACMAN_HIP_CHECK(hipGetLastError());
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

} // namespace pacman_amd
