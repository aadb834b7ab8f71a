// Moving-least-squares transfer (the reference's MLS method) on gfx950.
//
// The reference (src/mls/interpolator.hpp:47-75) hands the points to ArborX's
// MovingLeastSquares with its defaults: for every target y, the k nearest
// sources (k = 3, 6, 10 monomials of degree <= 2 in dimension 1, 2, 3),
// offsets s_j = x_j - y, radius rho = 1.1 max_j |s_j|, weights
// phi_j = (1 - |s_j| / rho)^2 (Wendland<0>) and the value
// c_0 = e_0^T (P^T Phi P)+ P^T Phi f, P the Vandermonde matrix at s_j.
//
// Setup (one thread per target, targets processed in cell order; the search
// and the least-squares weights are two kernels):
//   * k nearest sources by a shell search over the cell grid of the sources,
//     a sorted register list of (squared distance, index); ties go to the
//     lowest source index, and the search stops once the k-th distance is
//     below the distance to every unsearched cell;
//   * the weighted least-squares problem min |Phi^1/2 (P c - f)| solved by
//     Householder QR with column pivoting of B = Phi^1/2 P in registers
//     (mls_weights below).
// Apply: out_f(y) = sum_j w_j f(x_j), the k weights and indices of y kept in
// the plan (an HBM-bound gather).
// Compiled with -ffp-contract=off: squared distances round exactly like the
// oracle's, so the neighbour lists are identical.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "mls.hpp"

namespace pacman_amd {

namespace {

constexpr int kMlsBlock = 128;

template <int D> struct MlsBasis {
  static constexpr int K = (D + 1) * (D + 2) / 2;
};

} // namespace

// The same weights for a neighbourhood whose B is rank-deficient, where
// c_0 is not fixed by the fit (e.g. lattice neighbours on a conic): the
// minimum-norm fit of the pseudo-inverse, row 0 of B+ by a one-sided Jacobi
// SVD (B V = U S; B+ = V S+ U^T, singular values at or below K eps max
// dropped). Only row 0 of V is needed. Same operations, in the same order,
// as the restatement used to check this code.
template <int D>
__host__ __device__ inline void mls_weights_svd(const double *sx,
                                                const double *sy,
                                                const double *sz, double *w) {
  constexpr int K = (D + 1) * (D + 2) / 2;
  double rad = DBL_EPSILON;
  double nrm[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double n2 = sx[j] * sx[j];
    if (D > 1)
      n2 = n2 + sy[j] * sy[j];
    if (D > 2)
      n2 = n2 + sz[j] * sz[j];
    nrm[j] = sqrt(n2);
    rad = fmax(rad, nrm[j]);
  }
  rad *= 1.1;
  const double rinv = 1.0 / rad;
  double sw[K], B[K][K], v0[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double x = nrm[j] / rad;
    sw[j] = x < 1.0 ? 1.0 - x : 0.0;
    const double sc[3] = {sx[j] * rinv, sy[j] * rinv, sz[j] * rinv};
    int c = 0;
    B[j][c++] = 1.0;
#pragma unroll
    for (int a = 0; a < D; ++a)
      B[j][c++] = sc[a];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = a; b < D; ++b)
        B[j][c++] = sc[a] * sc[b];
#pragma unroll
    for (int u = 0; u < K; ++u)
      B[j][u] = sw[j] * B[j][u];
    v0[j] = j == 0 ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
#pragma unroll
    for (int p = 0; p < K; ++p)
#pragma unroll
      for (int q = p + 1; q < K; ++q) {
        double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          a += B[j][p] * B[j][p];
          b += B[j][q] * B[j][q];
          g += B[j][p] * B[j][q];
        }
        if (g == 0.0 || fabs(g) <= DBL_EPSILON * sqrt(a * b))
          continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                         (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const double bp = B[j][p], bq = B[j][q];
          B[j][p] = c * bp - sn * bq;
          B[j][q] = sn * bp + c * bq;
        }
        const double vp = v0[p], vq = v0[q];
        v0[p] = c * vp - sn * vq;
        v0[q] = sn * vp + c * vq;
      }
    if (!rotated)
      break;
  }
  double s2[K], s2max = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    s2[i] = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j)
      s2[i] += B[j][i] * B[j][i];
    s2max = fmax(s2max, s2[i]);
  }
  const double cut = (double)K * DBL_EPSILON;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (s2[i] > cut * cut * s2max)
        acc += v0[i] * B[j][i] / s2[i];
    w[j] = acc * sw[j];
  }
}

// MLS weights of one target from the offsets s_j = x_j - y of its K nearest
// sources (K = (D+1)(D+2)/2): value(y) = sum_j w_j f_j. Weighted least
// squares min |Phi^1/2 (P c - f)| with P the monomials at s_j / rho, by
// Householder QR of B = Phi^1/2 P with column pivoting (the column of
// largest remaining norm first). The factorisation stops at the first
// remaining norm below 1e-13 of the first pivot: those columns depend on the
// others (on regular grids some monomial columns vanish or are proportional)
// and the weights come from mls_weights_svd. Otherwise
// c_0 = e_0^T R^-1 Q^T Phi^1/2 f, so w = Phi^1/2 Q R^-T e_0' (e_0' the slot
// of the constant column). Host and device (the host build is for checks).
template <int D>
__host__ __device__ inline void mls_weights(const double *sx, const double *sy,
                                            const double *sz, double *w) {
  constexpr int K = (D + 1) * (D + 2) / 2;
  double rad = DBL_EPSILON;
  double nrm[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double n2 = sx[j] * sx[j];
    if (D > 1)
      n2 = n2 + sy[j] * sy[j];
    if (D > 2)
      n2 = n2 + sz[j] * sz[j];
    nrm[j] = sqrt(n2);
    rad = fmax(rad, nrm[j]);
  }
  rad *= 1.1;
  const double rinv = 1.0 / rad;
  double sw[K];  // phi_j^1/2 = 1 - |s_j| / rho
  double B[K][K];
  int perm[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double x = nrm[j] / rad;
    sw[j] = x < 1.0 ? 1.0 - x : 0.0;
    const double sc[3] = {sx[j] * rinv, sy[j] * rinv, sz[j] * rinv};
    double pb[K];
    int c = 0;
    pb[c++] = 1.0;
#pragma unroll
    for (int a = 0; a < D; ++a)
      pb[c++] = sc[a];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = a; b < D; ++b)
        pb[c++] = sc[a] * sc[b];
#pragma unroll
    for (int u = 0; u < K; ++u)
      B[j][u] = sw[j] * pb[u];
  }
#pragma unroll
  for (int u = 0; u < K; ++u)
    perm[u] = u;

  double beta[K], rdiag[K];
  int rank = 0;
  bool alive = true;
  double first = 0.0;
#pragma unroll
  for (int cc = 0; cc < K; ++cc) {
    beta[cc] = 0.0;
    rdiag[cc] = 0.0;
    if (alive) {
      // pivot: the remaining column of largest norm
      double best = -1.0;
      int piv = cc;
#pragma unroll
      for (int d = cc; d < K; ++d) {
        double n2 = 0.0;
#pragma unroll
        for (int i = cc; i < K; ++i)
          n2 += B[i][d] * B[i][d];
        if (n2 > best) {
          best = n2;
          piv = d;
        }
      }
      const double nrm_c = sqrt(best);
      if (cc == 0)
        first = nrm_c;
      if (!(nrm_c > 1e-13 * first)) {
        alive = false;
      } else {
#pragma unroll
        for (int d = cc + 1; d < K; ++d)
          if (d == piv) {
#pragma unroll
            for (int i = 0; i < K; ++i) {
              const double tmp = B[i][cc];
              B[i][cc] = B[i][d];
              B[i][d] = tmp;
            }
            const int tp = perm[cc];
            perm[cc] = perm[d];
            perm[d] = tp;
          }
        const double x0 = B[cc][cc];
        const double alpha = x0 >= 0.0 ? -nrm_c : nrm_c;
        const double v0 = x0 - alpha;
        const double vtv = best - x0 * x0 + v0 * v0;
        rdiag[cc] = alpha;
        rank = cc + 1;
        if (vtv > 0.0) {
          const double bt = 2.0 / vtv;
          beta[cc] = bt;
          B[cc][cc] = v0;  // v = (v0, B[cc+1..][cc])
#pragma unroll
          for (int d = cc + 1; d < K; ++d) {
            double dot = 0.0;
#pragma unroll
            for (int i = cc; i < K; ++i)
              dot += B[i][cc] * B[i][d];
            dot *= bt;
#pragma unroll
            for (int i = cc; i < K; ++i)
              B[i][d] -= dot * B[i][cc];
          }
        }
      }
    }
  }
  if (rank < K) {  // c_0 may not be fixed by the fit: minimum-norm fit
    mls_weights_svd<D>(sx, sy, sz, w);
    return;
  }
  // z = R^-T e (e: the slot of the constant column), R[l][i] = B[l][i], l < i
  double z[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double acc = perm[i] == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int l = 0; l < i; ++l)
      acc -= B[l][i] * z[l];
    z[i] = i < rank ? acc / rdiag[i] : 0.0;
  }
  // y = Q z = H_0 ... H_{rank-1} z
#pragma unroll
  for (int cc = K - 1; cc >= 0; --cc) {
    if (cc < rank && beta[cc] != 0.0) {
      double dot = 0.0;
#pragma unroll
      for (int i = cc; i < K; ++i)
        dot += B[i][cc] * z[i];
      dot *= beta[cc];
#pragma unroll
      for (int i = cc; i < K; ++i)
        z[i] -= dot * B[i][cc];
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j)
    w[j] = z[j] * sw[j];
}

namespace {

__device__ inline double outside_gap(const GridView &g, int a, double v,
                                     int c0, int c1, double h) {
  double gap = DBL_MAX;
  if (c0 > 0)
    gap = fmin(gap, v - (g.lo[a] + (double)c0 * h));
  if (c1 < g.dims[a] - 1)
    gap = fmin(gap, (g.lo[a] + (double)(c1 + 1) * h) - v);
  return gap;
}

__device__ __forceinline__ bool lex_less(double d1, int o1, double d2,
                                         int o2) {
  return d1 < d2 || (d1 == d2 && o1 < o2);
}

// k nearest sources of every target (targets in cell order): the original
// source index to idxOut and the cell-order position to posOut, both
// [t * K + j] in ascending (distance, index) order.
template <int D>
__global__ __launch_bounds__(kMlsBlock) void mls_knn_kernel(
    GridView g, GridView tg, int *idxOut, int *posOut) {
  constexpr int K = MlsBasis<D>::K;
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p0 >= tg.n)
    return;
  const int64_t t = tg.perm[p0];  // original target index
  const double q[3] = {tg.x[p0], tg.y[p0], tg.z[p0]};

  // ---- k nearest sources -------------------------------------------------
  double bd[K];
  int bo[K], bp[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    bd[j] = DBL_MAX;
    bo[j] = INT_MAX;
    bp[j] = 0;
  }
  const double h = 1.0 / g.inv_h;
  int c[3];
  for (int a = 0; a < 3; ++a)
    c[a] = grid_axis_cell(g, a, q[a]);
  for (int rho = 0;; ++rho) {
    const int x0 = max(c[0] - rho, 0), x1 = min(c[0] + rho, g.dims[0] - 1);
    const int y0 = max(c[1] - rho, 0), y1 = min(c[1] + rho, g.dims[1] - 1);
    const int z0 = max(c[2] - rho, 0), z1 = min(c[2] + rho, g.dims[2] - 1);
    for (int ix = x0; ix <= x1; ++ix) {
      const bool xin = ix > c[0] - rho && ix < c[0] + rho;
      for (int iy = y0; iy <= y1; ++iy) {
        const bool yin = iy > c[1] - rho && iy < c[1] + rho;
        // cells of this (x, y) column on the shell of Chebyshev radius rho
        const bool inner = xin && yin;
        const int nseg = inner ? (rho > 0 ? 2 : 1) : 1;
        for (int sgi = 0; sgi < nseg; ++sgi) {
          int za, zb;
          if (inner) {
            za = zb = sgi == 0 ? c[2] - rho : c[2] + rho;
            if (za < z0 || za > z1)
              continue;
          } else {
            za = z0;
            zb = z1;
          }
          const int cell0 = (ix * g.dims[1] + iy) * g.dims[2];
#pragma unroll 1
          for (int p = g.cell_start[cell0 + za]; p < g.cell_start[cell0 + zb + 1];
               ++p) {
            const double d2 = sqdiff3(q[0], q[1], q[2], g.x[p], g.y[p], g.z[p]);
            const int o = g.perm[p];
            if (lex_less(d2, o, bd[K - 1], bo[K - 1])) {
              // insert at pos = number of entries before it; shift the rest
              int pos = 0;
#pragma unroll
              for (int j = 0; j < K - 1; ++j)
                pos += lex_less(bd[j], bo[j], d2, o) ? 1 : 0;
#pragma unroll
              for (int j = K - 1; j > 0; --j) {
                const bool sh = j > pos, at = j == pos;
                bd[j] = sh ? bd[j - 1] : (at ? d2 : bd[j]);
                bo[j] = sh ? bo[j - 1] : (at ? o : bo[j]);
                bp[j] = sh ? bp[j - 1] : (at ? p : bp[j]);
              }
              if (pos == 0) {
                bd[0] = d2;
                bo[0] = o;
                bp[0] = p;
              }
            }
          }
        }
      }
    }
    double gap = DBL_MAX;
    gap = fmin(gap, outside_gap(g, 0, q[0], x0, x1, h));
    gap = fmin(gap, outside_gap(g, 1, q[1], y0, y1, h));
    gap = fmin(gap, outside_gap(g, 2, q[2], z0, z1, h));
    if (gap == DBL_MAX)
      break;  // the whole grid has been searched
    // conservative margin for the rounding of the cell assignment
    const double lb = gap - 1e-9 * (h + fabs(q[0]) + fabs(q[1]) + fabs(q[2]));
    if (bo[K - 1] != INT_MAX && lb > 0.0 && bd[K - 1] < lb * lb)
      break;
  }

#pragma unroll
  for (int j = 0; j < K; ++j) {
    idxOut[t * K + j] = bo[j];
    posOut[t * K + j] = bp[j];
  }
}

// Weights of every target from its neighbours (a separate kernel: the QR
// needs about 200 registers, the search few, so each runs at its own
// occupancy).
template <int D>
__global__ __launch_bounds__(kMlsBlock) void mls_weights_kernel(
    GridView g, GridView tg, const int *pos, double *coefOut) {
  constexpr int K = MlsBasis<D>::K;
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p0 >= tg.n)
    return;
  const int64_t t = tg.perm[p0];
  const double q[3] = {tg.x[p0], tg.y[p0], tg.z[p0]};
  double sx[K], sy[K], sz[K], w[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int bp = pos[t * K + j];
    sx[j] = g.x[bp] - q[0];
    sy[j] = g.y[bp] - q[1];
    sz[j] = g.z[bp] - q[2];
  }
  mls_weights<D>(sx, sy, sz, w);
#pragma unroll
  for (int j = 0; j < K; ++j)
    coefOut[t * K + j] = w[j];
}

template <int K>
__global__ __launch_bounds__(256) void mls_apply_kernel(
    const int *idx, const double *coef, const double *vals, int64_t N,
    int64_t M, int F, double *out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M)
    return;
  int id[K];
  double w[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    id[j] = idx[t * K + j];
    w[j] = coef[t * K + j];
  }
  for (int f = 0; f < F; ++f) {
    const double *v = vals + (int64_t)f * N;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j)
      acc = fma(w[j], v[id[j]], acc);
    out[(int64_t)f * M + t] = acc;
  }
}

} // namespace

MlsPlan::MlsPlan(int dim, const double *dSrc, int64_t N, const double *dTgt,
                 int64_t M, hipStream_t s)
    : dim_(dim), N_(N), M_(M), k_((dim + 1) * (dim + 2) / 2) {
  if (dim < 1 || dim > 3)
    throw std::invalid_argument(
        "The dimension of the points can only be: 1, 2 or 3.\n");
  if (N < k_)
    throw std::invalid_argument(
        "MLS needs at least " + std::to_string(k_) +
        " source points in dimension " + std::to_string(dim) + " (got " +
        std::to_string(N) + ")");
  if (M <= 0)
    return;
  if (N > INT_MAX || M > INT_MAX)
    throw std::invalid_argument("MLS: more than 2^31 - 1 points");
  double lo[3], hi[3], tlo[3], thi[3];
  bbox_device(dSrc, N, dim, lo, hi, s);
  bbox_device(dTgt, M, dim, tlo, thi, s);
  DeviceGrid g, tg;
  build_grid(dSrc, N, dim, lo, hi, 4.0, g, s);
  build_grid(dTgt, M, dim, tlo, thi, 4.0, tg, s);
  idx_.alloc((size_t)M * k_);
  coef_.alloc((size_t)M * k_);
  DevBuf<int> pos((size_t)M * k_);
  const unsigned blocks = (unsigned)((M + kMlsBlock - 1) / kMlsBlock);
  switch (dim) {
  case 1:
    mls_knn_kernel<1><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                   idx_.get(), pos.get());
    mls_weights_kernel<1><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                       pos.get(), coef_.get());
    break;
  case 2:
    mls_knn_kernel<2><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                   idx_.get(), pos.get());
    mls_weights_kernel<2><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                       pos.get(), coef_.get());
    break;
  default:
    mls_knn_kernel<3><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                   idx_.get(), pos.get());
    mls_weights_kernel<3><<<blocks, kMlsBlock, 0, s>>>(g.view, tg.view,
                                                       pos.get(), coef_.get());
  }
  PACMAN_HIP_CHECK(hipGetLastError());
  // the grids are released at the end of this scope: wait for the kernel
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

void MlsPlan::apply(const double *dVals, int F, double *dOut,
                    hipStream_t s) const {
  if (F < 1)
    throw std::invalid_argument("apply needs at least one field");
  if (M_ <= 0)
    return;
  const unsigned blocks = (unsigned)((M_ + 255) / 256);
  switch (k_) {
  case 3:
    mls_apply_kernel<3><<<blocks, 256, 0, s>>>(idx_.get(), coef_.get(), dVals,
                                               N_, M_, F, dOut);
    break;
  case 6:
    mls_apply_kernel<6><<<blocks, 256, 0, s>>>(idx_.get(), coef_.get(), dVals,
                                               N_, M_, F, dOut);
    break;
  default:
    mls_apply_kernel<10><<<blocks, 256, 0, s>>>(idx_.get(), coef_.get(),
                                                dVals, N_, M_, F, dOut);
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

void MlsPlan::neighbours(int *idx, double *coef, hipStream_t s) const {
  if (M_ <= 0)
    return;
  PACMAN_HIP_CHECK(hipMemcpyAsync(idx, idx_.get(), sizeof(int) * M_ * k_,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipMemcpyAsync(coef, coef_.get(),
                                  sizeof(double) * M_ * k_,
                                  hipMemcpyDeviceToHost, s));
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

} // namespace pacman_amd
