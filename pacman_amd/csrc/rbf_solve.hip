// Per-patch RBF-PUM solve and partition-of-unity blend (gfx950, FP64).
//
// Two kernels per size bin run the per-patch chain of
// src/rbf_pum/solver/solver.hpp:180-286 without writing any matrix to HBM.
//
// poly_kernel (one wave per patch): least-squares fit of the linear
// polynomial (SerialSolveQR, serial_solve.hpp:46-67) with the conditioning
// loop of SerialFillQ (serial_ops.hpp:108-139). The Gram matrix of the basis
// [1, (x_a - c_a) / r] (c = patch centre, r = patch radius) is reduced over
// the wave; it spans the same space as the reference's [1, x_a] and is well
// conditioned. The condition number of the reference's raw Q = [1, x_a] is
// that of R'' T (R'' = Cholesky factor of the scaled Gram matrix, T the
// basis change): a Frobenius bound settles cond <= 1e5 at once, the exact
// Jacobi singular values are computed only otherwise. While cond > 1e5 the
// active axis of smallest coordinate range is dropped (serial_ops.hpp:28-57)
// and the fit uses exactly the active columns. It writes p(x) = c0 +
// c1 (x - cx) + c2 (y - cy) + c3 (z - cz) per field and the active column
// count.
//
// solve_f64_kernel (rbf_solve_f64.hpp, one translation unit per Wendland
// kind; one workgroup of 16 x GC threads per patch):
// gather, residual, assembly into registers, L D L^T factorisation, back
// substitution and evaluation at the member targets.
// blend_kernel sums the slots of each target in ascending patch order with
// the Shepard-normalised Wendland-C2 weights (solver.hpp:128-169, 275-285),
// deterministically (the reference uses atomic_add in arbitrary order).

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <type_traits>

#include "rbf_solve_common.hpp"

namespace pacman_amd {

namespace {

using namespace solve_detail;

// 1 / PU radius of patch k: per patch (adaptive patches) or one for all.
template <class A> __device__ __forceinline__ double rinv_of(const A &a, int k) {
  return a.patchRinv ? a.patchRinv[k] : a.rinv;
}

// ---------------------------------------------------------------------------
// Polynomial fit
// ---------------------------------------------------------------------------

// Singular values (descending) of a p x p upper-triangular R by one-sided
// Jacobi; p <= 4. All register arrays are indexed by compile-time constants.
__device__ inline void small_svd(double (&U)[4][4], int p, double (&S)[4]) {
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = a + 1; b < 4; ++b) {
        if (b < p) {
          double al = 0, be = 0, ga = 0;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            al += U[i][a] * U[i][a];
            be += U[i][b] * U[i][b];
            ga += U[i][a] * U[i][b];
          }
          if (ga != 0.0) {
            const double rel = fabs(ga) / sqrt(al * be);
            off = fmax(off, rel);
            if (rel >= 1e-17) {
              const double zeta = (be - al) / (2.0 * ga);
              const double t = (zeta >= 0 ? 1.0 : -1.0) /
                               (fabs(zeta) + sqrt(1.0 + zeta * zeta));
              const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const double ua = U[i][a], ub = U[i][b];
                U[i][a] = cs * ua - sn * ub;
                U[i][b] = sn * ua + cs * ub;
              }
            }
          }
        }
      }
    if (off < 1e-16)
      break;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double s = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      s += U[i][a] * U[i][a];
    S[a] = (a < p) ? sqrt(s) : -1.0;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)  // sort descending (unused entries are -1)
#pragma unroll
    for (int b = a + 1; b < 4; ++b)
      if (S[b] > S[a]) {
        const double t = S[a];
        S[a] = S[b];
        S[b] = t;
      }
}

__device__ inline double pick4(const double (&G)[4][4], int r, int c) {
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      v = (i == r && j == c) ? G[i][j] : v;
  return v;
}

// Column c (0..3) of the active basis holds axis ca[c] (ca[0] = -1: the
// constant column).
__device__ inline int assign_columns(const bool (&active)[3], int (&ca)[4]) {
  int cols = 1;
  ca[0] = -1;
  ca[1] = ca[2] = ca[3] = -2;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (active[ax]) {
      if (cols == 1)
        ca[1] = ax;
      else if (cols == 2)
        ca[2] = ax;
      else
        ca[3] = ax;
      ++cols;
    }
  return cols;
}

// In-place Cholesky of the leading p x p block (lower, L L^T). Returns false
// when a pivot is not positive (rank-deficient basis).
__device__ inline bool chol4(double (&L)[4][4], int p) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j < p) {
      double d = L[j][j];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < j)
          d -= L[j][k] * L[j][k];
      ok = ok && (d > 0.0);
      const double ljj = sqrt(fmax(d, 0.0));
      L[j][j] = ljj;
      const double inv = ljj > 0.0 ? 1.0 / ljj : 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i > j && i < p) {
          double s = L[i][j];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (k < j)
              s -= L[i][k] * L[j][k];
          L[i][j] = s * inv;
        }
      }
    }
  }
  return ok;
}

// cond(raw Q) > 1e5 for the active columns? L is the Cholesky factor of the
// scaled Gram matrix, c the centre, sc = 1 / r.
__device__ inline bool ill_conditioned(const double (&L)[4][4], int pv,
                                       const int (&ca)[4], const double (&c)[3],
                                       double sc) {
  // raw basis [1, x_a] = [1, u_a] T, T_00 = 1, T_0j = c_j, T_jj = 1 / sc;
  // R = R'' T with R'' = L^T (upper triangular)
  double Rr[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double v = 0.0;
      if (i < pv && j < pv && i <= j) {
        double cj = 0.0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
          cj = (ca[j] == ax) ? c[ax] : cj;
        if (j == 0)
          v = L[0][0];
        else if (i == 0)
          v = L[0][0] * cj + L[j][0] / sc;
        else
          v = L[j][i] / sc;
      }
      Rr[i][j] = v;
    }
  // ||R||_F ||R^-1||_F >= cond_2(R): settles cond <= 1e5 cheaply; R^-1 is
  // accumulated one column at a time.
  double fr = 0.0, fi = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      fr += Rr[i][j] * Rr[i][j];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double col[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 3; i >= 0; --i) {
      if (i <= j && j < pv) {
        double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k > i && k <= j)
            s -= Rr[i][k] * col[k];
        col[i] = s / Rr[i][i];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      fi += col[i] * col[i];
  }
  if (sqrt(fr) * sqrt(fi) <= kCondMax)
    return false;
  double S[4];
  small_svd(Rr, pv, S);
  const double smin = pv == 4 ? S[3] : pv == 3 ? S[2] : S[1];
  return S[0] / fmax(smin, kEps) > kCondMax;
}

// Active polynomial columns of a patch and the Cholesky factor L of their
// scaled Gram matrix (one wave, all lanes get the same result): columns
// ca[0..pv), ca[0] = -1 for the constant, else the axis; the conditioning
// loop of SerialFillQ (serial_ops.hpp:108-139).
__device__ inline void poly_basis(const SolveArgs &a, int patch, int lane,
                                  const double (&c)[3], double sc, int &pv,
                                  int (&ca)[4], double (&L)[4][4]) {
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int dim = a.dim;
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double G[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      G[i][j] = 0.0;
  for (int t = lane; t < n; t += 64) {
    const int p = a.srcIdx[s0 + t];
    const double x[3] = {a.sx[p], a.sy[p], a.sz[p]};
    double u[4] = {1.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      mn[ax] = fmin(mn[ax], x[ax]);
      mx[ax] = fmax(mx[ax], x[ax]);
      u[ax + 1] = (x[ax] - c[ax]) * sc;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j <= i; ++j)
        G[i][j] += u[i] * u[j];
  }
  double range[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    range[ax] = fabs(wave_max(mx[ax]) - wave_min(mn[ax]));
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      G[i][j] = wave_sum(G[i][j]);
      G[j][i] = G[i][j];
    }
  bool active[3] = {dim > 0, dim > 1, dim > 2};
  pv = assign_columns(active, ca);
  for (;;) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        L[i][j] = (i < pv && j < pv) ? pick4(G, ca[i] + 1, ca[j] + 1)
                                     : (i == j ? 1.0 : 0.0);
    const bool ok = chol4(L, pv);
    if (pv <= 1)
      break;
    if (ok && !ill_conditioned(L, pv, ca, c, sc))
      break;
    double best = DBL_MAX;
    int bi = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double rg = (ax < dim && active[ax]) ? range[ax] : DBL_MAX;
      if (rg < best) {
        best = rg;
        bi = ax;
      }
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      if (ax == bi)
        active[ax] = false;
    pv = assign_columns(active, ca);
  }
}

// One wave per patch; coef[(patch * F + f) * 4 + k], pvOut[patch]. NF = 1
// for one field, kFMax for 2..kFMax (the per-field sums of the wider
// variant then fit three waves per SIMD only with the bound below).
template <int NF>
__global__ __launch_bounds__(256, NF > 1 ? 3 : 1) void poly_kernel(
    SolveArgs a, int K, double *coef) {
  const int patch = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (patch >= K)
    return;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F;
  const double c[3] = {a.cx[patch], a.cy[patch], a.cz[patch]};
  const double sc = a.patchRinv ? a.patchRinv[patch] : a.rinv;
  int pv, ca[4];
  double L[4][4];
  poly_basis(a, patch, lane, c, sc, pv, ca, L);
  // beta'' = G^-1 Q''^T b on the active columns (normal equations of the
  // well-conditioned scaled basis); coefficients returned around the centre.
  // Q''^T b of every field in one pass over the sources: each 64-source step
  // loads the coordinates and all F values together (one dependent round
  // trip per step, not one per field); a field's sums keep the source order.
  double hfa[NF][4];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      hfa[f][k] = 0.0;
  for (int t = lane; t < n; t += 64) {
    const int p = a.srcIdx[s0 + t];
    double b[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
      b[f] = f < F ? a.vals[(int64_t)f * a.N + p] : 0.0;
    const double dx = (a.sx[p] - c[0]) * sc, dy = (a.sy[p] - c[1]) * sc,
                 dz = (a.sz[p] - c[2]) * sc;
#pragma unroll
    for (int f = 0; f < NF; ++f)
      if (f < F) {
        hfa[f][0] += b[f];
        hfa[f][1] += dx * b[f];
        hfa[f][2] += dy * b[f];
        hfa[f][3] += dz * b[f];
      }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (f >= F)
      break;
    double hf[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      hf[k] = wave_sum(hfa[f][k]);
    double y[4], beta[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double hv = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        hv = (ca[i] + 1 == k) ? hf[k] : hv;
      y[i] = hv;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < pv) {
        double s = y[i];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k < i)
            s -= L[i][k] * y[k];
        y[i] = L[i][i] > 0.0 ? s / L[i][i] : 0.0;
      }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
      beta[i] = 0.0;
      if (i < pv) {
        double s = y[i];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k > i && k < pv)
            s -= L[k][i] * beta[k];
        beta[i] = L[i][i] > 0.0 ? s / L[i][i] : 0.0;
      }
    }
    if (lane == 0) {
      double cax[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        cax[ax] = ca[1] == ax ? beta[1]
                  : ca[2] == ax ? beta[2]
                  : ca[3] == ax ? beta[3]
                                : 0.0;
      double *o = coef + ((int64_t)patch * F + f) * 4;
      o[0] = beta[0];
      o[1] = cax[0] * sc;
      o[2] = cax[1] * sc;
      o[3] = cax[2] * sc;
    }
  }
  if (lane == 0 && a.pvOut)
    a.pvOut[patch] = pv;
}

// Shepard-normalised PU blend, one thread per target (cell order).
__global__ __launch_bounds__(256) void blend_kernel(BlendArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M)
    return;
  const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
  const double yx = a.tx[t], yy = a.ty[t], yz = a.tz[t];
  // the first kSlots kernel values are kept for the weighted sum (static
  // indices: registers); slots past them are evaluated again
  constexpr int kSlots = 8;
  double ph[kSlots];
  double sum = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    ph[j] = 0.0;
    if (s0 + j < s1) {
      const int k = a.slotPatch[s0 + j];
      // DistanceNoCheck(target, center), solver.hpp:158-161
      const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
      ph[j] = wendland<1, double, true>(d, rinv_of(a, k));
      sum += ph[j];
    }
  }
  for (int64_t s = s0 + kSlots; s < s1; ++s) {
    const int k = a.slotPatch[s];
    const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
    sum += wendland<1, double, true>(d, rinv_of(a, k));
  }
  const int64_t o = a.tperm[t];
  constexpr int kB = 8;
  for (int f0 = 0; f0 < a.F; f0 += kB) {
    double acc[kB];
#pragma unroll
    for (int f = 0; f < kB; ++f)
      acc[f] = 0.0;
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      if (s0 + j < s1 && a.slotMember[s0 + j]) {
        const double w = ph[j] / sum;
#pragma unroll
        for (int f = 0; f < kB; ++f)
          if (f0 + f < a.F)
            acc[f] = fma(w, a.contrib[(s0 + j) * a.F + f0 + f], acc[f]);
      }
    }
    for (int64_t s = s0 + kSlots; s < s1; ++s) {
      if (!a.slotMember[s])
        continue;
      const int k = a.slotPatch[s];
      const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
      const double w = wendland<1, double, true>(d, rinv_of(a, k)) / sum;
#pragma unroll
      for (int f = 0; f < kB; ++f)
        if (f0 + f < a.F)
          acc[f] = fma(w, a.contrib[s * a.F + f0 + f], acc[f]);
    }
#pragma unroll
    for (int f = 0; f < kB; ++f)
      if (f0 + f < a.F)
        a.out[(int64_t)(f0 + f) * a.M + o] = acc[f];
  }
}

// The same blend for 2 .. 8 fields with the plan's stored slot weights
// (slot_weights_kernel: the same Shepard sums and quotients as blend_kernel,
// 0 for non-member slots), 64 consecutive original targets per workgroup: a
// group of 8 lanes takes one target at a time (lane f, field f), reads each
// member slot's fields as one contiguous piece with no dependent loads, and
// the results are staged in LDS and written field by field as 64 consecutive
// outputs instead of F scattered writes per target.
__global__ __launch_bounds__(256) void blend_tile_kernel(BlendArgs a) {
  constexpr int kT = 64, kL = 8, kGroups = 256 / kL;
  __shared__ double tile[kT][kL + 1];
  const int64_t o0 = (int64_t)blockIdx.x * kT;
  const int tid = threadIdx.x, f = tid % kL, grp = tid / kL;
  const int F = a.F;
  for (int j = grp; j < kT; j += kGroups) {
    const int64_t o = o0 + j;
    double acc = 0.0;
    if (o < a.M) {
      const int64_t t = a.tinv[o];
      const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
      // weight and value loaded unconditionally (no load waits on another);
      // member slots only (their weight is > 0)
#pragma unroll 4
      for (int64_t s = s0; s < s1; ++s) {
        const double w = a.wslot[s];
        const double c = a.contrib[s * F + (f < F ? f : 0)];
        acc = (w != 0.0 && f < F) ? fma(w, c, acc) : acc;
      }
    }
    tile[j][f] = acc;
  }
  __syncthreads();
  for (int x = tid; x < kT * F; x += 256) {
    const int ff = x / kT, j = x % kT;
    if (o0 + j < a.M)
      a.out[(int64_t)ff * a.M + o0 + j] = tile[j][ff];
  }
}

// out[f][p] = vals[f][perm[p]], field f = blockIdx.y. The workgroups are
// dispatched field by field, so the random reads of one field (8 bytes out of
// each line) work on that field's rows alone, which the L2 and the
// last-level cache hold; all fields at once would not fit them.
__global__ void gather_values_kernel(const double *__restrict__ vals,
                                     const int *__restrict__ perm, int64_t N,
                                     double *__restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int64_t fo = (int64_t)blockIdx.y * N;
  out[fo + p] = vals[fo + perm[p]];
}

// Adjoint of the polynomial fit (one wave per patch). The forward transfer
// of a patch is t = E A^-1 (I - Q G^-1 Q^T) b + V G^-1 Q^T b (G = Q^T Q on
// the active columns of the scaled basis); its transpose applied to
// h = w g (member targets) is, with gamma = A^-1 E^T h,
//   gamma + Q G^-1 (V^T h - Q^T gamma).
__global__ __launch_bounds__(256) void adjoint_poly_kernel(
    SolveArgs a, int K, const double *gamma, double *contribS) {
  const int patch = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (patch >= K)
    return;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int64_t p0 = a.pairOff[patch];
  const int mall = (int)(a.pairOff[patch + 1] - p0);
  const int F = a.F;
  const double c[3] = {a.cx[patch], a.cy[patch], a.cz[patch]};
  const double sc = a.patchRinv ? a.patchRinv[patch] : a.rinv;
  int pv, ca[4];
  double L[4][4];
  poly_basis(a, patch, lane, c, sc, pv, ca, L);
  for (int f = 0; f < F; ++f) {
    double r4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = lane; e < mall; e += 64) {
      const int pe = a.pairTgt[p0 + e];
      if (((unsigned)pe >> 31) == 0)
        continue;
      const int tp = pe & 0x7fffffff;
      const double ws = a.wslot[a.pairSlot[p0 + e]];
      const double h = a.opRow0 >= 0 ? (e == a.opRow0 + f ? ws : 0.0)
                                     : ws * a.gSorted[f * a.M + tp];
      r4[0] += h;
      r4[1] += (a.tx[tp] - c[0]) * sc * h;
      r4[2] += (a.ty[tp] - c[1]) * sc * h;
      r4[3] += (a.tz[tp] - c[2]) * sc * h;
    }
    for (int t = lane; t < n; t += 64) {
      const int p = a.srcIdx[s0 + t];
      const double gm = gamma[(s0 + t) * F + f];
      r4[0] -= gm;
      r4[1] -= (a.sx[p] - c[0]) * sc * gm;
      r4[2] -= (a.sy[p] - c[1]) * sc * gm;
      r4[3] -= (a.sz[p] - c[2]) * sc * gm;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      r4[k] = wave_sum(r4[k]);
    // z = G^-1 r on the active columns (L L^T z = r)
    double y[4], z[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double hv = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        hv = (ca[i] + 1 == k) ? r4[k] : hv;
      y[i] = hv;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < pv) {
        double t = y[i];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k < i)
            t -= L[i][k] * y[k];
        y[i] = L[i][i] > 0.0 ? t / L[i][i] : 0.0;
      }
#pragma unroll
    for (int i = 3; i >= 0; --i) {
      z[i] = 0.0;
      if (i < pv) {
        double t = y[i];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k > i && k < pv)
            t -= L[k][i] * z[k];
        z[i] = L[i][i] > 0.0 ? t / L[i][i] : 0.0;
      }
    }
    double cu[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        v = (i < pv && ca[i] + 1 == k) ? z[i] : v;
      cu[k] = v;
    }
    for (int t = lane; t < n; t += 64) {
      const int p = a.srcIdx[s0 + t];
      const double q = cu[0] + cu[1] * ((a.sx[p] - c[0]) * sc) +
                       cu[2] * ((a.sy[p] - c[1]) * sc) +
                       cu[3] * ((a.sz[p] - c[2]) * sc);
      contribS[(s0 + t) * F + f] = gamma[(s0 + t) * F + f] + q;
    }
  }
}

// Shepard weight of each member slot of a target (cell order), as the blend
// applies it; 0 for the slots that are not members.
__global__ __launch_bounds__(256) void slot_weights_kernel(BlendArgs a,
                                                           double *wslot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M)
    return;
  const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
  const double yx = a.tx[t], yy = a.ty[t], yz = a.tz[t];
  // as blend_kernel: the first kSlots kernel values stay in registers
  constexpr int kSlots = 8;
  double ph[kSlots];
  double sum = 0.0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    ph[j] = 0.0;
    if (s0 + j < s1) {
      const int k = a.slotPatch[s0 + j];
      const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
      ph[j] = wendland<1, double, true>(d, rinv_of(a, k));
      sum += ph[j];
    }
  }
  for (int64_t s = s0 + kSlots; s < s1; ++s) {
    const int k = a.slotPatch[s];
    const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
    sum += wendland<1, double, true>(d, rinv_of(a, k));
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j)
    if (s0 + j < s1)
      wslot[s0 + j] = a.slotMember[s0 + j] ? ph[j] / sum : 0.0;
  for (int64_t s = s0 + kSlots; s < s1; ++s) {
    const int k = a.slotPatch[s];
    const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
    wslot[s] = a.slotMember[s] ? wendland<1, double, true>(d, rinv_of(a, k)) / sum
                               : 0.0;
  }
}

__global__ void src_count_kernel(const int *srcIdx, int64_t total, int *cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total)
    atomicAdd(&cnt[srcIdx[e]], 1);
}

__global__ void src_fill_kernel(const int *srcIdx, int64_t total,
                                const int64_t *off, int *cursor, int *ent) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total)
    return;
  const int p = srcIdx[e];
  const int k = atomicAdd(&cursor[p], 1);
  ent[off[p] + k] = (int)e;
}

// Entries of each source in ascending order (= ascending patch).
__global__ void src_sort_kernel(const int64_t *off, int64_t N, int *ent) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int64_t b = off[p], e = off[p + 1];
  for (int64_t i = b + 1; i < e; ++i) {
    const int v = ent[i];
    int64_t j = i - 1;
    while (j >= b && ent[j] > v) {
      ent[j + 1] = ent[j];
      --j;
    }
    ent[j + 1] = v;
  }
}

__global__ void src_reduce_kernel(const int64_t *off, const int *ent,
                                  int64_t N, int F, const double *contribS,
                                  const int *perm, double *out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int64_t b = off[p], e = off[p + 1];
  const int64_t o = perm[p];
  for (int f = 0; f < F; ++f) {
    double acc = 0.0;
    for (int64_t i = b; i < e; ++i)
      acc += contribS[(int64_t)ent[i] * F + f];
    out[(int64_t)f * N + o] = acc;
  }
}

// Evaluation of one chunk of up to 64 PPL (target, patch) pairs of a patch
// whose solve stored its alpha, by one wave (src/rbf_pum/solver/solver.hpp:
// 265-273): lane l holds pairs q0 + l + 64 h, h < PPL (the staged reads
// below serve all of them; from PPL = 2 on the VALU, not the LDS, sets the
// pace), sums over the patch's sources in source order. The wave stages
// its patch's sources kBlk at a time in its own LDS region (coalesced
// vector loads of the slot records: x y z 0 from slotXyz, NF alpha from
// alphaHeavy, zero past F), loading the next block while it computes this
// one; no workgroup barriers, so the waves of many patches stay in flight.
// NF = alpha_stride(F) products per source (1, 4 or 8); the phi code is
// the same for every NF, so a field's result does not depend on the field
// count.
constexpr int kBlk = 32;

template <int K, int NF, int PPL>
__global__ __launch_bounds__(256) void wave_eval_kernel(
    SolveArgs a, const double *__restrict__ coef,
    const int2 *__restrict__ chunks, int nChunks) {
  __shared__ double sxyz[4][kBlk][4];
  __shared__ double sal[4][kBlk * NF];
  // alpha doubles per lane and block (lanes below kBlk * NF / VW load)
  constexpr int VW = kBlk * NF >= 64 ? kBlk * NF / 64 : 1;
  constexpr int VL = kBlk * NF / VW;
  const int wv = threadIdx.x >> 6;
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + wv));
  if (w >= nChunks)
    return;  // no workgroup barriers below
  const int lane = threadIdx.x & 63;
  const int2 ch = chunks[w];
  const int patch = ch.x, q0 = ch.y;
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int64_t p0 = a.pairOff[patch];
  const int q1 = min((int)(a.pairOff[patch + 1] - p0), q0 + 64 * PPL);
  const int F = a.F;
  const double dinv = a.patchDinv ? a.patchDinv[patch] : a.dinv;
  const double *__restrict__ al = a.alphaHeavy + 8 * s0;
  const double *__restrict__ xyz = a.slotXyz + 4 * s0;
  bool mem[PPL];
  double yx[PPL], yy[PPL], yz[PPL];
  int slot[PPL];
#pragma unroll
  for (int h = 0; h < PPL; ++h) {
    const int q = q0 + lane + 64 * h;
    mem[h] = false;
    yx[h] = yy[h] = yz[h] = 0.0;
    slot[h] = 0;
    if (q < q1) {
      const int e = a.pairTgt[p0 + q];
      mem[h] = ((unsigned)e >> 31) != 0;
      const int tp = e & 0x7fffffff;
      yx[h] = a.tx[tp];
      yy[h] = a.ty[tp];
      yz[h] = a.tz[tp];
      slot[h] = a.pairSlot[p0 + q];
    }
  }
  // staged block: lane l < kBlk holds source b0 + l's x y z 0, lane l < VL
  // VW doubles of the block's alpha records (NF per source)
  double4 cxyz = {0.0, 0.0, 0.0, 0.0};
  double cv[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v)
    cv[v] = 0.0;
  auto load = [&](int b0) {
    const int i = b0 + lane;
    if (lane < kBlk && i < n)
      cxyz = *reinterpret_cast<const double4 *>(xyz + 4 * (int64_t)i);
    const int e = NF * b0 + VW * lane;
    if (lane < VL && e < NF * n) {
      if constexpr (VW == 4) {
        const double4 t = *reinterpret_cast<const double4 *>(al + e);
        cv[0] = t.x;
        cv[1] = t.y;
        cv[2] = t.z;
        cv[3] = t.w;
      } else if constexpr (VW == 2) {
        const double2 t = *reinterpret_cast<const double2 *>(al + e);
        cv[0] = t.x;
        cv[1] = t.y;
      } else {
        cv[0] = al[e];
      }
    }
  };
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  double acc[PPL][NF];
#pragma unroll
  for (int h = 0; h < PPL; ++h)
#pragma unroll
    for (int f = 0; f < NF; ++f)
      acc[h][f] = 0.0;
  load(0);
  for (int b0 = 0; b0 < n; b0 += kBlk) {
    if (lane < kBlk)
      *reinterpret_cast<double4 *>(&sxyz[wv][lane][0]) = cxyz;
    if (lane < VL)
#pragma unroll
      for (int v = 0; v < VW; ++v)
        sal[wv][VW * lane + v] = cv[v];
    wave_sync();
    if (b0 + kBlk < n)
      load(b0 + kBlk);
    const int bn = min(kBlk, n - b0);
    for (int j = 0; j < bn; ++j) {
      const double x = sxyz[wv][j][0], y = sxyz[wv][j][1],
                   z = sxyz[wv][j][2];
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        const double ph =
            wendland<K, double, true>(dist_fast(yx[h], yy[h], yz[h], x, y, z),
                                      dinv);
#pragma unroll
        for (int f = 0; f < NF; ++f)
          acc[h][f] = fma(ph, sal[wv][j * NF + f], acc[h][f]);
      }
    }
    wave_sync();  // this block read before the next one is written
  }
  const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];
#pragma unroll
  for (int h = 0; h < PPL; ++h) {
    if (!mem[h])
      continue;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (f >= F)
        break;
      const double *c = coef + ((int64_t)patch * F + f) * 4;
      const double vb = c[0] + c[1] * (yx[h] - cx) + c[2] * (yy[h] - cy) +
                        c[3] * (yz[h] - cz);
      a.contrib[(int64_t)slot[h] * F + f] = acc[h][f] + vb;
    }
  }
}

// slotXyz[4 i + d] = coordinate d of the source in member slot i (the
// fourth entry 0): the sources of a patch in slot order, contiguous.
__global__ void slot_xyz_kernel(const int *srcIdx, int64_t total,
                                const double *sx, const double *sy,
                                const double *sz, double *xyz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total)
    return;
  const int p = srcIdx[i];
  xyz[4 * i] = sx[p];
  xyz[4 * i + 1] = sy[p];
  xyz[4 * i + 2] = sz ? sz[p] : 0.0;
  xyz[4 * i + 3] = 0.0;
}

__global__ void zero_i32_kernel(int *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0;
}

} // namespace

int solve_fields_per_launch() { return kFMax; }

void launch_wave_eval(int kind, const SolveArgs &a, const double *coef,
                      const int2 *chunks, int nChunks, int pairsPerLane,
                      hipStream_t s) {
  if (nChunks <= 0)
    return;
  if (a.slotXyz == nullptr)
    throw std::runtime_error("wave evaluation: slot coordinates missing");
  const unsigned g = (unsigned)((nChunks + 3) / 4);
  if (pairsPerLane < 1 || pairsPerLane > 3)
    throw std::runtime_error("wave evaluation: 1 to 3 pairs per lane");
  auto go = [&](auto nf, auto ppl) {
    constexpr int NF = decltype(nf)::value, P = decltype(ppl)::value;
    switch (kind) {
    case 0: wave_eval_kernel<0, NF, P><<<g, 256, 0, s>>>(a, coef, chunks, nChunks); break;
    case 1: wave_eval_kernel<1, NF, P><<<g, 256, 0, s>>>(a, coef, chunks, nChunks); break;
    case 2: wave_eval_kernel<2, NF, P><<<g, 256, 0, s>>>(a, coef, chunks, nChunks); break;
    case 3: wave_eval_kernel<3, NF, P><<<g, 256, 0, s>>>(a, coef, chunks, nChunks); break;
    default: wave_eval_kernel<4, NF, P><<<g, 256, 0, s>>>(a, coef, chunks, nChunks); break;
    }
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I8 = std::integral_constant<int, 8>;
  auto byPpl = [&](auto nf) {
    if (pairsPerLane == 1)
      go(nf, I1{});
    else if (pairsPerLane == 2)
      go(nf, I2{});
    else
      go(nf, I3{});
  };
  const int S = alpha_stride(a.F);
  if (S == 8)
    byPpl(I8{});
  else if (S == 4)
    byPpl(I4{});
  else
    byPpl(std::integral_constant<int, 1>{});
  PACMAN_HIP_CHECK(hipGetLastError());
}

// Evaluation chunks of patch k: u = ceil(pairs / 64) units of 64 pairs ->
// one chunk of one unit (class 0), or chunks of two units (class 1) and, for
// an odd u, a last one of three (class 2); patches above tileMaxN sources are
// group 1, the others group 0. cnt holds the six (group, class) counts of
// every patch, class-major: cnt[(3 g + i) K + k].
__device__ inline void eval_chunk_shape(const int64_t *srcOff,
                                        const int64_t *pairOff, int64_t k,
                                        int tileMaxN, int &g, int64_t &u) {
  g = srcOff[k + 1] - srcOff[k] > tileMaxN ? 1 : 0;
  u = (pairOff[k + 1] - pairOff[k] + 63) / 64;
}

__global__ void eval_chunk_count_kernel(const int64_t *srcOff,
                                        const int64_t *pairOff, int64_t K,
                                        int tileMaxN, int *cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K)
    return;
  int g;
  int64_t u;
  eval_chunk_shape(srcOff, pairOff, k, tileMaxN, g, u);
  const int c[3] = {u == 1 ? 1 : 0,
                    u > 1 ? (int)((u & 1) ? (u - 3) / 2 : u / 2) : 0,
                    u > 1 ? (int)(u & 1) : 0};
  for (int gg = 0; gg < 2; ++gg)
    for (int i = 0; i < 3; ++i)
      cnt[(3 * gg + i) * K + k] = gg == g ? c[i] : 0;
}

// The six class boundaries of the scanned counts (pos has 6 K + 1 entries).
__global__ void eval_chunk_bounds_kernel(const int64_t *pos, int64_t K,
                                         int64_t *bounds) {
  const int c = threadIdx.x;
  if (c <= 6)
    bounds[c] = pos[c * K];
}

// Patch k's chunks at its scanned positions, first unit first.
__global__ void eval_chunk_fill_kernel(const int64_t *srcOff,
                                       const int64_t *pairOff, int64_t K,
                                       int tileMaxN, const int64_t *pos,
                                       int2 *chunks) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K)
    return;
  int g;
  int64_t u;
  eval_chunk_shape(srcOff, pairOff, k, tileMaxN, g, u);
  if (u == 1) {
    chunks[pos[(3 * g) * K + k]] = make_int2((int)k, 0);
    return;
  }
  int64_t p2 = pos[(3 * g + 1) * K + k], p3 = pos[(3 * g + 2) * K + k];
  int64_t q = 0;
  for (int64_t left = u; left > 0;) {
    const int take = left == 3 ? 3 : 2;
    chunks[take == 2 ? p2++ : p3++] = make_int2((int)k, (int)q);
    q += 64 * take;
    left -= take;
  }
}

void plan_eval_chunks(const int64_t *srcOff, const int64_t *pairOff, int64_t K,
                      int tileMaxN, DevBuf<int2> &chunks, int64_t first[2][3],
                      int count[2][3], hipStream_t s) {
  for (int g = 0; g < 2; ++g)
    for (int i = 0; i < 3; ++i) {
      first[g][i] = 0;
      count[g][i] = 0;
    }
  if (K <= 0)
    return;
  DevBuf<int> cnt((size_t)6 * K);
  DevBuf<int64_t> pos((size_t)6 * K + 1), bounds(7);
  const unsigned g = (unsigned)((K + 255) / 256);
  eval_chunk_count_kernel<<<g, 256, 0, s>>>(srcOff, pairOff, K, tileMaxN,
                                            cnt.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  exclusive_scan_i32_to_i64_async(cnt.get(), pos.get(), 6 * K, s);
  eval_chunk_bounds_kernel<<<1, 64, 0, s>>>(pos.get(), K, bounds.get());
  PACMAN_HIP_CHECK(hipGetLastError());
  int64_t b[7];
  copy_to_host(b, bounds.get(), sizeof(b), s);
  for (int c = 0; c < 6; ++c) {
    if (b[c + 1] - b[c] > INT_MAX)
      throw std::runtime_error("evaluation: too many chunks in one class");
    first[c / 3][c % 3] = b[c];
    count[c / 3][c % 3] = (int)(b[c + 1] - b[c]);
  }
  chunks.alloc((size_t)std::max<int64_t>(b[6], 1));
  eval_chunk_fill_kernel<<<g, 256, 0, s>>>(srcOff, pairOff, K, tileMaxN,
                                           pos.get(), chunks.get());
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_slot_xyz(const int *srcIdx, int64_t total, const double *sx,
                     const double *sy, const double *sz, double *xyz,
                     hipStream_t s) {
  if (total <= 0)
    return;
  slot_xyz_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
      srcIdx, total, sx, sy, sz, xyz);
  PACMAN_HIP_CHECK(hipGetLastError());
}


void launch_poly(const SolveArgs &a, int64_t K, double *coef, hipStream_t s) {
  if (K <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("poly: 1 <= fields <= 8 per launch");
  if (a.F == 1)
    poly_kernel<1><<<(unsigned)((K + 3) / 4), 256, 0, s>>>(a, (int)K, coef);
  else
    poly_kernel<kFMax><<<(unsigned)((K + 3) / 4), 256, 0, s>>>(a, (int)K,
                                                                coef);
  PACMAN_HIP_CHECK(hipGetLastError());
}

int solve_bin_rows(int n, int F) { return (n + F + 15) / 16; }

bool solve_tile_variant(int R, int waves) {
  return (waves == 8 && R >= 1 && R <= 13) ||
         (waves == 16 && R >= 11 && R <= kTileMaxR);
}

void launch_solve_f64(int R, int waves, int kind, const SolveArgs &a,
                      const double *coef, hipStream_t s) {
  if (a.count <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("solve: 1 <= fields <= 8 per launch");
  if (kind < 0 || kind > 4)
    throw std::runtime_error("solve: unknown Wendland kind");
  if (!solve_tile_variant(R, waves))
    throw std::runtime_error("solve: unsupported size bin");
  if (a.alphaHeavy == nullptr)
    throw std::runtime_error("solve: alpha records (the evaluation runs "
                             "apart) missing");
  if (waves == 16) {
    switch (R) {
    case 11: launch_solve_tile<11, 16>(kind, a, coef, s); break;
    case 12: launch_solve_tile<12, 16>(kind, a, coef, s); break;
    case 13: launch_solve_tile<13, 16>(kind, a, coef, s); break;
    case 14: launch_solve_tile<14, 16>(kind, a, coef, s); break;
    case 15: launch_solve_tile<15, 16>(kind, a, coef, s); break;
    case 16: launch_solve_tile<16, 16>(kind, a, coef, s); break;
    default: launch_solve_tile<17, 16>(kind, a, coef, s); break;
    }
    return;
  }
  switch (R) {
  case 1: launch_solve_tile<1, 8>(kind, a, coef, s); break;
  case 2: launch_solve_tile<2, 8>(kind, a, coef, s); break;
  case 3: launch_solve_tile<3, 8>(kind, a, coef, s); break;
  case 4: launch_solve_tile<4, 8>(kind, a, coef, s); break;
  case 5: launch_solve_tile<5, 8>(kind, a, coef, s); break;
  case 6: launch_solve_tile<6, 8>(kind, a, coef, s); break;
  case 7: launch_solve_tile<7, 8>(kind, a, coef, s); break;
  case 8: launch_solve_tile<8, 8>(kind, a, coef, s); break;
  case 9: launch_solve_tile<9, 8>(kind, a, coef, s); break;
  case 10: launch_solve_tile<10, 8>(kind, a, coef, s); break;
  case 11: launch_solve_tile<11, 8>(kind, a, coef, s); break;
  case 12: launch_solve_tile<12, 8>(kind, a, coef, s); break;
  default: launch_solve_tile<13, 8>(kind, a, coef, s); break;
  }
}

void launch_solve_f64_adjoint(int R, int kind, const SolveArgs &a,
                              hipStream_t s) {
  if (a.count <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("solve: 1 <= fields <= 8 per launch");
  if (kind < 0 || kind > 4)
    throw std::runtime_error("solve: unknown Wendland kind");
  switch (R) {
  case 1: launch_solve_tile_adjoint<1>(kind, a, s); break;
  case 2: launch_solve_tile_adjoint<2>(kind, a, s); break;
  case 3: launch_solve_tile_adjoint<3>(kind, a, s); break;
  case 4: launch_solve_tile_adjoint<4>(kind, a, s); break;
  case 5: launch_solve_tile_adjoint<5>(kind, a, s); break;
  case 6: launch_solve_tile_adjoint<6>(kind, a, s); break;
  case 7: launch_solve_tile_adjoint<7>(kind, a, s); break;
  case 8: launch_solve_tile_adjoint<8>(kind, a, s); break;
  case 9: launch_solve_tile_adjoint<9>(kind, a, s); break;
  case 10: launch_solve_tile_adjoint<10>(kind, a, s); break;
  case 11: launch_solve_tile_adjoint<11>(kind, a, s); break;
  case 12: launch_solve_tile_adjoint<12>(kind, a, s); break;
  case 13: launch_solve_tile_adjoint<13>(kind, a, s); break;
  default:
    throw std::runtime_error("solve: unsupported size bin");
  }
}

void launch_adjoint_poly(const SolveArgs &a, int64_t K, const double *gamma,
                         double *contribS, hipStream_t s) {
  if (K <= 0)
    return;
  adjoint_poly_kernel<<<(unsigned)((K + 3) / 4), 256, 0, s>>>(a, (int)K, gamma,
                                                               contribS);
  PACMAN_HIP_CHECK(hipGetLastError());
}

template <class T>
__global__ __launch_bounds__(256) void op_store_kernel(
    SolveArgs a, int K, const double *contribS, const int64_t *opOff, T *op) {
  const int patch = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (patch >= K)
    return;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int mall = (int)(a.pairOff[patch + 1] - a.pairOff[patch]);
  for (int f = 0; f < a.F; ++f) {
    const int q = a.opRow0 + f;
    if (q >= mall)
      break;
    T *row = op + opOff[patch] + (int64_t)q * n;
    for (int e = lane; e < n; e += 64)
      row[e] = (T)contribS[(s0 + e) * a.F + f];
  }
}

// One workgroup per patch: the patch's source values staged in LDS (zero
// past the patch's n sources); 16 lanes per operator row, lane l reading
// columns l, l + 16, ... (each row read as contiguous 128-byte pieces), all
// of a row's loads and its pair's target and slot issued before the first
// use; partial sums over the lanes added by a fixed butterfly.
constexpr int kOpBlock = 256;
template <int FM, class T>
__global__ __launch_bounds__(kOpBlock) __attribute__((amdgpu_waves_per_eu(
    FM <= 2 ? 4 : 3))) void op_apply_kernel(
    SolveArgs a, int K, const int64_t *opOff, const T *op) {
  constexpr int kCols = kMaxPatch / 16;  // loads per lane and row
  __shared__ double fv[FM][kMaxPatch];
  const int patch = xcd_remap(blockIdx.x, K);
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int64_t p0 = a.pairOff[patch];
  const int mall = (int)(a.pairOff[patch + 1] - p0);
  const int F = a.F;
  const int tid = threadIdx.x;
  if (n == 0 || mall == 0)  // the whole block: nothing to read or write
    return;
  // fields F .. FM - 1 and columns past n are staged as zeros
  for (int e = tid; e < kMaxPatch; e += kOpBlock) {
    const int p = e < n ? a.srcIdx[s0 + e] : 0;
#pragma unroll
    for (int f = 0; f < FM; ++f)
      fv[f][e] = (e < n && f < F) ? a.vals[(int64_t)f * a.N + p] : 0.0;
  }
  __syncthreads();
  const int lane = tid & 63, g = lane >> 4, l16 = lane & 15;
  constexpr int kRowsPerPass = kOpBlock / 16;
  const T *base = op + opOff[patch];
  for (int q0 = (tid >> 6) * 4; q0 < mall; q0 += kRowsPerPass) {
    const int q = q0 + g;
    const int qc = min(q, mall - 1);
    const int pe = a.pairTgt[p0 + qc];
    const int slot = a.pairSlot[p0 + qc];
    // columns past n read the row's last entry; their staged value is 0
    const T *row = base + (int64_t)qc * n;
    double acc[FM];
#pragma unroll
    for (int f = 0; f < FM; ++f)
      acc[f] = 0.0;
    {
      // every load of the row in flight at once
      T raw[kCols];
#pragma unroll
      for (int j = 0; j < kCols; ++j) {
        raw[j] = (T)0;
        if (16 * j < n)  // uniform over the block
          raw[j] = row[min(l16 + 16 * j, n - 1)];
      }
#pragma unroll
      for (int j = 0; j < kCols; ++j) {
        const double gv = (double)raw[j];
#pragma unroll
        for (int f = 0; f < FM; ++f)
          acc[f] = fma(gv, fv[f][l16 + 16 * j], acc[f]);
      }
    }
#pragma unroll
    for (int f = 0; f < FM; ++f)
      if (f < F) {
        double v = acc[f];
        v += __shfl_xor(v, 8, 16);
        v += __shfl_xor(v, 4, 16);
        v += __shfl_xor(v, 2, 16);
        v += __shfl_xor(v, 1, 16);
        acc[f] = v;
      }
    if (q < mall && ((unsigned)pe >> 31) != 0 && l16 == 0) {
#pragma unroll
      for (int f = 0; f < FM; ++f)
        if (f < F)
          a.contrib[(int64_t)slot * F + f] = acc[f];
    }
  }
}

// MFMA form of the stored-operator apply: per patch, Out (rows x fields) =
// G (rows x n) . V (n x fields), in 16 x 16 output tiles with FP64
// 16x16x4 MFMAs. A wave takes a 16-row tile; lane (i + 16 k) loads the
// 32-byte run G[i][c + 4k .. c + 4k + 3] of each 16-column chunk c, and step
// s of the chunk multiplies column c + 4k + s of row i by the staged values
// of the same column (the K dimension visits the chunk's columns in the
// order k-major within each step; every column is used once). FT tiles of
// 16 fields per pass (up to 32 fields).
typedef double opd4 __attribute__((ext_vector_type(4)));
template <int FT, class T>
__global__ __launch_bounds__(kOpBlock) void op_apply_mfma_kernel(
    SolveArgs a, int K, const int64_t *opOff, const T *op) {
  constexpr int FW = 16 * FT;
  __shared__ double fv[kMaxPatch][FW];
  const int patch = xcd_remap(blockIdx.x, K);
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int64_t p0 = a.pairOff[patch];
  const int mall = (int)(a.pairOff[patch + 1] - p0);
  const int F = a.F;
  const int tid = threadIdx.x;
  if (n == 0 || mall == 0)  // the whole block: nothing to read or write
    return;
  // rows n .. (n rounded up to 32) - 1 are zero: the B operand needs no
  // column mask
  const int nr = (n + 31) & ~31;
  for (int x = tid; x < nr * FW; x += kOpBlock) {
    const int e = x / FW, f = x % FW;
    fv[e][f] = (f < F && e < n)
                   ? a.vals[(int64_t)f * a.N + a.srcIdx[s0 + e]]
                   : 0.0;
  }
  __syncthreads();
  const int lane = tid & 63, i = lane & 15, k = lane >> 4;
  const T *base = op + opOff[patch];
  for (int rt = (tid >> 6); rt * 16 < mall; rt += kOpBlock / 64) {
    const int r = rt * 16 + i;
    const bool rok = r < mall;
    const T *row = base + (int64_t)(rok ? r : 0) * n;
    // two accumulator sets over alternate 16-column chunks: independent
    // MFMA chains, added once at the end
    opd4 acc[FT], acc2[FT];
#pragma unroll
    for (int t = 0; t < FT; ++t) {
      acc[t] = opd4{0.0, 0.0, 0.0, 0.0};
      acc2[t] = opd4{0.0, 0.0, 0.0, 0.0};
    }
    // Eight operator loads per 32-column chunk, issued unconditionally: the
    // column is clamped into the row (its B entry is zero, so the finite
    // value read there adds nothing) and rows past the patch's read row 0
    // (their outputs are not stored). The next chunk's loads are issued
    // before the current chunk's MFMAs.
    // Loads run two chunks ahead (ra: even chunks, rb: odd chunks).
    auto load_chunk = [&](int c, T *r8) {
#pragma unroll
      for (int s = 0; s < 8; ++s)
        r8[s] = row[min(c + 16 * (s >> 2) + 4 * k + (s & 3), n - 1)];
    };
    auto mfma_chunk = [&](int c, const T *r8) {
      double g[8];
#pragma unroll
      for (int s = 0; s < 8; ++s)
        g[s] = (double)r8[s];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int col0 = c + 4 * k + s, col1 = col0 + 16;
#pragma unroll
        for (int t = 0; t < FT; ++t) {
          const double b0 = fv[col0][16 * t + i];
          const double b1 = fv[col1][16 * t + i];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(g[s], b0, acc[t], 0, 0,
                                                        0);
          acc2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(g[4 + s], b1,
                                                         acc2[t], 0, 0, 0);
        }
      }
    };
    T ra[8], rb[8];
    load_chunk(0, ra);
    load_chunk(32, rb);
    // the output rows of this lane (l / 16 + 4 v): target flag and slot,
    // loaded before the MFMAs; slot -1 for rows not stored
    int oslot[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = rt * 16 + k + 4 * v;
      const int rc = min(rr, mall - 1);
      const int pe = a.pairTgt[p0 + rc];
      const int sl = a.pairSlot[p0 + rc];
      oslot[v] = (rr < mall && ((unsigned)pe >> 31) != 0) ? sl : -1;
    }
    for (int c = 0; c < n; c += 64) {
      // chunk c (c < n), then chunk c + 32 if it starts inside the patch
      // (staged rows up to n rounded up to 32 are zero)
      T cur[8];
#pragma unroll
      for (int s = 0; s < 8; ++s)
        cur[s] = ra[s];
      if (c + 64 < n)
        load_chunk(c + 64, ra);
      mfma_chunk(c, cur);
      if (c + 32 < n) {
#pragma unroll
        for (int s = 0; s < 8; ++s)
          cur[s] = rb[s];
        if (c + 96 < n)
          load_chunk(c + 96, rb);
        mfma_chunk(c + 32, cur);
      }
    }
#pragma unroll
    for (int t = 0; t < FT; ++t)
      acc[t] += acc2[t];
    // lane l, register v: row l / 16 + 4 v of the tile, field l % 16
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (oslot[v] >= 0) {
        const int slot = oslot[v];
#pragma unroll
        for (int t = 0; t < FT; ++t) {
          const int f = 16 * t + i;
          if (f < F)
            a.contrib[(int64_t)slot * F + f] = acc[t][v];
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void blend_sum_kernel(BlendArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M)
    return;
  const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
  const int64_t o = a.tperm[t];
  constexpr int kB = 8;
  for (int f0 = 0; f0 < a.F; f0 += kB) {
    double acc[kB];
#pragma unroll
    for (int f = 0; f < kB; ++f)
      acc[f] = 0.0;
    for (int64_t s = s0; s < s1; ++s) {
      if (!a.slotMember[s])
        continue;
#pragma unroll
      for (int f = 0; f < kB; ++f)
        if (f0 + f < a.F)
          acc[f] += a.contrib[s * a.F + f0 + f];
    }
#pragma unroll
    for (int f = 0; f < kB; ++f)
      if (f0 + f < a.F)
        a.out[(int64_t)(f0 + f) * a.M + o] = acc[f];
  }
}

void launch_op_store(const SolveArgs &a, int64_t K, const double *contribS,
                     const int64_t *opOff, const OpStore &op, hipStream_t s) {
  if (K <= 0)
    return;
  const unsigned blocks = (unsigned)((K + 3) / 4);
  if (op.f32)
    op_store_kernel<<<blocks, 256, 0, s>>>(a, (int)K, contribS, opOff, op.f32);
  else
    op_store_kernel<<<blocks, 256, 0, s>>>(a, (int)K, contribS, opOff, op.f64);
  PACMAN_HIP_CHECK(hipGetLastError());
}

template <class T>
static void op_apply_launch(const SolveArgs &a, int64_t K,
                            const int64_t *opOff, const T *op, hipStream_t s) {
  // a plain dot per row is as fast for few fields
  if (a.F == 1)
    op_apply_kernel<1, T><<<(unsigned)K, kOpBlock, 0, s>>>(a, (int)K, opOff,
                                                           op);
  else if (a.F == 2)
    op_apply_kernel<2, T><<<(unsigned)K, kOpBlock, 0, s>>>(a, (int)K, opOff,
                                                           op);
  else if (a.F <= 4)
    op_apply_kernel<4, T><<<(unsigned)K, kOpBlock, 0, s>>>(a, (int)K, opOff,
                                                           op);
  else if (a.F <= 16)
    op_apply_mfma_kernel<1, T><<<(unsigned)K, kOpBlock, 0, s>>>(a, (int)K,
                                                                opOff, op);
  else
    op_apply_mfma_kernel<2, T><<<(unsigned)K, kOpBlock, 0, s>>>(a, (int)K,
                                                                opOff, op);
}

void launch_op_apply(const SolveArgs &a, int64_t K, const int64_t *opOff,
                     const OpStore &op, hipStream_t s) {
  if (K <= 0)
    return;
  if (a.F < 1 || a.F > kOpFMax)
    throw std::runtime_error("operator apply: 1 <= fields <= 32 per launch");
  if (op.f32)
    op_apply_launch(a, K, opOff, (const float *)op.f32, s);
  else
    op_apply_launch(a, K, opOff, (const double *)op.f64, s);
  PACMAN_HIP_CHECK(hipGetLastError());
}

// The same sum for 2 .. 32 fields, 64 consecutive original targets per
// workgroup: a group of L lanes (L >= F) takes one target at a time and reads
// each of its slots' F contiguous values in one coalesced piece (lane f, field
// f; slots in ascending order, as above), the sums are staged in LDS and
// written out field by field as 64 consecutive outputs.
template <int L>
__global__ __launch_bounds__(256) void blend_sum_tile_kernel(BlendArgs a) {
  constexpr int kT = 64;  // targets per workgroup
  __shared__ double tile[kT][L + 1];
  const int64_t o0 = (int64_t)blockIdx.x * kT;
  const int tid = threadIdx.x;
  const int f = tid % L, grp = tid / L;
  constexpr int kGroups = 256 / L;
  const int F = a.F;
  for (int j = grp; j < kT; j += kGroups) {
    const int64_t o = o0 + j;
    double acc = 0.0;
    if (o < a.M) {
      const int64_t t = a.tinv[o];
      const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
      // flag and value loaded unconditionally (no load waits on another)
#pragma unroll 4
      for (int64_t sl = s0; sl < s1; ++sl) {
        const bool mem = a.slotMember[sl] != 0;
        const double c = a.contrib[sl * F + (f < F ? f : 0)];
        acc = (mem && f < F) ? acc + c : acc;
      }
    }
    tile[j][f] = acc;
  }
  __syncthreads();
  for (int x = tid; x < kT * F; x += 256) {
    const int ff = x / kT, j = x % kT;
    if (o0 + j < a.M)
      a.out[(int64_t)ff * a.M + o0 + j] = tile[j][ff];
  }
}

// Transpose through the stored operator: contribS[(s0 + e) * F + f] =
// sum over the member rows q of patch k of op[q][e] g_f(y_q) (rows in
// ascending order, staged 256 at a time); one thread per source column
// (consecutive threads read consecutive columns of a row).
template <class T>
__global__ __launch_bounds__(256) void op_transpose_kernel(
    SolveArgs a, int K, const int64_t *opOff, const T *op, double *contribS) {
  __shared__ double gv[kFMax][256];
  const int patch = xcd_remap(blockIdx.x, K);
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int64_t p0 = a.pairOff[patch];
  const int mall = (int)(a.pairOff[patch + 1] - p0);
  const int F = a.F;
  const int tid = threadIdx.x;
  const T *base = op + opOff[patch];
  double acc[kFMax];
#pragma unroll
  for (int f = 0; f < kFMax; ++f)
    acc[f] = 0.0;
  for (int q0 = 0; q0 < mall; q0 += 256) {
    const int q = q0 + tid;
    if (q < mall) {
      const int pe = a.pairTgt[p0 + q];
      const bool mem = ((unsigned)pe >> 31) != 0;
      const int tp = pe & 0x7fffffff;
      for (int f = 0; f < F; ++f)
        gv[f][tid] = mem ? a.gSorted[(int64_t)f * a.M + tp] : 0.0;
    }
    __syncthreads();
    const int qn = min(256, mall - q0);
    if (tid < n)
      for (int j = 0; j < qn; ++j) {
        const double g = (double)base[(int64_t)(q0 + j) * n + tid];
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          if (f < F)
            acc[f] = fma(g, gv[f][j], acc[f]);
      }
    __syncthreads();
  }
  if (tid < n)
    for (int f = 0; f < F; ++f)
      contribS[(s0 + tid) * F + f] = acc[f];
}

void launch_op_transpose(const SolveArgs &a, int64_t K, const int64_t *opOff,
                         const OpStore &op, double *contribS, hipStream_t s) {
  if (K <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("operator transpose: 1 <= fields <= 8 per launch");
  if (op.f32)
    op_transpose_kernel<<<(unsigned)K, 256, 0, s>>>(
        a, (int)K, opOff, (const float *)op.f32, contribS);
  else
    op_transpose_kernel<<<(unsigned)K, 256, 0, s>>>(
        a, (int)K, opOff, (const double *)op.f64, contribS);
  PACMAN_HIP_CHECK(hipGetLastError());
}

__global__ void invert_perm_kernel(const int *perm, int64_t n, int *inv) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n)
    inv[perm[p]] = (int)p;
}

void launch_blend_sum(const BlendArgs &a, hipStream_t s) {
  if (a.M <= 0)
    return;
  const unsigned tiles = (unsigned)((a.M + 63) / 64);
  if (a.F >= 2 && a.F <= 32 && a.tinv) {
    if (a.F <= 8)
      blend_sum_tile_kernel<8><<<tiles, 256, 0, s>>>(a);
    else if (a.F <= 16)
      blend_sum_tile_kernel<16><<<tiles, 256, 0, s>>>(a);
    else
      blend_sum_tile_kernel<32><<<tiles, 256, 0, s>>>(a);
  } else {
    blend_sum_kernel<<<(unsigned)((a.M + 255) / 256), 256, 0, s>>>(a);
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_invert_perm(const int *perm, int64_t n, int *inv,
                        hipStream_t s) {
  if (n <= 0)
    return;
  invert_perm_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(perm, n, inv);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_slot_weights(const BlendArgs &a, double *wslot, hipStream_t s) {
  if (a.M <= 0)
    return;
  slot_weights_kernel<<<(unsigned)((a.M + 255) / 256), 256, 0, s>>>(a, wslot);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void build_source_transpose(const int *srcIdx, int64_t total, int64_t N,
                            int64_t *srcTOff, int *srcTEnt, hipStream_t s) {
  DevBuf<int> cnt(std::max<int64_t>(N, 1)), cur(std::max<int64_t>(N, 1));
  const unsigned bn = (unsigned)((std::max<int64_t>(N, 1) + 255) / 256);
  zero_i32_kernel<<<bn, 256, 0, s>>>(cnt.get(), N);
  zero_i32_kernel<<<bn, 256, 0, s>>>(cur.get(), N);
  PACMAN_HIP_CHECK(hipGetLastError());
  if (total > 0) {
    src_count_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
        srcIdx, total, cnt.get());
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  exclusive_scan_i32_to_i64(cnt.get(), srcTOff, N, s);
  if (total > 0) {
    src_fill_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
        srcIdx, total, srcTOff, cur.get(), srcTEnt);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  if (N > 0) {
    src_sort_kernel<<<bn, 256, 0, s>>>(srcTOff, N, srcTEnt);
    PACMAN_HIP_CHECK(hipGetLastError());
  }
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

void launch_source_reduce(const int64_t *srcTOff, const int *srcTEnt,
                          int64_t N, int F, const double *contribS,
                          const int *perm, double *out, hipStream_t s) {
  if (N <= 0)
    return;
  src_reduce_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(
      srcTOff, srcTEnt, N, F, contribS, perm, out);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_blend(const BlendArgs &a, hipStream_t s) {
  if (a.M <= 0)
    return;
  if (a.F >= 2 && a.F <= 8 && a.tinv && a.wslot)
    blend_tile_kernel<<<(unsigned)((a.M + 63) / 64), 256, 0, s>>>(a);
  else
    blend_kernel<<<(unsigned)((a.M + 255) / 256), 256, 0, s>>>(a);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_gather_values(const double *vals, const int *perm, int64_t N,
                          int F, double *out, hipStream_t s) {
  if (N <= 0 || F <= 0)
    return;
  if (F > 65535)
    throw std::runtime_error("gather: at most 65535 fields per call");
  gather_values_kernel<<<dim3((unsigned)((N + 255) / 256), (unsigned)F), 256, 0,
                         s>>>(vals, perm, N, out);
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace pacman_amd
