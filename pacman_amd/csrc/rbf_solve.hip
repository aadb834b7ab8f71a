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
// solve_f64_kernel (one workgroup of 16 x GC threads per patch):
//   1. gathers the patch's source coordinates and values into LDS and
//      subtracts the polynomial (the residual b - Q beta, solver.hpp:255-263);
//   2. assembles the Wendland matrix A straight into registers: thread
//      (tr, tc) owns rows tr + 16a and columns tc + GC b of every block that
//      holds lower-triangle entries (cyclic, so every wave keeps work while
//      the trailing matrix shrinks);
//   3. factors A = L D L^T, the symmetric form of the reference's unpivoted
//      Doolittle LU (serial_solve.hpp:76-84): the same pivots, half the flops,
//      no upper triangle. KGRP columns per barrier: their owners publish the
//      raw columns, every thread factors the small diagonal group itself,
//      transforms its own row and column values and applies the rank-KGRP
//      update; the right-hand sides are eliminated in the same sweep;
//   4. back substitution L^T alpha = D^{-1} y, 16 rows at a time;
//   5. t(y) = E alpha + V beta at every member target (solver.hpp:265-273),
//      written to that (target, patch) slot.
// blend_kernel sums the slots of each target in ascending patch order with
// the Shepard-normalised Wendland-C2 weights (solver.hpp:128-169, 275-285),
// deterministically (the reference uses atomic_add in arbitrary order).

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <type_traits>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

constexpr int kFMax = 4;   // fields per launch
constexpr int KGRP = 2;    // columns factored per barrier

// Compile-time loop: f(integral_constant<int, I>) for I in [B, E). Keeps
// register-array indices constant without relying on unrolling heuristics.
template <int B, int E, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// sqrt for the dense arithmetic: hardware reciprocal square root refined by
// two Newton-Raphson (Goldschmidt) steps; within an ulp for the distances
// used here (0 <= x < 1e300). Membership decisions never use it.
__device__ inline double fast_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y;
  double h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  const double d = fma(-s, s, x);
  s = fma(d, h, s);
  return x > 0.0 ? s : 0.0;
}

__device__ inline double dist_fast(double ax, double ay, double az, double bx,
                                   double by, double bz) {
  const double dx = bx - ax, dy = by - ay, dz = bz - az;
  return fast_sqrt(dx * dx + dy * dy + dz * dz);
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

// One wave per patch; coef[(patch * F + f) * 4 + k], pvOut[patch].
__global__ __launch_bounds__(256) void poly_kernel(SolveArgs a, int K,
                                                   double *coef) {
  const int patch = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (patch >= K)
    return;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F, dim = a.dim;
  const double c[3] = {a.cx[patch], a.cy[patch], a.cz[patch]};
  const double sc = a.rinv;
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
  int ca[4];
  int pv = assign_columns(active, ca);
  double L[4][4];
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
  // beta'' = G^-1 Q''^T b on the active columns (normal equations of the
  // well-conditioned scaled basis); coefficients returned around the centre.
  for (int f = 0; f < F; ++f) {
    double hf[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = lane; t < n; t += 64) {
      const int p = a.srcIdx[s0 + t];
      const double b = a.vals[(int64_t)f * a.N + p];
      hf[0] += b;
      hf[1] += (a.sx[p] - c[0]) * sc * b;
      hf[2] += (a.sy[p] - c[1]) * sc * b;
      hf[3] += (a.sz[p] - c[2]) * sc * b;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      hf[k] = wave_sum(hf[k]);
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

// ---------------------------------------------------------------------------
// Per-patch solve
// ---------------------------------------------------------------------------

// Register layout for a 16 x GC thread grid and R 16-row blocks: column
// block b (GC columns) holds lower-triangle entries from row block S b on,
// S = GC / 16. Entries (a, b), a >= S b, are stored at off(b) + a - S b.
template <int R, int GC> struct Layout {
  static constexpr int S = GC / 16;
  static constexpr int NMAX = 16 * R;
  static constexpr int RB = (NMAX + GC - 1) / GC;
  static constexpr int amin(int b) { return S * b; }
  static constexpr int off(int b) { return b * R - S * b * (b - 1) / 2; }
  static constexpr int idx(int a, int b) { return off(b) + a - S * b; }
  static constexpr int NA = off(RB);
};

template <int R, int GC, int KIND>
__global__ __launch_bounds__(16 * GC, GC == 32 ? 4 : 2) void solve_f64_kernel(SolveArgs a,
                                                            const double *coef) {
  using LY = Layout<R, GC>;
  constexpr int NT = 16 * GC;
  constexpr int NMAX = LY::NMAX;
  constexpr int RB = LY::RB;
  constexpr int NSL = NT / 32;  // evaluation slices

  __shared__ double px[NMAX], py[NMAX], pz[NMAX];
  __shared__ double rhs[kFMax][NMAX];
  __shared__ double yv[kFMax][NMAX];
  __shared__ double gcol[2][KGRP][NMAX];
  __shared__ double dinvs[NMAX];
  __shared__ double dblk[16][17];
  __shared__ double spart[kFMax][16];
  __shared__ double scoef[kFMax][4];
  __shared__ double epart[NSL][kFMax][32];
  __shared__ int spivot;

  const int patch = a.patches[xcd_remap(blockIdx.x, a.count)];
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F;
  const int tid = threadIdx.x;
  const int tr = tid & 15;
  const int tc = tid >> 4;
  const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];

  if (tid < kFMax * 4)
    scoef[tid >> 2][tid & 3] =
        (tid >> 2) < F ? coef[((int64_t)patch * F + (tid >> 2)) * 4 + (tid & 3)]
                       : 0.0;
  if (tid == 0)
    spivot = 0;
  __syncthreads();

  // 1. gather; right-hand side = value - polynomial fit
  for (int t = tid; t < NMAX; t += NT) {
    double x = 0.0, y = 0.0, z = 0.0;
    int p = 0;
    if (t < n) {
      p = a.srcIdx[s0 + t];
      x = a.sx[p];
      y = a.sy[p];
      z = a.sz[p];
    }
    px[t] = x;
    py[t] = y;
    pz[t] = z;
    for (int f = 0; f < F; ++f) {
      double v = 0.0;
      if (t < n)
        v = a.vals[(int64_t)f * a.N + p] -
            (scoef[f][0] + scoef[f][1] * (x - cx) + scoef[f][2] * (y - cy) +
             scoef[f][3] * (z - cz));
      rhs[f][t] = v;
      yv[f][t] = 0.0;
    }
    dinvs[t] = 1.0;
  }
  __syncthreads();

  // 2. assemble A into registers
  double A[LY::NA];
  static_for<0, RB>([&](auto bc) {
    constexpr int b = decltype(bc)::value;
    const int l = tc + GC * b;
    const int lc = l < NMAX ? l : 0;
    const double xl = px[lc], yl = py[lc], zl = pz[lc];
    static_for<LY::amin(b), R>([&](auto ac) {
      constexpr int ai = decltype(ac)::value;
      const int i = tr + 16 * ai;
      double v;
      if (i < n && l < n)
        v = wendland<KIND, double, true>(
            dist_fast(px[i], py[i], pz[i], xl, yl, zl), a.dinv);
      else
        v = (i == l) ? 1.0 : 0.0;
      A[LY::idx(ai, b)] = v;
    });
  });

  // 3. A = L D L^T, KGRP columns per barrier. Group g of column block cb
  // holds columns j0 = GC cb + KGRP g ... j0 + KGRP - 1; their owners (tc in
  // the group) publish the raw columns (rows >= j0). Every thread factors
  // the KGRP x KGRP diagonal group itself, transforms its own row and column
  // values (z = a L_g^-T, w = z D_g^-1) and applies A(i, l) -= z_i . w_l;
  // the owners keep z (unnormalised L) as the final column values.
  static_for<0, RB>([&](auto cbc) {
    constexpr int cb = decltype(cbc)::value;
    if (GC * cb >= n)
      return;
    static_for<0, GC / KGRP>([&](auto gc) {
      constexpr int g = decltype(gc)::value;
      constexpr int j0 = GC * cb + KGRP * g;
      constexpr int je = j0 + KGRP;
      if (j0 >= n)
        return;
      double(*buf)[NMAX] = gcol[(cb * (GC / KGRP) + g) & 1];
      const bool owner = (tc / KGRP) == g;
      const int qo = tc % KGRP;
      if (owner) {
        static_for<LY::amin(cb), R>([&](auto ac) {
          constexpr int ai = decltype(ac)::value;
          if constexpr (16 * ai + 15 >= j0) {
            const int i = tr + 16 * ai;
            if (i >= j0)
              buf[qo][i] = A[LY::idx(ai, cb)];
          }
        });
      }
      __syncthreads();
      // small L D L^T of the diagonal group (identical in every thread)
      double Gc[KGRP][KGRP], Lg[KGRP][KGRP], dv[KGRP], di[KGRP];
#pragma unroll
      for (int p = 0; p < KGRP; ++p)
#pragma unroll
        for (int q = 0; q <= p; ++q)
          Gc[p][q] = buf[q][j0 + p];
#pragma unroll
      for (int q = 0; q < KGRP; ++q) {
        dv[q] = Gc[q][q];
        di[q] = 1.0 / dv[q];
#pragma unroll
        for (int p = q + 1; p < KGRP; ++p)
          Lg[p][q] = Gc[p][q] * di[q];
#pragma unroll
        for (int p = q + 1; p < KGRP; ++p)
#pragma unroll
          for (int r2 = q + 1; r2 <= p; ++r2)
            Gc[p][r2] = fma(-Lg[p][q], Gc[r2][q], Gc[p][r2]);
      }
      if (tid == 0) {
#pragma unroll
        for (int q = 0; q < KGRP; ++q) {
          dinvs[j0 + q] = di[q];
          if (dv[q] == 0.0)
            spivot = 1;
        }
      }
      // column factors of this thread's columns l >= je
      double w[RB][KGRP];
      static_for<cb, RB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        const int l = tc + GC * b;
        const bool act = b > cb || tc >= KGRP * g + KGRP;
        double z[KGRP];
#pragma unroll
        for (int q = 0; q < KGRP; ++q) {
          double v = act ? buf[q][l] : 0.0;
#pragma unroll
          for (int p = 0; p < q; ++p)
            v = fma(-z[p], Lg[q][p], v);
          z[q] = v;
        }
#pragma unroll
        for (int q = 0; q < KGRP; ++q)
          w[b][q] = z[q] * di[q];
      });
      // rows: transform, rank-KGRP update, owners keep the final columns
      static_for<LY::amin(cb), R>([&](auto ac) {
        constexpr int ai = decltype(ac)::value;
        constexpr bool rowNever = 16 * ai + 15 < je;
        constexpr bool rowAlways = 16 * ai >= je;
        const int i = tr + 16 * ai;
        if constexpr (!rowNever) {
          const bool act = rowAlways || i >= je;
          double z[KGRP];
#pragma unroll
          for (int q = 0; q < KGRP; ++q) {
            double v = act ? buf[q][i] : 0.0;
#pragma unroll
            for (int p = 0; p < q; ++p)
              v = fma(-z[p], Lg[q][p], v);
            z[q] = v;
          }
          static_for<cb, RB>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            if constexpr (ai >= LY::amin(b)) {
              double acc = A[LY::idx(ai, b)];
#pragma unroll
              for (int q = 0; q < KGRP; ++q)
                acc = fma(-z[q], w[b][q], acc);
              A[LY::idx(ai, b)] = acc;
            }
          });
          if (owner && act) {
            double zq = z[0];
#pragma unroll
            for (int q = 1; q < KGRP; ++q)
              zq = (qo == q) ? z[q] : zq;
            A[LY::idx(ai, cb)] = zq;
          }
        }
        if constexpr (16 * ai + 15 >= j0 && 16 * ai < je) {
          // rows inside the group: unnormalised L below, pivots on the diagonal
          if (owner && i >= j0 && i < je) {
            const int pr = i - j0;
            double v = A[LY::idx(ai, cb)];
#pragma unroll
            for (int p = 0; p < KGRP; ++p)
#pragma unroll
              for (int q = 0; q < KGRP; ++q) {
                if (p == pr && q == qo && p > q)
                  v = Lg[p][q] * dv[q];
                if (p == pr && q == qo && p == q)
                  v = dv[q];
              }
            A[LY::idx(ai, cb)] = v;
          }
        }
      });
      // right-hand sides: y_g = L_g^-1 b_g; later rows b_t -= w_t . y_g
      if (tid < NMAX) {
        const bool rowAct = tid >= je && tid < n;
        const bool inGrp = tid >= j0 && tid < je;
        double wt[KGRP];
        {
          double z[KGRP];
#pragma unroll
          for (int q = 0; q < KGRP; ++q) {
            double v = rowAct ? buf[q][tid] : 0.0;
#pragma unroll
            for (int p = 0; p < q; ++p)
              v = fma(-z[p], Lg[q][p], v);
            z[q] = v;
          }
#pragma unroll
          for (int q = 0; q < KGRP; ++q)
            wt[q] = z[q] * di[q];
        }
        for (int f = 0; f < F; ++f) {
          double yq[KGRP];
#pragma unroll
          for (int q = 0; q < KGRP; ++q) {
            double v = rhs[f][j0 + q];
#pragma unroll
            for (int p = 0; p < q; ++p)
              v = fma(-Lg[q][p], yq[p], v);
            yq[q] = v;
          }
          if (rowAct) {
            double b = rhs[f][tid];
#pragma unroll
            for (int q = 0; q < KGRP; ++q)
              b = fma(-wt[q], yq[q], b);
            rhs[f][tid] = b;
          }
          if (inGrp) {
            double v = yq[0];
#pragma unroll
            for (int q = 1; q < KGRP; ++q)
              v = (tid - j0 == q) ? yq[q] : v;
            yv[f][tid] = v;
          }
        }
      }
    });
  });
  __syncthreads();

  // 4. back substitution, 16 rows per block, bottom-up
  static_for<0, R>([&](auto bc) {
    constexpr int B = R - 1 - decltype(bc)::value;
    constexpr int bb = (16 * B) / GC;        // column block of rows 16B..
    constexpr int toff = 16 * B - GC * bb;   // first tc of those columns
    if (16 * B >= n)
      return;
    const bool mine = tc >= toff && tc < toff + 16;
    double pf[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      pf[f] = 0.0;
    if (mine) {
      static_for<B + 1, R>([&](auto ac) {
        constexpr int ai = decltype(ac)::value;
        const double c = A[LY::idx(ai, bb)];
        const int i = tr + 16 * ai;
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          if (f < F)
            pf[f] = fma(c, rhs[f][i], pf[f]);
      });
    }
#pragma unroll
    for (int f = 0; f < kFMax; ++f) {
      if (f < F) {
        double v = pf[f];
        v += __shfl_xor(v, 8, kWave);
        v += __shfl_xor(v, 4, kWave);
        v += __shfl_xor(v, 2, kWave);
        v += __shfl_xor(v, 1, kWave);
        if (mine && tr == 0)
          spart[f][tc - toff] = v;
      }
    }
    if (mine)
      dblk[tr][tc - toff] = A[LY::idx(B, bb)];
    __syncthreads();
    // 16 lanes of wave 0 solve the 16 x 16 unit-lower system, one row per
    // lane; the solved value of row jj is broadcast from lane jj.
    if (tid < kWave) {
      const int k = tid & 15;
      double col[16];
#pragma unroll
      for (int jj = 0; jj < 16; ++jj)
        col[jj] = dblk[jj][k];
      const double dk = dinvs[16 * B + k];
      for (int f = 0; f < F; ++f) {
        double sv = yv[f][16 * B + k] - spart[f][k];
#pragma unroll
        for (int jj = 15; jj >= 0; --jj) {
          const double al = __shfl(sv * dk, jj, kWave);
          if (k < jj)
            sv = fma(-col[jj], al, sv);
          if (k == jj && tid < 16)
            rhs[f][16 * B + jj] = al;
        }
      }
    }
    __syncthreads();
  });

  if (tid == 0 && a.status)
    a.status[patch] = spivot;

  // 5. evaluate at the member targets and write each (target, patch) slot.
  // Tiles of 32 pairs; thread = (pair, slice): slice s of NSL sums sources
  // [s n / NSL, (s + 1) n / NSL); slices are added in order (deterministic).
  {
    const int64_t p0 = a.pairOff[patch], p1 = a.pairOff[patch + 1];
    const int mall = (int)(p1 - p0);
    const int lane = tid & 31, sl = tid >> 5;
    const int qn = (n + NSL - 1) / NSL;
    const int l0 = min(n, sl * qn), l1 = min(n, l0 + qn);
    for (int tile = 0; tile < mall; tile += 32) {
      const int pi = tile + lane;
      int e = 0;
      if (pi < mall)
        e = a.pairTgt[p0 + pi];
      const bool mem = pi < mall && ((unsigned)e >> 31) != 0;
      const int tp = e & 0x7fffffff;
      double yx = 0.0, yy = 0.0, yz = 0.0;
      if (mem) {
        yx = a.tx[tp];
        yy = a.ty[tp];
        yz = a.tz[tp];
      }
      double acc[kFMax];
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        acc[f] = 0.0;
      if (mem) {
        for (int l = l0; l < l1; ++l) {
          const double phi = wendland<KIND, double, true>(
              dist_fast(yx, yy, yz, px[l], py[l], pz[l]), a.dinv);
#pragma unroll
          for (int f = 0; f < kFMax; ++f)
            if (f < F)
              acc[f] = fma(phi, rhs[f][l], acc[f]);
        }
      }
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          epart[sl][f][lane] = acc[f];
      __syncthreads();
      if (sl == 0 && mem) {
        const int slot = a.pairSlot[p0 + pi];
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          if (f < F) {
            double t = epart[0][f][lane];
#pragma unroll
            for (int q = 1; q < NSL; ++q)
              t += epart[q][f][lane];
            // V beta (solver.hpp:272), around the centre
            const double vb = scoef[f][0] + scoef[f][1] * (yx - cx) +
                              scoef[f][2] * (yy - cy) + scoef[f][3] * (yz - cz);
            a.contrib[(int64_t)slot * F + f] = t + vb;
          }
      }
      __syncthreads();
    }
  }
}

// Shepard-normalised PU blend, one thread per target (cell order).
__global__ __launch_bounds__(256) void blend_kernel(BlendArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M)
    return;
  const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
  const double yx = a.tx[t], yy = a.ty[t], yz = a.tz[t];
  double sum = 0.0;
  for (int64_t s = s0; s < s1; ++s) {
    const int k = a.slotPatch[s];
    // DistanceNoCheck(target, center), solver.hpp:158-161
    const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
    sum += wendland<1, double, true>(d, a.rinv);
  }
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
      const int k = a.slotPatch[s];
      const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
      const double w = wendland<1, double, true>(d, a.rinv) / sum;
#pragma unroll
      for (int f = 0; f < kB; ++f)
        if (f0 + f < a.F)
          acc[f] += w * a.contrib[s * a.F + f0 + f];
    }
#pragma unroll
    for (int f = 0; f < kB; ++f)
      if (f0 + f < a.F)
        a.out[(int64_t)(f0 + f) * a.M + o] = acc[f];
  }
}

__global__ void gather_values_kernel(const double *vals, const int *perm,
                                     int64_t N, int F, double *out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N)
    return;
  const int64_t i = perm[p];
  for (int f = 0; f < F; ++f)
    out[(int64_t)f * N + p] = vals[(int64_t)f * N + i];
}

template <int R, int GC>
void launch_r(int kind, const SolveArgs &a, const double *coef,
              hipStream_t s) {
  const dim3 grid((unsigned)a.count), block(16 * GC);
  switch (kind) {
  case 0:
    solve_f64_kernel<R, GC, 0><<<grid, block, 0, s>>>(a, coef);
    break;
  case 1:
    solve_f64_kernel<R, GC, 1><<<grid, block, 0, s>>>(a, coef);
    break;
  case 2:
    solve_f64_kernel<R, GC, 2><<<grid, block, 0, s>>>(a, coef);
    break;
  case 3:
    solve_f64_kernel<R, GC, 3><<<grid, block, 0, s>>>(a, coef);
    break;
  default:
    solve_f64_kernel<R, GC, 4><<<grid, block, 0, s>>>(a, coef);
    break;
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace

int solve_fields_per_launch() { return kFMax; }

void launch_poly(const SolveArgs &a, int64_t K, double *coef, hipStream_t s) {
  if (K <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("poly: 1 <= fields <= 4 per launch");
  poly_kernel<<<(unsigned)((K + 3) / 4), 256, 0, s>>>(a, (int)K, coef);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_solve_f64(int R, int kind, const SolveArgs &a, const double *coef,
                      hipStream_t s) {
  if (a.count <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("solve: 1 <= fields <= 4 per launch");
  switch (R) {
  case 2:
    launch_r<2, 16>(kind, a, coef, s);
    break;
  case 4:
    launch_r<4, 16>(kind, a, coef, s);
    break;
  case 6:
    launch_r<6, 16>(kind, a, coef, s);
    break;
  case 8:
    launch_r<8, 32>(kind, a, coef, s);
    break;
  case 10:
    launch_r<10, 32>(kind, a, coef, s);
    break;
  case 12:
    launch_r<12, 32>(kind, a, coef, s);
    break;
  default:
    throw std::runtime_error("solve: unsupported register tile R");
  }
}

void launch_blend(const BlendArgs &a, hipStream_t s) {
  if (a.M <= 0)
    return;
  blend_kernel<<<(unsigned)((a.M + 255) / 256), 256, 0, s>>>(a);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_gather_values(const double *vals, const int *perm, int64_t N,
                          int F, double *out, hipStream_t s) {
  if (N <= 0)
    return;
  gather_values_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(vals, perm,
                                                                   N, F, out);
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace pacman_amd
