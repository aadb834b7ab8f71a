// Per-patch RBF-PUM solve and partition-of-unity blend (gfx950, FP64).
//
// One 256-thread workgroup owns one patch (n sources, m targets) and runs
// the whole per-patch chain of src/rbf_pum/solver/solver.hpp:180-286 without
// writing any matrix to HBM:
//   1. gather the patch's source coordinates and values into LDS;
//   2. wave 0: least-squares fit of the linear polynomial
//      (SerialSolveQR, serial_solve.hpp:46-67) with the conditioning loop of
//      SerialFillQ (serial_ops.hpp:108-139). The Gram matrix of the basis
//      [1, (x_a - c_a) / r] (c = patch centre, r = patch radius) is formed in
//      one batched wave reduction; it spans the same space as the
//      reference's [1, x_a] and is well conditioned. The condition number of
//      the reference's raw Q = [1, x_a] is that of R'' T (R'' = Cholesky
//      factor of the scaled Gram matrix, T the basis change): a Frobenius
//      bound settles cond <= 1e5 at once, the exact Jacobi singular values
//      are computed only otherwise. While cond > 1e5 the active axis with
//      the smallest coordinate range is dropped (serial_ops.hpp:28-57); the
//      fit then uses exactly the active columns. Residual b - Q beta
//      (solver.hpp:255-263);
//   3. all threads: assemble the Wendland matrix A straight into registers.
//      Thread (tr, tc) of the 16 x 16 thread grid owns rows tr + 16a and
//      columns tc + 16b for b <= a (every lower-triangle entry, cyclically);
//   4. factor A with the symmetric form of the reference's unpivoted
//      Doolittle LU (serial_solve.hpp:76-84): A = L D L^T has exactly the
//      pivots d_j of that LU, half its flops and no upper triangle. The
//      column loop is split into a compile-time loop over 16-column blocks
//      and a run-time loop inside the block, so every thread's active
//      register range is static; one barrier per column, the column and its
//      pivot reciprocal double-buffered in LDS. The right-hand sides are
//      eliminated in the same sweep;
//   5. back substitution L^T alpha = D^{-1} y, 16 rows at a time;
//   6. t(y) = E alpha + V beta at every member target (solver.hpp:265-273),
//      written to that (target, patch) slot; the blend kernel sums the slots
//      of each target in ascending patch order with the Shepard-normalised
//      Wendland-C2 weights (solver.hpp:128-169, 275-285), deterministically.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "rbf_internal.hpp"

namespace pacman_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kG = 16;      // thread grid is kG x kG
constexpr int kFMax = 8;    // fields per launch

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

// Column c (0..3) of the active basis holds axis ca[c] (ca[0] = -1: the
// constant). Entries of a 4 x 4 matrix over the full basis [1, u0, u1, u2]
// are gathered with selects so every register index stays static.
__device__ inline int basis_index(int axis) { return axis + 1; }

__device__ inline double pick4(const double (&G)[4][4], int r, int c) {
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      v = (i == r && j == c) ? G[i][j] : v;
  return v;
}

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

template <int R, int KIND>
__global__ __launch_bounds__(kThreads) void solve_f64_kernel(SolveArgs a) {
  constexpr int NMAX = kG * R;
  constexpr int NR = (NMAX + 63) / 64;
  constexpr int NA = R * (R + 1) / 2;

  __shared__ double px[NMAX], py[NMAX], pz[NMAX];
  __shared__ double rhs[kFMax][NMAX];
  __shared__ double colbuf[2][NMAX];
  __shared__ double dinvs[NMAX];
  __shared__ double cbinv[2];
  __shared__ double dblk[kG][kG + 1];
  __shared__ double spart[kFMax][kG];
  __shared__ double scoef[kFMax][4];
  __shared__ int spv, spivot;

  const int patch = a.patches[xcd_remap(blockIdx.x, a.count)];
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F;
  const int tid = threadIdx.x;
  const int tr = tid & (kG - 1);
  const int tc = tid >> 4;
  const int dim = a.dim;

  // 1. gather
  for (int t = tid; t < NMAX; t += kThreads) {
    if (t < n) {
      const int p = a.srcIdx[s0 + t];
      px[t] = a.sx[p];
      py[t] = a.sy[p];
      pz[t] = a.sz[p];
      for (int f = 0; f < F; ++f)
        rhs[f][t] = a.vals[(int64_t)f * a.N + p];
    } else {
      px[t] = 0.0;
      py[t] = 0.0;
      pz[t] = 0.0;
      for (int f = 0; f < F; ++f)
        rhs[f][t] = 0.0;
    }
    dinvs[t] = 1.0;
  }
  if (tid == 0)
    spivot = 0;
  __syncthreads();

  // 2. polynomial least squares (wave 0)
  if (tid < kWave) {
    const int lane = tid;
    const double c[3] = {a.cx[patch], a.cy[patch], a.cz[patch]};
    const double sc = a.rinv;
    double U3[3][NR];
    bool valid[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int t = lane + 64 * s;
      valid[s] = t < n;
      U3[0][s] = valid[s] ? (px[t] - c[0]) * sc : 0.0;
      U3[1][s] = valid[s] ? (py[t] - c[1]) * sc : 0.0;
      U3[2][s] = valid[s] ? (pz[t] - c[2]) * sc : 0.0;
    }
    // one batched reduction: coordinate ranges and the Gram matrix
    double mn[3], mx[3], G[4][4];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      mn[ax] = DBL_MAX;
      mx[ax] = -DBL_MAX;
#pragma unroll
      for (int s = 0; s < NR; ++s)
        if (valid[s]) {
          const int t = lane + 64 * s;
          const double x = ax == 0 ? px[t] : ax == 1 ? py[t] : pz[t];
          mn[ax] = fmin(mn[ax], x);
          mx[ax] = fmax(mx[ax], x);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        G[i][j] = 0.0;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const double u[4] = {valid[s] ? 1.0 : 0.0, U3[0][s], U3[1][s], U3[2][s]};
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
    bool ok = true;
    for (;;) {
      // Gram of the active columns, its Cholesky factor, cond(raw Q)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          L[i][j] = (i < pv && j < pv)
                        ? pick4(G, ca[i] + 1, ca[j] + 1)
                        : (i == j ? 1.0 : 0.0);
      ok = chol4(L, pv);
      if (pv <= 1)
        break;
      bool keep = false;
      if (ok) {
        // raw basis [1, x_a] = [1, u_a] T with T = [[1, c_a], [0, 1/sc]]
        double Rr[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            double v = 0.0;
            if (i < pv && j < pv && i <= j) {
              // (R'' T)_ij = sum_m R''_im T_mj with R'' = L^T, T_00 = 1,
              // T_0j = c_j, T_jj = 1 / sc (j >= 1), zero elsewhere.
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
        // ||R||_F * ||R^-1||_F >= cond_2(R): settles cond <= 1e5 cheaply
        double Ri[4][4];
        double fr = 0.0, fi = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int i = 3; i >= 0; --i) {
            double v = 0.0;
            if (i <= j && j < pv) {
              double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
              for (int k = 0; k < 4; ++k)
                if (k > i && k <= j)
                  s -= Rr[i][k] * Ri[k][j];
              v = s / Rr[i][i];
            }
            Ri[i][j] = v;
          }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            fr += Rr[i][j] * Rr[i][j];
            fi += Ri[i][j] * Ri[i][j];
          }
        if (sqrt(fr) * sqrt(fi) <= kCondMax) {
          keep = true;
        } else {
          double S[4];
          small_svd(Rr, pv, S);
          const double smin = pv == 4 ? S[3] : pv == 3 ? S[2] : S[1];
          keep = !(S[0] / fmax(smin, kEps) > kCondMax);
        }
      }
      if (keep)
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
    // beta'' = G^-1 h on the active columns (normal equations of the
    // well-conditioned scaled basis), then the residual b - Q beta.
#pragma unroll
    for (int f = 0; f < kFMax; ++f) {
      if (f >= F)
        continue;
      // Q''^T b over the full scaled basis
      double hf[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int t = lane + 64 * s;
        const double b = valid[s] ? rhs[f][t] : 0.0;
        hf[0] += b;
        hf[1] += U3[0][s] * b;
        hf[2] += U3[1][s] * b;
        hf[3] += U3[2][s] * b;
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
      // per-axis coefficients of the scaled basis (inactive axes: zero)
      double cax[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        cax[ax] = ca[1] == ax ? beta[1]
                  : ca[2] == ax ? beta[2]
                  : ca[3] == ax ? beta[3]
                                : 0.0;
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int t = lane + 64 * s;
        if (valid[s])
          rhs[f][t] -= beta[0] + cax[0] * U3[0][s] + cax[1] * U3[1][s] +
                       cax[2] * U3[2][s];
      }
      if (lane == 0) {
        scoef[f][0] = beta[0];
        scoef[f][1] = cax[0] * sc;
        scoef[f][2] = cax[1] * sc;
        scoef[f][3] = cax[2] * sc;
      }
    }
    if (lane == 0)
      spv = pv;
  }

  // 3. assemble A (lower block triangle, registers)
  double A[NA];
#pragma unroll
  for (int ai = 0; ai < R; ++ai) {
    const int i = tr + kG * ai;
    const double xi = px[i], yi = py[i], zi = pz[i];
#pragma unroll
    for (int bi = 0; bi <= ai; ++bi) {
      const int l = tc + kG * bi;
      double v;
      if (i < n && l < n)
        v = wendland<KIND, double>(dist_fast(xi, yi, zi, px[l], py[l], pz[l]),
                                   a.dinv);
      else
        v = (i == l) ? 1.0 : 0.0;
      A[ai * (ai + 1) / 2 + bi] = v;
    }
  }
  __syncthreads();

  // 4. A = L D L^T with forward elimination of the right-hand sides
#pragma unroll
  for (int jb = 0; jb < R; ++jb) {
    if (kG * jb >= n)
      break;
    for (int jt = 0; jt < kG; ++jt) {
      const int j = kG * jb + jt;
      if (j >= n)
        break;
      double *cb = colbuf[j & 1];
      if (tc == jt) {
#pragma unroll
        for (int ai = jb; ai < R; ++ai) {
          const int i = tr + kG * ai;
          if (ai > jb || i >= j) {
            const double v = A[ai * (ai + 1) / 2 + jb];
            cb[i] = v;
            if (ai == jb && i == j) {
              const double iv = 1.0 / v;
              cbinv[j & 1] = iv;
              dinvs[j] = iv;
              if (v == 0.0)
                spivot = 1;
            }
          }
        }
      }
      __syncthreads();
      const double inv = cbinv[j & 1];
      if (tid > j && tid < n) {
        const double ct = cb[tid];
        for (int f = 0; f < F; ++f)
          rhs[f][tid] -= ct * (rhs[f][j] * inv);
      }
      double wl[R];
#pragma unroll
      for (int bi = jb; bi < R; ++bi) {
        const int l = tc + kG * bi;
        wl[bi] = (bi > jb || l > j) ? cb[l] * inv : 0.0;
      }
#pragma unroll
      for (int ai = jb; ai < R; ++ai) {
        const int i = tr + kG * ai;
        const double ci = (ai > jb || i > j) ? cb[i] : 0.0;
#pragma unroll
        for (int bi = jb; bi <= ai; ++bi)
          A[ai * (ai + 1) / 2 + bi] = fma(-ci, wl[bi], A[ai * (ai + 1) / 2 + bi]);
      }
    }
  }
  __syncthreads();

  // 5. back substitution, 16 rows per block, bottom-up
#pragma unroll
  for (int B = R - 1; B >= 0; --B) {
    if (kG * B >= n)
      continue;
    double pf[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      pf[f] = 0.0;
#pragma unroll
    for (int ai = B + 1; ai < R; ++ai) {
      const int i = tr + kG * ai;
      const double c = A[ai * (ai + 1) / 2 + B];
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          pf[f] = fma(c, rhs[f][i], pf[f]);
    }
#pragma unroll
    for (int f = 0; f < kFMax; ++f) {
      if (f < F) {
        double v = pf[f];
        v += __shfl_xor(v, 8, kWave);
        v += __shfl_xor(v, 4, kWave);
        v += __shfl_xor(v, 2, kWave);
        v += __shfl_xor(v, 1, kWave);
        if (tr == 0)
          spart[f][tc] = v;
      }
    }
    dblk[tr][tc] = A[B * (B + 1) / 2 + B];
    __syncthreads();
    if (tid < F) {
      const int f = tid;
      double sv[kG];
#pragma unroll
      for (int jj = 0; jj < kG; ++jj)
        sv[jj] = rhs[f][kG * B + jj] - spart[f][jj];
#pragma unroll
      for (int jj = kG - 1; jj >= 0; --jj) {
        const double al = sv[jj] * dinvs[kG * B + jj];
        rhs[f][kG * B + jj] = al;
#pragma unroll
        for (int k = 0; k < jj; ++k)
          sv[k] = fma(-dblk[jj][k], al, sv[k]);
      }
    }
    __syncthreads();
  }

  if (tid == 0) {
    if (a.pvOut)
      a.pvOut[patch] = spv;
    if (a.status)
      a.status[patch] = spivot;
  }

  // 6. evaluate at the member targets and write each (target, patch) slot
  const int64_t p0 = a.pairOff[patch], p1 = a.pairOff[patch + 1];
  for (int64_t p = p0 + tid; p < p1; p += kThreads) {
    const int e = a.pairTgt[p];
    if (((unsigned)e >> 31) == 0)
      continue;
    const int tp = e & 0x7fffffff;
    const double yx = a.tx[tp], yy = a.ty[tp], yz = a.tz[tp];
    double acc[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      acc[f] = 0.0;
    for (int l = 0; l < n; ++l) {
      const double phi = wendland<KIND, double>(
          dist_fast(yx, yy, yz, px[l], py[l], pz[l]), a.dinv);
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          acc[f] = fma(phi, rhs[f][l], acc[f]);
    }
    const int slot = a.pairSlot[p];
    const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      if (f < F) {
        // V beta (solver.hpp:272) in the scaled basis around the centre
        const double vb = scoef[f][0] + scoef[f][1] * (yx - cx) +
                          scoef[f][2] * (yy - cy) + scoef[f][3] * (yz - cz);
        a.contrib[(int64_t)slot * F + f] = acc[f] + vb;
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
    sum += wendland<1, double>(d, a.rinv);
  }
  const int64_t o = a.tperm[t];
  for (int f0 = 0; f0 < a.F; f0 += kFMax) {
    double acc[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      acc[f] = 0.0;
    for (int64_t s = s0; s < s1; ++s) {
      if (!a.slotMember[s])
        continue;
      const int k = a.slotPatch[s];
      const double d = dist_fast(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
      const double w = wendland<1, double>(d, a.rinv) / sum;
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f0 + f < a.F)
          acc[f] += w * a.contrib[s * a.F + f0 + f];
    }
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
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

template <int R>
void launch_r(int kind, const SolveArgs &a, hipStream_t s) {
  const dim3 grid((unsigned)a.count), block(kThreads);
  switch (kind) {
  case 0:
    solve_f64_kernel<R, 0><<<grid, block, 0, s>>>(a);
    break;
  case 1:
    solve_f64_kernel<R, 1><<<grid, block, 0, s>>>(a);
    break;
  case 2:
    solve_f64_kernel<R, 2><<<grid, block, 0, s>>>(a);
    break;
  case 3:
    solve_f64_kernel<R, 3><<<grid, block, 0, s>>>(a);
    break;
  default:
    solve_f64_kernel<R, 4><<<grid, block, 0, s>>>(a);
    break;
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace

void launch_solve_f64(int R, int kind, const SolveArgs &a, hipStream_t s) {
  if (a.count <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::runtime_error("solve: 1 <= fields <= 8 per launch");
  switch (R) {
  case 2:
    launch_r<2>(kind, a, s);
    break;
  case 4:
    launch_r<4>(kind, a, s);
    break;
  case 6:
    launch_r<6>(kind, a, s);
    break;
  case 8:
    launch_r<8>(kind, a, s);
    break;
  case 10:
    launch_r<10>(kind, a, s);
    break;
  case 12:
    launch_r<12>(kind, a, s);
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
