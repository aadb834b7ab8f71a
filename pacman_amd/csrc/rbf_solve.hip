// Per-patch RBF-PUM solve and partition-of-unity blend (gfx950, FP64).
//
// One 256-thread workgroup owns one patch (n sources, m targets) and runs
// the whole per-patch chain of src/rbf_pum/solver/solver.hpp:180-286 without
// touching HBM for any matrix:
//   1. gather the patch's source coordinates and values into LDS;
//   2. wave 0: polynomial least squares on Q = [1, x_a] by Householder QR,
//      with the conditioning loop of SerialFillQ (serial_ops.hpp:108-139):
//      singular values of R (= those of Q); cond > 1e5 drops the active axis
//      with the smallest coordinate range, then Q is refitted on exactly the
//      active columns; residual b - Q beta (solver.hpp:255-263);
//   3. all threads: assemble the Wendland matrix A straight into registers.
//      Thread (tr, tc) of the 16 x 16 thread grid owns rows tr + 16a and
//      columns tc + 16b for b <= a (every lower-triangle entry, cyclically);
//   4. factor A with the symmetric form of the reference's unpivoted
//      Doolittle LU (serial_solve.hpp:76-84): A = L D L^T has exactly the
//      pivots d_j of that LU, half its flops and no upper triangle. The
//      right-hand sides are eliminated in the same sweep (one barrier per
//      column, the column broadcast double-buffered in LDS);
//   5. blocked back substitution L^T alpha = D^{-1} y, 16 rows at a time;
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

__device__ inline double dist3(double ax, double ay, double az, double bx,
                               double by, double bz) {
  return sqrt(sqdiff3(ax, ay, az, bx, by, bz));
}

// Singular values (descending) of a p x p upper-triangular R by one-sided
// Jacobi; p <= 4. Every lane runs it on identical data. All register arrays
// are indexed by compile-time constants.
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

// Householder QR of the n x pv matrix held as q[c][s] (rows lane + 64 s),
// applied to the first F right-hand sides x[f][s]. Same textbook reflector
// as the oracle: v = col - alpha e_j with alpha = -sign(x0) |col|.
template <int NR, int NF>
__device__ inline void wave_householder(double (&q)[4][NR], int pv, int lane,
                                        double (&x)[NF][NR], int F) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j >= pv)
      continue;
    double loc = 0.0;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int t = lane + 64 * s;
      if (t >= j)
        loc += q[j][s] * q[j][s];
    }
    const double nrm = sqrt(wave_sum(loc));
    if (nrm == 0.0)
      continue;
    const double x0 = __shfl(q[j][0], j, kWave);
    const double alpha = (x0 >= 0 ? -nrm : nrm);
    double v[NR];
    double vl = 0.0;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int t = lane + 64 * s;
      v[s] = (t > j) ? q[j][s] : (t == j ? x0 - alpha : 0.0);
      vl += v[s] * v[s];
    }
    const double vv = wave_sum(vl);
    if (vv == 0.0)
      continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= j && c < pv) {
        double sl = 0.0;
#pragma unroll
        for (int s = 0; s < NR; ++s)
          sl += v[s] * q[c][s];
        const double f = 2.0 * wave_sum(sl) / vv;
#pragma unroll
        for (int s = 0; s < NR; ++s)
          q[c][s] -= f * v[s];
      }
    }
#pragma unroll
    for (int ff = 0; ff < NF; ++ff) {
      if (ff < F) {
        double sl = 0.0;
#pragma unroll
        for (int s = 0; s < NR; ++s)
          sl += v[s] * x[ff][s];
        const double f = 2.0 * wave_sum(sl) / vv;
#pragma unroll
        for (int s = 0; s < NR; ++s)
          x[ff][s] -= f * v[s];
      }
    }
  }
}

// Column c (1..3) of Q holds axis ca[c]; selects keep the indexing static.
template <int NR>
__device__ inline void fill_q(double (&q)[4][NR], const double (&X)[3][NR],
                              const bool (&valid)[NR], const int (&ca)[4]) {
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    q[0][s] = valid[s] ? 1.0 : 0.0;
#pragma unroll
    for (int c = 1; c < 4; ++c)
      q[c][s] = ca[c] == 0 ? X[0][s]
                : ca[c] == 1 ? X[1][s]
                : ca[c] == 2 ? X[2][s]
                             : 0.0;
  }
}

__device__ inline int assign_columns(const bool (&active)[3], int (&ca)[4]) {
  int cols = 1;
  ca[0] = -1;
  ca[1] = ca[2] = ca[3] = -1;
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

template <int R, int KIND>
__global__ __launch_bounds__(kThreads) void solve_f64_kernel(SolveArgs a) {
  constexpr int NMAX = kG * R;
  constexpr int NR = (NMAX + 63) / 64;
  constexpr int NA = R * (R + 1) / 2;

  __shared__ double px[NMAX], py[NMAX], pz[NMAX];
  __shared__ double rhs[kFMax][NMAX];
  __shared__ double colbuf[2][NMAX];
  __shared__ double dblk[kG][kG + 1];
  __shared__ double spart[kFMax][kG];
  __shared__ double scoef[kFMax][4];
  __shared__ int spv;

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
  }
  __syncthreads();

  // 2. polynomial least squares (wave 0)
  if (tid < kWave) {
    const int lane = tid;
    double X[3][NR];
    bool valid[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) {
      const int t = lane + 64 * s;
      valid[s] = t < n;
      X[0][s] = valid[s] ? px[t] : 0.0;
      X[1][s] = valid[s] ? py[t] : 0.0;
      X[2][s] = valid[s] ? pz[t] : 0.0;
    }
    // coordinate range per axis (SerialDeactivateOneAxis, serial_ops.hpp:28-57)
    double range[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      double mn = DBL_MAX, mx = -DBL_MAX;
#pragma unroll
      for (int s = 0; s < NR; ++s)
        if (valid[s]) {
          mn = fmin(mn, X[ax][s]);
          mx = fmax(mx, X[ax][s]);
        }
      range[ax] = fabs(wave_max(mx) - wave_min(mn));
    }
    bool active[3] = {dim > 0, dim > 1, dim > 2};
    int ca[4];
    double q[4][NR];
    double xr[kFMax][NR];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
#pragma unroll
      for (int s = 0; s < NR; ++s)
        xr[f][s] = 0.0;
    int pv = assign_columns(active, ca);
    while (pv > 1 && n > 0) {
      fill_q<NR>(q, X, valid, ca);
      wave_householder<NR, kFMax>(q, pv, lane, xr, 0);
      double U[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double v = __shfl(q[c][0], i, kWave);
          U[i][c] = (i < pv && c < pv && i <= c && i < n) ? v : 0.0;
        }
      double S[4];
      small_svd(U, pv, S);
      const double smin = pv == 4 ? S[3] : pv == 3 ? S[2] : S[1];
      const double cond = S[0] / fmax(smin, kEps);
      if (!(cond > kCondMax))
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
    // final fit on the active columns
    fill_q<NR>(q, X, valid, ca);
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      if (f < F)
#pragma unroll
        for (int s = 0; s < NR; ++s) {
          const int t = lane + 64 * s;
          xr[f][s] = valid[s] ? rhs[f][t] : 0.0;
        }
    wave_householder<NR, kFMax>(q, pv, lane, xr, F);
    const int pvr = min(pv, n);
#pragma unroll
    for (int f = 0; f < kFMax; ++f) {
      if (f >= F)
        continue;
      double beta[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 3; j >= 0; --j) {
        if (j < pvr) {
          double s = __shfl(xr[f][0], j, kWave);
#pragma unroll
          for (int c = j + 1; c < 4; ++c)
            if (c < pv)
              s -= __shfl(q[c][0], j, kWave) * beta[c];
          beta[j] = s / __shfl(q[j][0], j, kWave);
        }
      }
      // per-axis coefficients: inactive axes contribute exact zeros
      double cax[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
        cax[ax] = ca[1] == ax ? beta[1]
                  : ca[2] == ax ? beta[2]
                  : ca[3] == ax ? beta[3]
                                : 0.0;
      // residual b - Q beta (solver.hpp:258-263)
#pragma unroll
      for (int s = 0; s < NR; ++s) {
        const int t = lane + 64 * s;
        if (valid[s]) {
          const double qb =
              beta[0] + cax[0] * X[0][s] + cax[1] * X[1][s] + cax[2] * X[2][s];
          rhs[f][t] -= qb;
        }
      }
      if (lane == 0) {
        scoef[f][0] = beta[0];
        scoef[f][1] = cax[0];
        scoef[f][2] = cax[1];
        scoef[f][3] = cax[2];
      }
    }
    if (lane == 0)
      spv = pv;
  }

  // 3. assemble A (lower block triangle, registers)
  double A[NA];
  {
#pragma unroll
    for (int ai = 0; ai < R; ++ai) {
      const int i = tr + kG * ai;
      const double xi = px[i], yi = py[i], zi = pz[i];
#pragma unroll
      for (int bi = 0; bi <= ai; ++bi) {
        const int l = tc + kG * bi;
        double v;
        if (i < n && l < n)
          v = wendland<KIND, double>(dist3(xi, yi, zi, px[l], py[l], pz[l]),
                                     a.dinv);
        else
          v = (i == l) ? 1.0 : 0.0;
        A[ai * (ai + 1) / 2 + bi] = v;
      }
    }
  }
  __syncthreads();

  // 4. A = L D L^T with forward elimination of the right-hand sides
  bool zeroPivot = false;
  for (int j = 0; j < n; ++j) {
    const int jb = j >> 4, jt = j & (kG - 1);
    double *cb = colbuf[j & 1];
    if (tc == jt) {
#pragma unroll
      for (int ai = 0; ai < R; ++ai)
#pragma unroll
        for (int bi = 0; bi <= ai; ++bi)
          if (bi == jb) {
            const int i = tr + kG * ai;
            if (i >= j)
              cb[i] = A[ai * (ai + 1) / 2 + bi];
          }
    }
    __syncthreads();
    const double d = cb[j];
    zeroPivot |= (d == 0.0);
    const double inv = 1.0 / d;
    if (tid > j && tid < n) {
      const double ct = cb[tid];
      for (int f = 0; f < F; ++f)
        rhs[f][tid] -= ct * (rhs[f][j] * inv);
    }
    double wl[R];
#pragma unroll
    for (int bi = 0; bi < R; ++bi) {
      const int l = tc + kG * bi;
      wl[bi] = (kG * bi + kG - 1 > j && l > j) ? cb[l] * inv : 0.0;
    }
#pragma unroll
    for (int ai = 0; ai < R; ++ai) {
      if (kG * ai + kG - 1 > j) {
        const int i = tr + kG * ai;
        const double ci = (i > j) ? cb[i] : 0.0;
#pragma unroll
        for (int bi = 0; bi <= ai; ++bi)
          if (kG * bi + kG - 1 > j)
            A[ai * (ai + 1) / 2 + bi] -= ci * wl[bi];
      }
    }
  }
  __syncthreads();

  // 5. back substitution, 16 rows per block, bottom-up
#pragma unroll
  for (int B = R - 1; B >= 0; --B) {
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
          pf[f] += c * rhs[f][i];
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
      double al[kG];
#pragma unroll
      for (int jj = kG - 1; jj >= 0; --jj) {
        const int j = kG * B + jj;
        double s = rhs[f][j] - spart[f][jj];
#pragma unroll
        for (int ii = jj + 1; ii < kG; ++ii)
          s -= dblk[ii][jj] * al[ii];
        al[jj] = s / dblk[jj][jj];
        rhs[f][j] = al[jj];
      }
    }
    __syncthreads();
  }

  if (tid == 0) {
    if (a.pvOut)
      a.pvOut[patch] = spv;
    if (a.status)
      a.status[patch] = 0;
  }
  if (zeroPivot && a.status)
    a.status[patch] = 1;

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
      const double phi =
          wendland<KIND, double>(dist3(yx, yy, yz, px[l], py[l], pz[l]),
                                 a.dinv);
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          acc[f] += phi * rhs[f][l];
    }
    const int slot = a.pairSlot[p];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      if (f < F) {
        // V beta (solver.hpp:272): inactive axes carry zero coefficients
        const double vb = scoef[f][0] + scoef[f][1] * yx + scoef[f][2] * yy +
                          scoef[f][3] * yz;
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
    const double d = dist3(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
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
      const double d = dist3(yx, yy, yz, a.cx[k], a.cy[k], a.cz[k]);
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
