// Workspace solve: the per-patch chain for patches of any size (the
// register-resident kernel of rbf_solve_tile.hpp takes n <= 192).
//
// One persistent workgroup (8 waves) per slot; each slot owns a private
// workspace in device memory (L2 / MALL resident for the patch sizes of
// uneven clouds) and pulls patches from a queue (largest first, so the tail
// of the launch is short). Per patch, with T the storage precision:
//
//   1. gather the coordinates and the right-hand sides b_f = v_f - Q beta_f
//      (FP64) into the slot's vectors;
//   2. assemble the lower block triangle of the bordered matrix
//      [[A, B^T], [B, I]] in 32 x 32 blocks (B = the F right-hand sides as
//      extra rows, as in the register kernel), A_ij = phi(|x_i - x_j| / delta)
//      evaluated in T;
//   3. factor A = L D L^T by 32-column panels: one wave eliminates the
//      diagonal block in FP64 registers (lane j holds column j of
//      [A_pp | I]; the identity becomes Linv = L_pp^-1; pivots d), every
//      wave turns the rows below into
//      L_Ip = A_Ip Linv^T D^-1 and applies A_IJ -= (L_Ip D) L_Jp^T to the
//      trailing blocks, both on 16 x 16 x 4 matrix cores in T. This is the
//      block form of the reference's unpivoted Doolittle LU
//      (src/rbf_pum/solver/utils/serial_solve.hpp:76-84): the same pivots
//      in the same order. The bordered rows end as y_f = D^-1 L^-1 b_f;
//   4. back substitution L^T alpha = y, block by block from the bottom with
//      the stored Linv blocks (FP64 arithmetic);
//   5. T = float only: iterative refinement in FP64. The residual
//      r = b - A alpha is matrix-free (phi re-evaluated in FP64 from the
//      coordinates), the correction solves L D L^T delta = r with the FP32
//      factors; it stops when the estimated remaining error
//      |delta_k| rho / (1 - rho), rho = |delta_k| / |delta_k-1|, is below
//      1e-12 |alpha|. A patch that does not converge within the iteration
//      limit (or whose FP32 pivots break down) is appended to the fallback
//      list and solved again by the FP64 variant;
//   6. evaluate t = E alpha + V beta at the member targets in FP64 (fixed
//      summation order) and write the (target, patch) slot
//      (src/rbf_pum/solver/solver.hpp:265-273).
//
// Pivots with d <= 1e-12 phi(0) (or not finite) mark a numerically singular
// patch (duplicate or near-duplicate sources): FP64 counts it, FP32 falls
// back to FP64.

#include <hip/hip_runtime.h>

#include "rbf_solve_common.hpp"

namespace pacman_amd {
namespace {

using namespace solve_detail;

constexpr int TB = kWsBlock;      // block edge
constexpr int NT = 512;           // threads per workgroup
constexpr int NW = NT / 64;       // waves
constexpr double kPivotTol = 1e-12;
constexpr double kIrTol = 1e-12;
constexpr int kLdsRows = 608;

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mma;
template <> struct Mma<double> {
  typedef dbl4 V;
  __device__ static V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row (within the 16 x 16 tile) of accumulator register v of lane l; the
  // column is l & 15
  __device__ static int row(int l, int v) { return (l >> 4) + 4 * v; }
};
template <> struct Mma<float> {
  typedef flt4 V;
  __device__ static V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int l, int v) { return 4 * (l >> 4) + v; }
};

// First element of block (I, J), J <= I, of the lower block triangle.
__device__ __forceinline__ int64_t blk(int I, int J) {
  return ((int64_t)I * (I + 1) / 2 + J) * (TB * TB);
}

template <int K, typename T>
__device__ __forceinline__ T phi_t(double xi, double yi, double zi, double xj,
                                   double yj, double zj, double dinv) {
  if constexpr (sizeof(T) == 8) {
    return wendland<K, double, true>(dist_fast(xi, yi, zi, xj, yj, zj), dinv);
  } else {
    const float dx = (float)(xj - xi), dy = (float)(yj - yi),
                dz = (float)(zj - zi);
    const float d = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    return wendland<K, float, true>(d, (float)dinv);
  }
}

template <int K>
__device__ __forceinline__ double phi64(double xi, double yi, double zi,
                                        double xj, double yj, double zj,
                                        double dinv) {
  return wendland<K, double, true>(dist_fast(xi, yi, zi, xj, yj, zj), dinv);
}

// Workgroup maxima of two values (every thread gets both).
__device__ void block_max2(double &u, double &v, double (*red)[2]) {
  u = wave_max(u);
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wave][0] = u;
    red[wave][1] = v;
  }
  __syncthreads();
  u = red[0][0];
  v = red[0][1];
  for (int w = 1; w < NW; ++w) {
    u = fmax(u, red[w][0]);
    v = fmax(v, red[w][1]);
  }
}

// SMALL: the bordered matrix has at most kLdsRows rows, so coordinates,
// alpha, the pivots and the solve vectors live in LDS (about 138 KB for
// T = float); otherwise in the slot's vectors in device memory.
template <typename T, int K, bool SMALL>
__global__ __launch_bounds__(NT, 2) void ws_solve_kernel(WsArgs w) {
  using M = Mma<T>;
  using V = typename M::V;
  const SolveArgs &a = w.a;
  __shared__ int sItem, sFlag;
  __shared__ T linv[TB][TB + 1];      // Linv of the current panel
  __shared__ double sd[TB], sdinv[TB];
  __shared__ double scoef[kFMax][4];
  __shared__ double red[NW][2];
  constexpr bool FP32 = sizeof(T) == 4;
  constexpr int LR = SMALL ? kLdsRows : 1;
  constexpr int LA = SMALL ? kFMax * kLdsRows : 1;
  constexpr int LZ = SMALL && FP32 ? kFMax * kLdsRows : 1;
  __shared__ double lxyz[3][LR];
  __shared__ double lal[LA];   // alpha
  __shared__ double lsv[LA];   // solve vector
  __shared__ double lzv[LZ];   // correction; residual partial sums before
  __shared__ double ldv[LR];   // pivots

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int np = w.nbMax * TB;  // vector stride of the slot
  T *W = static_cast<T *>(w.work) + (int64_t)blockIdx.x * w.slotBlocks;
  T *Li = W + w.linvOffset;  // stored Linv blocks, TB x TB each
  double *V0 = w.vec + (int64_t)blockIdx.x * w.slotVec;
  double *gx = V0, *gy = V0 + np, *gz = V0 + 2 * np, *dv = V0 + 3 * np;
  double *bv = V0 + 4 * np;                   // [kFMax][np] b - Q beta
  double *al = bv + (int64_t)kFMax * np;      // [kFMax][np] alpha
  double *sv = al + (int64_t)kFMax * np;      // [kFMax][np] work vector
  double *zv = sv + (int64_t)kFMax * np;      // [kFMax][np] work vector
  const int count = w.countPtr ? *w.countPtr : a.count;
  // solve vectors: [F][vst] with stride vst, pivots vd
  double *const vs = SMALL ? lsv : sv;
  double *const vz = SMALL && FP32 ? lzv : zv;
  double *const vd = SMALL ? ldv : dv;
  constexpr int vstC = SMALL ? kLdsRows : 0;
  const int vst = SMALL ? vstC : np;

  for (;;) {
    if (tid == 0)
      sItem = atomicAdd(&w.counters[w.queueIdx], 1);
    __syncthreads();
    const int item = sItem;
    if (item >= count)
      break;
    const int patch = a.patches[item];
    const int64_t s0 = a.srcOff[patch];
    const int n = (int)(a.srcOff[patch + 1] - s0);
    const int F = a.F;
    const int NB = (n + F + TB - 1) / TB, P = (n + TB - 1) / TB;
    const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];
    const double dinv = a.patchDinv ? a.patchDinv[patch] : a.dinv;
    double *px, *py, *pz, *alp;  // coordinates; alpha with stride ast
    int ast;
    if constexpr (SMALL) {
      px = lxyz[0];
      py = lxyz[1];
      pz = lxyz[2];
      alp = lal;
      ast = kLdsRows;
    } else {
      px = gx;
      py = gy;
      pz = gz;
      alp = al;
      ast = np;
    }
    if (tid < kFMax * 4)
      scoef[tid >> 2][tid & 3] =
          (tid >> 2) < F
              ? w.coef[((int64_t)patch * F + (tid >> 2)) * 4 + (tid & 3)]
              : 0.0;
    if (tid == 0)
      sFlag = 0;
    __syncthreads();  // sItem read by all; scoef ready

    // 1. gather
    for (int t = tid; t < NB * TB; t += NT) {
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
      for (int f = 0; f < F; ++f)
        bv[(int64_t)f * np + t] =
            t < n ? a.vals[(int64_t)f * a.N + p] -
                        (scoef[f][0] + scoef[f][1] * (x - cx) +
                         scoef[f][2] * (y - cy) + scoef[f][3] * (z - cz))
                  : 0.0;
    }
    __syncthreads();

    // 2. assemble the blocks (I, J), J < P, of the bordered matrix
    {
      int I = 0, J = 0;
      const int64_t nblk = (int64_t)NB * (NB + 1) / 2;
      for (int64_t bi = 0; bi < nblk; ++bi) {
        if (bi > 0) {
          if (++J > I) {
            ++I;
            J = 0;
          }
        }
        if (J >= P)
          continue;
        T *B = W + bi * (TB * TB);
        for (int rc = tid; rc < TB * TB; rc += NT) {
          const int i = I * TB + (rc >> 5), j = J * TB + (rc & 31);
          T v;
          if (i < n && j < n)
            v = phi_t<K, T>(px[i], py[i], pz[i], px[j], py[j], pz[j], dinv);
          else if (i >= n && i < n + F && j < n)
            v = (T)bv[(int64_t)(i - n) * np + j];
          else
            v = (i == j) ? T(1) : T(0);
          B[rc] = v;
        }
      }
    }
    __syncthreads();

    // 3. factorisation by panels
    for (int p = 0; p < P; ++p) {
      const int kk = min(TB, n - TB * p);
      const int64_t bpp = blk(p, p);
      if (wave == 0) {
        // [A_pp | I] in registers: lane j owns column j of it (0..31: the
        // block, 32..63: the identity, which the row operations turn into
        // Linv), e[r] = row r. Pivot k: the multipliers A_ik / d come from
        // lane k (readlane), the pivot row is each lane's own e[k]; every
        // index is a compile-time constant.
        double e[TB];
#pragma unroll
        for (int r = 0; r < TB; ++r)
          e[r] = lane < TB ? (double)W[bpp + r * TB + lane]
                           : (lane - TB == r ? 1.0 : 0.0);
        bool bad = false;
        static_for<0, TB>([&](auto KC) {
          constexpr int k = decltype(KC)::value;
          if (k < kk) {
            const double d = readlane_f64(e[k], k);
            const double dinv = 1.0 / d;
            bad = bad || !(d > kPivotTol * w.phi0);
            if (lane == 0) {
              sd[k] = d;
              sdinv[k] = dinv;
            }
            static_for<k + 1, TB>([&](auto IC) {
              constexpr int i = decltype(IC)::value;
              if (i < kk) {
                const double m = readlane_f64(e[i], k) * dinv;
                e[i] = fma(-m, e[k], e[i]);
              }
            });
          }
        });
        if (lane == 0 && bad)
          sFlag = 1;
        if (lane < TB && lane >= kk) {
          sd[lane] = 1.0;
          sdinv[lane] = 0.0;
        }
        if (lane >= TB) {
          const int c = lane - TB;
#pragma unroll
          for (int i = 0; i < TB; ++i) {
            const T val = (i < kk && c < kk) ? (T)e[i] : T(0);
            linv[i][c] = val;
            Li[(int64_t)p * TB * TB + i * TB + c] = val;
          }
        }
      }
      __syncthreads();
      if (tid < kk)
        vd[TB * p + tid] = sd[tid];
      // rows below the pivots: L = A Linv^T D^-1, one 32-row block per wave
      // step, every operand load issued before the first product
      const int firstRow = kk < TB ? p : p + 1;
      for (int I = firstRow + wave; I < NB; I += NW) {
        T *B = W + blk(I, p);
        T av[2][TB / 4];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int q = 0; q < TB / 4; ++q)
            av[h][q] = B[(16 * h + (lane & 15)) * TB + 4 * q + (lane >> 4)];
        V acc[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int c2 = 0; c2 < 2; ++c2)
            acc[h][c2] = V{0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < TB / 4; ++q) {
          const int kq = 4 * q + (lane >> 4);
          const T l0 = linv[lane & 15][kq], l1 = linv[16 + (lane & 15)][kq];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            acc[h][0] = M::mma(av[h][q], l0, acc[h][0]);
            acc[h][1] = M::mma(av[h][q], l1, acc[h][1]);
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 16 * h + M::row(lane, v), c = lane & 15;
            if (TB * I + r >= TB * p + kk) {
              B[r * TB + c] = (T)(acc[h][0][v] * sdinv[c]);
              B[r * TB + 16 + c] = (T)(acc[h][1][v] * sdinv[16 + c]);
            }
          }
      }
      __syncthreads();
      if (p + 1 < P) {
        // trailing update of blocks (I, J), p < J < P, J <= I < NB, one
        // block per wave step: A_IJ -= (L_Ip D) L_Jp^T for all four 16 x 16
        // tiles (the upper one of a diagonal block included: the next
        // elimination reads whole rows of the diagonal block), every operand
        // loaded before the first product
        int64_t total = 0;
        for (int J = p + 1; J < P; ++J)
          total += NB - J;
        for (int64_t it = wave; it < total; it += NW) {
          int64_t rem = it;
          int J = p + 1;
          while (rem >= NB - J) {
            rem -= NB - J;
            ++J;
          }
          const int I = J + (int)rem;
          T *C = W + blk(I, J);
          const T *LI = W + blk(I, p), *LJ = W + blk(J, p);
          T a[2][TB / 4], b[2][TB / 4];
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < TB / 4; ++q) {
              const int o = (16 * h + (lane & 15)) * TB + 4 * q + (lane >> 4);
              a[h][q] = LI[o];
              b[h][q] = LJ[o];
            }
          V c[2][2];
#pragma unroll
          for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                c[ti][tj][v] = C[(16 * ti + M::row(lane, v)) * TB + 16 * tj +
                                 (lane & 15)];
#pragma unroll
          for (int q = 0; q < TB / 4; ++q) {
            const T dk = (T)(-sd[4 * q + (lane >> 4)]);
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) {
              const T av = a[ti][q] * dk;
#pragma unroll
              for (int tj = 0; tj < 2; ++tj)
                c[ti][tj] = M::mma(av, b[tj][q], c[ti][tj]);
            }
          }
#pragma unroll
          for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
              for (int v = 0; v < 4; ++v)
                C[(16 * ti + M::row(lane, v)) * TB + 16 * tj + (lane & 15)] =
                    c[ti][tj][v];
        }
      }
      __syncthreads();
    }

    auto Lat = [&](int i, int c) -> double {
      return (double)W[blk(i / TB, c / TB) + (i % TB) * TB + (c % TB)];
    };
    // L^T x = s (s destroyed), block by block from the bottom; x = dst.
    // Vectors are [F][stride].
    auto back_subst = [&](double *s, int sStride, double *dst, int dstStride) {
      for (int B = P - 1; B >= 0; --B) {
        const int kb = min(TB, n - TB * B);
        const T *LB = Li + (int64_t)B * TB * TB;
        for (int t = tid; t < F * TB; t += NT) {
          const int f = t / TB, cl = t % TB;
          const double *sb = s + (int64_t)f * sStride + TB * B;
          double acc = 0.0;
          // Linv is lower triangular with zeros past kb: every load is in
          // range and the masked terms add exact zeros
#pragma unroll 8
          for (int c2 = 0; c2 < TB; ++c2)
            acc = fma((double)LB[c2 * TB + cl], c2 < kb ? sb[c2] : 0.0, acc);
          if (cl < kb)
            dst[(int64_t)f * dstStride + TB * B + cl] = acc;
        }
        __syncthreads();
        const int wdt = TB * B;
        for (int t = tid; t < F * wdt; t += NT) {
          const int f = t / wdt, c = t % wdt;
          // column c of the block row B: L[TB B + r][c], r < kb
          const T *col = W + blk(B, c / TB) + (c % TB);
          const double *xb = dst + (int64_t)f * dstStride + TB * B;
          double acc = 0.0;
#pragma unroll 8
          for (int r = 0; r < TB; ++r)
            acc = fma(r < kb ? (double)col[r * TB] : 0.0, r < kb ? xb[r] : 0.0,
                      acc);
          s[(int64_t)f * sStride + c] -= acc;
        }
        __syncthreads();
      }
    };
    // L x = s (s destroyed), block by block from the top; x = dst (both
    // [F][stride]).
    auto fwd_subst = [&](double *s, double *dst, int stride) {
      for (int B = 0; B < P; ++B) {
        const int kb = min(TB, n - TB * B);
        const T *LB = Li + (int64_t)B * TB * TB;
        for (int t = tid; t < F * TB; t += NT) {
          const int f = t / TB, cl = t % TB;
          const double *sb = s + (int64_t)f * stride + TB * B;
          double acc = 0.0;
#pragma unroll 8
          for (int c2 = 0; c2 < TB; ++c2)
            acc = fma((double)LB[cl * TB + c2], c2 < kb ? sb[c2] : 0.0, acc);
          if (cl < kb)
            dst[(int64_t)f * stride + TB * B + cl] = acc;
        }
        __syncthreads();
        const int r0 = TB * (B + 1), rows = n - r0;
        for (int t = tid; t < F * max(rows, 0); t += NT) {
          const int f = t / rows, i = r0 + t % rows;
          // row i of the block column B: L[i][TB B + c], c < kb
          const T *row = W + blk(i / TB, B) + (i % TB) * TB;
          const double *xb = dst + (int64_t)f * stride + TB * B;
          double acc = 0.0;
#pragma unroll 8
          for (int c = 0; c < TB; ++c)
            acc = fma(c < kb ? (double)row[c] : 0.0, c < kb ? xb[c] : 0.0, acc);
          s[(int64_t)f * stride + i] -= acc;
        }
        __syncthreads();
      }
    };

    // 4. first solve: y_f = bordered row n + f
    for (int t = tid; t < F * n; t += NT) {
      const int f = t / n, c = t % n;
      vs[(int64_t)f * vst + c] = Lat(n + f, c);
    }
    __syncthreads();
    back_subst(vs, vst, alp, ast);

    int status = 0;
    if constexpr (FP32) {
      // 5. iterative refinement in FP64
      bool ok = sFlag == 0;
      double prev = 0.0;
      int iters = 0;
      bool conv = false;
      while (ok && !conv && iters < w.maxIR) {
        ++iters;
        // residual r = b - A alpha (phi in FP64); in LDS the j range of
        // every row is split over S parts so that all waves work, and the
        // parts' sums (in vz, free until the forward solve) are added in
        // part order: S F n <= max(NT F, 8 n) <= LZ
        const int S = SMALL ? max(1, NT / n) : 1;
        const int chunk = (n + S - 1) / S;
        for (int u = tid; u < n * S; u += NT) {
          const int i = u % n, part = u / n;
          const int j0 = part * chunk, j1 = min(n, j0 + chunk);
          double acc[kFMax];
#pragma unroll
          for (int f = 0; f < kFMax; ++f)
            acc[f] = 0.0;
          const double xi = px[i], yi = py[i], zi = pz[i];
#pragma unroll 4
          for (int j = j0; j < j1; ++j) {
            const double ph = phi64<K>(xi, yi, zi, px[j], py[j], pz[j], dinv);
#pragma unroll
            for (int f = 0; f < kFMax; ++f)
              if (f < F)
                acc[f] = fma(ph, alp[f * ast + j], acc[f]);
          }
          for (int f = 0; f < F; ++f) {
            if constexpr (SMALL)
              vz[((int64_t)part * F + f) * n + i] = acc[f];
            else
              vs[(int64_t)f * vst + i] = bv[(int64_t)f * np + i] - acc[f];
          }
        }
        __syncthreads();
        if constexpr (SMALL) {
          for (int t = tid; t < F * n; t += NT) {
            const int f = t / n, i = t % n;
            double sum = 0.0;
            for (int part = 0; part < S; ++part)
              sum += vz[((int64_t)part * F + f) * n + i];
            vs[(int64_t)f * vst + i] = bv[(int64_t)f * np + i] - sum;
          }
          __syncthreads();
        }
        fwd_subst(vs, vz, vst);
        for (int t = tid; t < F * n; t += NT) {
          const int f = t / n, c = t % n;
          vs[(int64_t)f * vst + c] = vz[(int64_t)f * vst + c] / vd[c];
        }
        __syncthreads();
        back_subst(vs, vst, vz, vst);
        double dmax = 0.0, amax = 0.0;
        for (int t = tid; t < F * n; t += NT) {
          const int f = t / n, c = t % n;
          const double dl = vz[(int64_t)f * vst + c];
          const double x = alp[f * ast + c] + dl;
          alp[f * ast + c] = x;
          dmax = fmax(dmax, fabs(dl));
          amax = fmax(amax, fabs(x));
        }
        block_max2(dmax, amax, red);
        if (!(dmax <= 1e300) || !(amax <= 1e300)) {
          ok = false;
        } else if (dmax == 0.0) {
          conv = true;
        } else if (iters >= 2) {
          const double rho = dmax / prev;
          if (rho >= 0.5)
            ok = false;
          else if (dmax * rho / (1.0 - rho) <= kIrTol * amax)
            conv = true;
        }
        prev = dmax;
        __syncthreads();
      }
      if (tid == 0)
        atomicAdd(&w.counters[kWsIrIters], iters);
      status = iters << 8;
      if (!conv) {
        if (tid == 0) {
          const int slot = atomicAdd(&w.counters[kWsFallbacks], 1);
          w.fallback[slot] = patch;
          atomicAdd(&w.counters[kWsFallbackTotal], 1);
          if (a.status)
            a.status[patch] = status | (1 << 16);
        }
        __syncthreads();
        continue;  // solved again in FP64
      }
    } else {
      if (sFlag) {
        status = 1;
        if (tid == 0)
          atomicAdd(&w.counters[kWsSingular], 1);
      }
    }

    if (a.alphaOut)
      for (int t = tid; t < F * n; t += NT) {
        const int f = t / n, c = t % n;
        a.alphaOut[(s0 + c) * F + f] = alp[f * ast + c];
      }
    if (a.alphaHeavy) {
      // evaluated in chunks across many waves; alpha of source c at
      // ah[c * S + f], zero past F
      double *ah = a.alphaHeavy + 8 * s0;
      const int S = alpha_stride(F);
      for (int t = tid; t < S * n; t += NT) {
        const int f = t / n, c = t % n;
        ah[c * S + f] = f < F ? alp[f * ast + c] : 0.0;
      }
      if (tid == 0 && a.status)
        a.status[patch] = status;
      __syncthreads();
      continue;
    }
    // 6. evaluation at the member targets
    const int64_t p0 = a.pairOff[patch];
    const int mall = (int)(a.pairOff[patch + 1] - p0);
    for (int q = tid; q < mall; q += NT) {
      const int e = a.pairTgt[p0 + q];
      if (((unsigned)e >> 31) == 0)
        continue;
      const int tp = e & 0x7fffffff;
      const double yx = a.tx[tp], yy = a.ty[tp], yz = a.tz[tp];
      double acc[kFMax];
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        acc[f] = 0.0;
#pragma unroll 4
      for (int j = 0; j < n; ++j) {
        const double ph = phi64<K>(yx, yy, yz, px[j], py[j], pz[j], dinv);
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          if (f < F)
            acc[f] = fma(ph, alp[f * ast + j], acc[f]);
      }
      const int slot = a.pairSlot[p0 + q];
      for (int f = 0; f < F; ++f) {
        const double vb = scoef[f][0] + scoef[f][1] * (yx - cx) +
                          scoef[f][2] * (yy - cy) + scoef[f][3] * (yz - cz);
        a.contrib[(int64_t)slot * F + f] = acc[f] + vb;
      }
    }
    if (tid == 0 && a.status)
      a.status[patch] = status;
    __syncthreads();
  }
}

// mode 2: every counter; 1: queues and fallback list; 0: first-pass queue
__global__ void ws_reset_kernel(int *counters, int n, int mode) {
  const int i = threadIdx.x;
  if (i < n && (mode == 2 || i == kWsQueue ||
                (mode == 1 && (i == kWsQueue2 || i == kWsFallbacks))))
    counters[i] = 0;
}

template <typename T, bool SMALL>
void launch_ws_t(int kind, const WsArgs &w, int grid, hipStream_t s) {
  switch (kind) {
  case 0: ws_solve_kernel<T, 0, SMALL><<<grid, NT, 0, s>>>(w); break;
  case 1: ws_solve_kernel<T, 1, SMALL><<<grid, NT, 0, s>>>(w); break;
  case 2: ws_solve_kernel<T, 2, SMALL><<<grid, NT, 0, s>>>(w); break;
  case 3: ws_solve_kernel<T, 3, SMALL><<<grid, NT, 0, s>>>(w); break;
  default: ws_solve_kernel<T, 4, SMALL><<<grid, NT, 0, s>>>(w); break;
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace

int64_t ws_slot_blocks(int nbMax) {
  return ((int64_t)nbMax * (nbMax + 1) / 2 + nbMax) * (TB * TB);
}
int64_t ws_linv_offset(int nbMax) {
  return (int64_t)nbMax * (nbMax + 1) / 2 * (TB * TB);
}
int ws_lds_block_rows() { return kLdsRows / TB; }

int64_t ws_slot_vec(int nbMax) {
  return (int64_t)(4 + 4 * kFMax) * nbMax * TB;
}

void launch_ws_reset(int *counters, bool all, hipStream_t s,
                     bool keepFallbacks) {
  ws_reset_kernel<<<1, 64, 0, s>>>(counters, kWsCounters,
                                   all ? 2 : keepFallbacks ? 0 : 1);
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_ws_solve(bool fp32, int kind, const WsArgs &w, int grid,
                     hipStream_t s) {
  if (grid <= 0)
    return;
  if (w.a.F < 1 || w.a.F > kFMax)
    throw std::runtime_error("workspace solve: 1 <= fields <= 8 per launch");
  // the LDS variant covers slots of up to kLdsRows rows
  const bool small = w.nbMax * TB <= kLdsRows;
  if (fp32) {
    if (small)
      launch_ws_t<float, true>(kind, w, grid, s);
    else
      launch_ws_t<float, false>(kind, w, grid, s);
  } else {
    if (small)
      launch_ws_t<double, true>(kind, w, grid, s);
    else
      launch_ws_t<double, false>(kind, w, grid, s);
  }
}

} // namespace pacman_amd
