// Blocked per-patch solve on 16 x 16 tiles with FP64 matrix-core updates.
//
// One workgroup of NW waves per patch (8, or 16 for the larger bins). The
// patch matrix A (n x n, n <= 16R) is held in registers as its
// lower-triangle 16 x 16 tiles (I, J), I >= J, in the accumulator layout of
// v_mfma_f64_16x16x4f64: lane l, register v holds element
// (16 I + (l >> 4) + 4 v, 16 J + (l & 15)) of tile (I, J). Waves 1..NW-1
// (the tile waves) hold the tiles, dealt in column-major order, cyclically,
// so every wave keeps work while the trailing matrix shrinks. Wave 0 (the
// control wave) holds none: it eliminates the diagonal blocks. With 16 waves
// (one workgroup per CU, 128 registers per lane) bins of up to R = 17 keep
// every tile in registers; 8 waves run two workgroups per CU.
//
// The right-hand sides ride along as F extra rows n .. n+F-1 holding the
// residual b - Q beta (the bordered matrix [[A, B^T], [B, I]]): the
// factorisation then leaves L[n + f][j] = (D^-1 L^-1 b_f)_j in those rows,
// exactly the right-hand side of the back substitution
// L^T alpha = D^-1 L^-1 b. The extra rows never pivot, and all of them take
// the same arithmetic wherever they sit.
//
// Factorisation: A = L D L^T by 16-column panels, the block form of the
// reference's unpivoted Doolittle LU (src/rbf_pum/solver/utils/
// serial_solve.hpp:76-84): the same pivots in the same order. The control
// wave eliminates the diagonal block of panel p in registers (one row per
// lane, pivot rows broadcast by DPP), applying the same
// row operations to an identity, which yields Linv = L_pp^-1 and the
// pivots d. Per
// panel, three barriers:
//   A. the tile waves turn their panel tiles into X = A_Ip Linv^T = L_Ip D
//      (four MFMAs each) and publish W = X D^-1 = L_Ip to LDS;
//   B. the owner of the next diagonal tile applies the panel's update to it
//      first and hands it to the control wave;
//   C. the control wave eliminates that block while the tile waves update
//      their trailing tiles, A_IJ -= (W_I D) W_J^T (four MFMAs each), and
//      publish the next panel's tiles: the sequential elimination overlaps
//      the trailing update.
// Back substitution over 16-row blocks from the bottom, one barrier per
// block: the control wave computes alpha_B = Linv_B^T x_B for all fields at
// once as 16 x 16 x 16 MFMA products (with the sub-diagonal tile staged in
// LDS), the tile waves add
// L_IJ^T alpha_I to the sums of the blocks further up as soon as alpha_I is
// known. The forward kernel then stores alpha, one record per source, for
// the evaluation at the member targets (src/rbf_pum/solver/solver.hpp:
// 265-273), which runs apart in wave_eval_kernel (rbf_solve.hip).
#pragma once

#include "rbf_solve_common.hpp"

namespace pacman_amd {
namespace {

using namespace solve_detail;

typedef double dbl4 __attribute__((ext_vector_type(4)));

// 1 / d: hardware reciprocal refined by two Newton steps (about 1 ulp).
__device__ __forceinline__ double fast_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// x += (x of lane K of this 16-lane row) * nl, as one FP64 FMA with a DPP
// row_newbcast source (the compiler forms a separate DPP move and FMA through
// one temporary, which serialises a pivot step). The x read through DPP was
// last written by the previous pivot step, several instructions earlier.
template <int K>
__device__ __forceinline__ void fmac_row_bcast(double &x, double nl) {
  asm volatile("v_fmac_f64_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf "
               "bank_mask:0xf"
               : "+v"(x)
               : "v"(nl), "i"(K));
}
// x of lane K of this 16-lane row; the two wait states cover a VALU write
// of x just before.
template <int K> __device__ __forceinline__ double row_bcast_asm(double x) {
  double r;
  asm volatile("s_nop 1\n\tv_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf "
               "bank_mask:0xf"
               : "=v"(r)
               : "v"(x), "i"(K));
  return r;
}

// Owners of the lower-triangle 16x16 tiles among the NWT tile waves. Tile
// (J+1, J) and the diagonal tile (J+1, J+1) go to the same wave: that wave
// computes W_{J+1} during panel J and applies it to the diagonal tile at
// once, so the control wave gets that tile without a further barrier. Pairs
// are dealt first, then the tile (0, 0), then the other tiles, each to the
// least-loaded wave; every wave lists its tiles column by column.
constexpr int kMaxTileWaves = 15;     // tile waves of a 16-wave workgroup
constexpr int kMaxTilesPerWave = 16;  // R <= 13 at 7 tile waves, 17 at 15

struct TileTable {
  int I[kMaxTileWaves][kMaxTilesPerWave], J[kMaxTileWaves][kMaxTilesPerWave],
      cnt[kMaxTileWaves];
};

constexpr TileTable make_tile_table(int R, int NWT) {
  TileTable t{};
  int load[kMaxTileWaves] = {};
  auto pick = [&]() {
    int w = 0;
    for (int x = 1; x < NWT; ++x)
      if (load[x] < load[w])
        w = x;
    return w;
  };
  auto put = [&](int w, int I, int J) {
    t.I[w][t.cnt[w]] = I;
    t.J[w][t.cnt[w]] = J;
    ++t.cnt[w];
    ++load[w];
  };
  for (int J = 0; J + 1 < R; ++J) {
    const int w = pick();
    put(w, J + 1, J);
    put(w, J + 1, J + 1);
  }
  put(pick(), 0, 0);
  for (int J = 0; J < R; ++J)
    for (int I = J + 2; I < R; ++I)
      put(pick(), I, J);
  for (int w = 0; w < NWT; ++w)  // column-major order within a wave
    for (int a = 1; a < t.cnt[w]; ++a)
      for (int b = a; b > 0; --b) {
        const bool less = t.J[w][b] < t.J[w][b - 1] ||
                          (t.J[w][b] == t.J[w][b - 1] &&
                           t.I[w][b] < t.I[w][b - 1]);
        if (!less)
          break;
        const int ti = t.I[w][b], tj = t.J[w][b];
        t.I[w][b] = t.I[w][b - 1];
        t.J[w][b] = t.J[w][b - 1];
        t.I[w][b - 1] = ti;
        t.J[w][b - 1] = tj;
      }
  return t;
}

constexpr int max_tiles_per_wave(const TileTable &t, int NWT) {
  int m = 1;
  for (int w = 0; w < NWT; ++w)
    m = t.cnt[w] > m ? t.cnt[w] : m;
  return m;
}

template <int R, int NW_ = 8> struct TileShape {
  static constexpr int NMAX = 16 * R;
  static constexpr int NW = NW_;            // waves per workgroup
  static constexpr int NT = 64 * NW;
  static constexpr int NWT = NW - 1;        // tile waves
  static constexpr int NTILE = R * (R + 1) / 2;
  static constexpr TileTable TAB = make_tile_table(R, NWT);
  static constexpr int TPW = max_tiles_per_wave(TAB, NWT);
  static_assert(NWT <= kMaxTileWaves, "tile waves");
  static_assert(TPW <= kMaxTilesPerWave, "tiles per wave");
  static constexpr int LDW = NMAX + 4;      // panel column stride (doubles)
  static constexpr int PANEL = 2 * 16 * LDW;
  static constexpr int BACK = (R - 1) * 16 * 17 + R * kFMax * 16;
  // the transpose chain stages its right-hand side in the same buffer:
  // target coordinates [3][64], h [kFMax][64], second halves [kFMax][256]
  static constexpr int ADJRHS = (3 + kFMax) * 64 + kFMax * 256;
  static constexpr int SBUF0 = PANEL > BACK ? PANEL : BACK;
  static constexpr int SBUF = SBUF0 > ADJRHS ? SBUF0 : ADJRHS;
};

// ADJ: the transpose chain. The right-hand side is E^T h with h = w g at
// the member targets (instead of the residual b - Q beta), and the kernel
// stores gamma = A^-1 E^T h per patch source instead of evaluating (8 waves
// only: its right-hand-side pass splits 512 threads into two halves).
template <int R, bool ADJ, int NW>
__global__ __launch_bounds__(64 * NW, 4) void solve_tile_kernel(
    SolveArgs a, const double *coef, int kind) {
  static_assert(!ADJ || NW == 8, "the transpose chain runs 8 waves");
  using TS = TileShape<R, NW>;
  constexpr int NMAX = TS::NMAX;
  constexpr int NT = TS::NT;
  constexpr int NWT = TS::NWT;
  constexpr int NTILE = TS::NTILE;
  constexpr int TPW = TS::TPW;
  constexpr int LDW = TS::LDW;

  __shared__ double px[NMAX], py[NMAX], pz[NMAX];
  __shared__ double rhs[kFMax][NMAX];
  __shared__ double sbuf[TS::SBUF];  // panels, back substitution, rhs (ADJ)
  __shared__ double linv[2][16][17];         // Linv of the diagonal block
  __shared__ double linvAll[R][120];         // strictly-lower Linv, packed
  __shared__ double dtile[16][17];           // diagonal tile hand-over
  __shared__ double dpan[2][16], dinvp[2][16];
  __shared__ double scoef[kFMax][4];
  __shared__ int spivot;
  double(*wp)[16][LDW] = reinterpret_cast<double(*)[16][LDW]>(sbuf);

  const int patch = a.patches[xcd_remap(blockIdx.x, a.count)];
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // lane row / column in a 16 x 16 tile (not const: the tile waves launder
  // them once per panel phase, below)
  int lr = lane >> 4, lc = lane & 15;
  const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];
  const double dinv = a.patchDinv ? a.patchDinv[patch] : a.dinv;

  // gather; right-hand side = value - polynomial fit. Thread t holds source
  // row t. The fit coefficients, its coordinates and every field's value are
  // loaded into registers first, all in flight together; the barrier below
  // publishes the coefficients, then the rows are formed and written to LDS.
  static_assert(NMAX <= NT, "one source row per thread");
  double cf = 0.0;
  if (!ADJ && tid < kFMax * 4 && (tid >> 2) < F)
    cf = coef[((int64_t)patch * F + (tid >> 2)) * 4 + (tid & 3)];
  double x = 0.0, y = 0.0, z = 0.0;
  double gv[kFMax];
#pragma unroll
  for (int f = 0; f < kFMax; ++f)
    gv[f] = 0.0;
  if (tid < n) {
    const int p = a.srcIdx[s0 + tid];
    x = a.sx[p];
    y = a.sy[p];
    z = a.sz[p];
    if constexpr (!ADJ) {
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          gv[f] = a.vals[(int64_t)f * a.N + p];
    }
  }
  if (!ADJ && tid < kFMax * 4)
    scoef[tid >> 2][tid & 3] = cf;
  if (tid == 0)
    spivot = 0;
  __syncthreads();  // scoef, spivot
  if (tid < NMAX) {
    px[tid] = x;
    py[tid] = y;
    pz[tid] = z;
#pragma unroll
    for (int f = 0; f < kFMax; ++f) {
      if (f < F) {
        double v = 0.0;
        if (!ADJ && tid < n)
          v = gv[f] - (scoef[f][0] + scoef[f][1] * (x - cx) +
                       scoef[f][2] * (y - cy) + scoef[f][3] * (z - cz));
        rhs[f][tid] = v;
      }
    }
  }
  __syncthreads();

  const int NB = (n + F + 15) >> 4;  // row blocks holding data
  const int P = (n + 15) >> 4;       // pivot panels

  // fn(integral_constant<int, kind>): one code path per Wendland kind, so
  // the kernel evaluations inside the loops carry no branch.
  auto with_kind = [&](auto &&fn) __attribute__((always_inline)) {
    switch (kind) {
    case 0:
      fn(std::integral_constant<int, 0>{});
      break;
    case 1:
      fn(std::integral_constant<int, 1>{});
      break;
    case 2:
      fn(std::integral_constant<int, 2>{});
      break;
    case 3:
      fn(std::integral_constant<int, 3>{});
      break;
    default:
      fn(std::integral_constant<int, 4>{});
      break;
    }
  };

  // Tile of the bordered, padded matrix: A among the patch rows, the
  // residual rows below them, identity on the padding.
  auto assemble = [&](auto kc, int I, int J) __attribute__((always_inline)) {
    constexpr int K = decltype(kc)::value;
    dbl4 t4;
    const int j = 16 * J + lc;
    const double xj = px[j], yj = py[j], zj = pz[j];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * I + lr + 4 * v;
      double val;
      if (i < n && j < n)
        val = wendland<K, double, true>(
            dist_fast(px[i], py[i], pz[i], xj, yj, zj), dinv);
      else if (i >= n && i < n + F && j < n)
        val = rhs[i - n][j];
      else
        val = (i == j) ? 1.0 : 0.0;
      t4[v] = val;
    }
    return t4;
  };

  // t - (W_I D) W_J^T for the panel in buffer `buf` (four 16x16x4 MFMAs;
  // operand lane (i + 16 q) holds column 4 s + q of row i).
  auto schur = [&](dbl4 t4, int I, int J, int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * s + lr;
      const double av = -wp[buf][k][16 * I + lc] * dpan[buf][k];
      const double bv = wp[buf][k][16 * J + lc];
      t4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, t4, 0, 0, 0);
    }
    return t4;
  };

  // W = (A_rows Linv^T) D^-1 for 16 raw rows stored column-major at
  // base[k * stride + i]; zero past the pivots.
  auto to_w = [&](const double *base, int stride, int buf, int kk) __attribute__((always_inline)) {
    dbl4 t4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int k = 4 * s + lr;
      t4 = __builtin_amdgcn_mfma_f64_16x16x4f64(base[k * stride + lc],
                                                linv[buf][lc][k], t4, 0, 0, 0);
    }
    const double sc = lc < kk ? dinvp[buf][lc] : 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      t4[v] *= sc;
    return t4;
  };

  if constexpr (ADJ) {
    // right-hand side E^T h, h = w g at the member targets: tiles of 64
    // pairs staged in LDS; thread (j, half) sums half of each tile for
    // source j; the halves are added in a fixed order.
    static_assert(TS::SBUF >= TS::ADJRHS, "rhs staging fits the buffer");
    double *sty = sbuf;                        // [3][64] target coordinates
    double *sth = sbuf + 3 * 64;               // [kFMax][64] h
    double *sred = sbuf + (3 + kFMax) * 64;    // [kFMax][256] second halves
    const int64_t p0 = a.pairOff[patch];
    const int mall = (int)(a.pairOff[patch + 1] - p0);
    const int j = tid & 255, half = tid >> 8;
    double acc[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      acc[f] = 0.0;
    const int jc = j < n ? j : 0;
    const double xj = px[jc], yj = py[jc], zj = pz[jc];
    for (int tile = 0; tile < mall; tile += 64) {
      if (tid < 64) {
        const int e = tile + tid;
        double y0 = 0.0, y1 = 0.0, y2 = 0.0, w = 0.0;
        int tp = 0;
        if (e < mall) {
          const int pe = a.pairTgt[p0 + e];
          tp = pe & 0x7fffffff;
          y0 = a.tx[tp];
          y1 = a.ty[tp];
          y2 = a.tz[tp];
          w = a.wslot[a.pairSlot[p0 + e]];  // 0 unless a member slot
        }
        const int orow = a.opRow0 >= 0 ? e - a.opRow0 : -1;
        sty[tid] = y0;
        sty[64 + tid] = y1;
        sty[128 + tid] = y2;
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          sth[f * 64 + tid] =
              a.opRow0 >= 0 ? ((f < F && orow == f) ? w : 0.0)
              : (f < F && w != 0.0) ? w * a.gSorted[f * a.M + tp]
                                    : 0.0;
      }
      __syncthreads();
      if (j < n) {
        const int q0 = half * 32, q1 = min(q0 + 32, mall - tile);
        with_kind([&](auto kc) __attribute__((always_inline)) {
          constexpr int K = decltype(kc)::value;
          for (int q = q0; q < q1; ++q) {
            const double ph = wendland<K, double, true>(
                dist_fast(xj, yj, zj, sty[q], sty[64 + q], sty[128 + q]),
                dinv);
#pragma unroll
            for (int f = 0; f < kFMax; ++f)
              if (f < F)
                acc[f] = fma(ph, sth[f * 64 + q], acc[f]);
          }
        });
      }
      __syncthreads();
    }
    if (half == 1 && j < n) {
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          sred[f * 256 + j] = acc[f];
    }
    __syncthreads();
    if (half == 0 && j < NMAX) {
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          rhs[f][j] = j < n ? acc[f] + sred[f * 256 + j] : 0.0;
    }
    __syncthreads();
  }

  if (wave == 0) {
    // ---------------- control wave: eliminations -------------------------
    // Elimination of the diagonal block in dtile (rows < kk pivot): lane i
    // (mod 16) holds row i of the block and a quarter of row i of an
    // identity; the pivot row is broadcast along each 16-lane row (DPP
    // row_newbcast). The identity turns into Linv. Writes Linv, d and 1/d to
    // LDS buffer `buf` and the packed Linv of block p.
    auto factor = [&](int kk, int buf, int p) __attribute__((always_inline)) {
      // Lane (q, i) = 16 q + i holds all 16 columns of row i of the block
      // (the four 16-lane rows hold the same copy and update it the same
      // way) and columns 4q .. 4q + 3 of row i of the identity. The pivot
      // and the multiplier of row i are then formed in every 16-lane row
      // from its own copy, with no exchange between rows: a pivot step is a
      // chain of one broadcast, the reciprocal and the update of column
      // k + 1. Columns at or left of the pivot are not updated (never read
      // again).
      const int i = lane & 15, q = lane >> 4;
      double v[16], w[4];
#pragma unroll
      for (int c = 0; c < 16; ++c)
        v[c] = dtile[i][c];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        w[c] = (4 * q + c == i) ? 1.0 : 0.0;
      static_for<0, 16>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value;
        if (k < kk) {
          const double d = row_bcast_asm<k>(v[k]);
          const double nl = (i > k && i < kk) ? -(v[k] * fast_rcp(d)) : 0.0;
#pragma unroll
          for (int c = k + 1; c < 16; ++c)
            fmac_row_bcast<k>(v[c], nl);
#pragma unroll
          for (int c = 0; c < 4; ++c)
            fmac_row_bcast<k>(w[c], nl);
        }
      });
      if (q == 0) {  // the diagonal entry of row i
        double dd = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c)
          dd = (c == i) ? v[c] : dd;
        dpan[buf][i] = dd;
        dinvp[buf][i] = i < kk ? fast_rcp(dd) : 1.0;
        if (i < kk && !(dd > 1e-12 * a.phi0))
          spivot = 1;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int col = 4 * q + c;
        linv[buf][i][col] = w[c];
        if (col < i)
          linvAll[p][i * (i - 1) / 2 + col] = w[c];
      }
    };

    // The eliminations are the critical path of the factorisation: let this
    // wave win the instruction arbitration against the tile waves.
    __builtin_amdgcn_s_setprio(3);
    __syncthreads();  // first diagonal tile handed over
    if (P > 0)
      factor(min(16, n), 0, 0);
    for (int p = 0; p < P; ++p) {
      __syncthreads();  // A
      __syncthreads();  // B
      if (p + 1 < P)
        factor(min(16, n - 16 * (p + 1)), (p + 1) & 1, p + 1);
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();  // E
    __syncthreads();  // F
    // Back substitution L^T alpha = D^-1 L^-1 b, one barrier per block:
    // alpha_B = Linv_B^T (x_B - S_B - L_{B+1,B}^T alpha_{B+1}), where S_B holds
    // the contributions of the blocks below B + 1 (added by the tile waves as
    // soon as their alpha was known) and the sub-diagonal tile is staged in
    // LDS, so the chain from one block to the next stays in this wave.
    double(*sub)[16][17] = reinterpret_cast<double(*)[16][17]>(sbuf);
    double(*sacc)[kFMax][16] = reinterpret_cast<double(*)[kFMax][16]>(
        sbuf + (R - 1) * 16 * 17);
    // Per block, two 16 x 16 x 16 FP64 matrix-core products with the fields
    // as columns (up to 16, F used): y = L_{B+1,B}^T alpha_{B+1}, then
    // alpha_B = Linv_B^T (x_B - S_B - y). In the accumulator layout lane l,
    // register v holds row (l >> 4) + 4 v, field l & 15 -- exactly the B
    // operand of step v of the next product, so alpha and x feed the matrix
    // cores from registers. The A operands (lane l, step s: row l & 15,
    // column 4 s + (l >> 4)) come from the staged sub-diagonal tile and the
    // packed Linv, off the chain; rows past n and entries above the diagonal
    // carry weight 0, the diagonal of Linv weight 1. A field's column is
    // formed in the same order whatever the field count.
    dbl4 aD = {0.0, 0.0, 0.0, 0.0};  // alpha_{B+1}
    for (int B = P - 1; B >= 0; --B) {
      double sA[4], cA[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = 4 * s + lr;
        sA[s] = (B + 1 < P && 16 * (B + 1) + k < n) ? sub[B][k][lc] : 0.0;
        cA[s] = k > lc ? linvAll[B][k * (k - 1) / 2 + lc]
                       : (k == lc ? 1.0 : 0.0);
      }
      dbl4 y = {0.0, 0.0, 0.0, 0.0};
      if (B + 1 < P) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
          y = __builtin_amdgcn_mfma_f64_16x16x4f64(sA[s], aD[s], y, 0, 0, 0);
      }
      dbl4 x;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = lr + 4 * v;
        const bool ok = 16 * B + row < n && lc < F;
        x[v] = ok ? (rhs[lc][16 * B + row] - sacc[B][lc][row]) - y[v] : 0.0;
      }
      dbl4 al = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s)
        al = __builtin_amdgcn_mfma_f64_16x16x4f64(cA[s], x[s], al, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = lr + 4 * v;
        if (16 * B + row < n && lc < F)
          rhs[lc][16 * B + row] = al[v];
      }
      aD = al;
      __builtin_amdgcn_wave_barrier();
      __syncthreads();  // alpha_B ready
    }
  } else {
    // ---------------- tile waves: the lower-triangle tiles ---------------
    const int tw = wave - 1;
    int tI[TPW], tJ[TPW];
    static_for<0, TPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      tI[s] = R;
      tJ[s] = R;
      static_for<0, NWT>([&](auto wc) {
        constexpr int w = decltype(wc)::value;
        constexpr TileTable tab = TS::TAB;
        if (tw == w && s < tab.cnt[w]) {
          tI[s] = tab.I[w][s < kMaxTilesPerWave ? s : 0];
          tJ[s] = tab.J[w][s < kMaxTilesPerWave ? s : 0];
        }
      });
    });

    dbl4 acc[TPW];
    with_kind([&](auto kc) __attribute__((always_inline)) {
    static_for<0, TPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const int I = tI[s], J = tJ[s];
      acc[s] = (I < NB && J < P) ? assemble(kc, I, J)
                                 : dbl4{0.0, 0.0, 0.0, 0.0};
      // hand over the first diagonal tile, publish the first panel (raw)
      if (J == 0 && I == 0 && P > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          dtile[lr + 4 * v][lc] = acc[s][v];
      } else if (J == 0 && I < NB) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          wp[0][lc][16 * I + lr + 4 * v] = acc[s][v];
      }
    });
    });
    __syncthreads();  // first diagonal tile handed over

    for (int p = 0; p < P; ++p) {
      const int buf = p & 1;
      const int kk = min(16, n - 16 * p);
      // Opaque copies of the lane indices for this phase: the LDS addresses
      // of every tile (16 I + lc, lc * LDW + 16 I, ...) are then formed
      // where they are used instead of being hoisted out of the panel loop
      // and kept live across it, one or two registers per tile, which
      // spilled the larger bins (R = 11 at 8 waves: 104 B per lane).
      asm volatile("" : "+v"(lc), "+v"(lr));
      __syncthreads();  // A
      // panel tiles: W = (A_Ip Linv^T) D^-1, kept and published
      static_for<0, TPW>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int I = tI[s], J = tJ[s];
        if (J == p && I > p && I < NB) {
          const dbl4 w4 = to_w(&wp[buf][0][16 * I], LDW, buf, kk);
          acc[s] = w4;
          // every lane has read its raw values before any lane writes
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int v = 0; v < 4; ++v)
            wp[buf][lc][16 * I + lr + 4 * v] = w4[v];
          if (I == p + 1 && I < P) {
            // this wave owns the next diagonal tile: update it with its own
            // W_{p+1} and hand it to the control wave
            __builtin_amdgcn_wave_barrier();
            static_for<0, TPW>([&](auto dc) {
              constexpr int d = decltype(dc)::value;
              if (tI[d] == p + 1 && tJ[d] == p + 1) {
                acc[d] = schur(acc[d], p + 1, p + 1, buf);
#pragma unroll
                for (int v = 0; v < 4; ++v)
                  dtile[lr + 4 * v][lc] = acc[d][v];
              }
            });
          }
        } else if (J == p && I == p && kk < 16 && 16 * p + kk < n + F) {
          // bordered rows inside the last diagonal block
#pragma unroll
          for (int v = 0; v < 4; ++v)
            dtile[lr + 4 * v][lc] = acc[s][v];
          __builtin_amdgcn_wave_barrier();
          dbl4 t4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int k = 4 * q + lr;
            t4 = __builtin_amdgcn_mfma_f64_16x16x4f64(dtile[lc][k],
                                                      linv[buf][lc][k], t4, 0,
                                                      0, 0);
          }
          const double sc2 = lc < kk ? dinvp[buf][lc] : 0.0;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (lr + 4 * v >= kk)
              acc[s][v] = t4[v] * sc2;
        }
      });
      __syncthreads();  // B: panel published, next diagonal tile handed over
      asm volatile("" : "+v"(lc), "+v"(lr));
      // trailing update; publish the next panel's tiles
      static_for<0, TPW>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int I = tI[s], J = tJ[s];
        if (I < NB && J > p && J < P && !(I == p + 1 && J == p + 1)) {
          acc[s] = schur(acc[s], I, J, buf);
          if (J == p + 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              wp[buf ^ 1][lc][16 * I + lr + 4 * v] = acc[s][v];
          }
        }
      });
    }
    __syncthreads();  // E
    // bordered rows -> rhs
    static_for<0, TPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const int I = tI[s], J = tJ[s];
      if (I < NB && J < P) {
        const int c = 16 * J + lc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = 16 * I + lr + 4 * v;
          if (r >= n && r < n + F && c < n)
            rhs[r - n][c] = acc[s][v];
        }
      }
    });
    // stage the sub-diagonal tiles for the control wave; the back-
    // substitution sums S_J follow them in the same (now free) panel buffer
    double(*sub)[16][17] = reinterpret_cast<double(*)[16][17]>(sbuf);
    double(*sacc)[kFMax][16] = reinterpret_cast<double(*)[kFMax][16]>(
        sbuf + (R - 1) * 16 * 17);
    static_for<0, TPW>([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      const int I = tI[s], J = tJ[s];
      if (I == J + 1 && I < P) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          sub[J][lr + 4 * v][lc] = acc[s][v];
      }
    });
    for (int t = tid - 64; t < R * kFMax * 16; t += NT - 64)
      (&sacc[0][0][0])[t] = 0.0;
    __syncthreads();  // F
    for (int B = P - 1; B >= 0; --B) {
      // alpha_{B+1} is known: add L_{B+1,J}^T alpha_{B+1} to S_J, J < B, as
      // one 16 x 16 x 16 FP64 MFMA product for all fields at once: the
      // tile's registers are already the A operand of L^T (lane (lc, lr),
      // step v: L[4 v + lr][lc]); B = alpha (fields past F zero); D row
      // lr + 4 v, column lc = field.
      static_for<0, TPW>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int I = tI[s], J = tJ[s];
        if (I == B + 1 && J < B && I < P) {
          dbl4 d4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 16 * I + 4 * v + lr;
            const double bv = (lc < F && r < n) ? rhs[lc < F ? lc : 0][r]
                                                : 0.0;
            // rows past n (bordered right-hand sides) masked on both sides
            const double av = r < n ? acc[s][v] : 0.0;
            d4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d4, 0, 0, 0);
          }
          if (lc < F) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              sacc[J][lc][lr + 4 * v] += d4[v];
          }
        }
      });
      __syncthreads();  // alpha_B ready
    }
  }

  if (tid == 0 && a.status)
    a.status[patch] = spivot;
  if (tid == 0 && spivot && a.singular)
    atomicAdd(a.singular, 1);

  if constexpr (ADJ) {
    for (int t = tid; t < n; t += NT)
      for (int f = 0; f < F; ++f)
        a.gammaOut[(s0 + t) * F + f] = rhs[f][t];
    return;
  } else {
    if (a.alphaOut)  // saved for the target-coordinate gradient
      for (int t = tid; t < n; t += NT)
        for (int f = 0; f < F; ++f)
          a.alphaOut[(s0 + t) * F + f] = rhs[f][t];
    // the evaluation runs apart (wave_eval_kernel): alpha of source t at
    // ah[t * S + f], zero past F (one record of S doubles per source)
    double *ah = a.alphaHeavy + 8 * s0;
    const int S = alpha_stride(F);
    for (int t = tid; t < n; t += NT)
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < S)
          ah[t * S + f] = f < F ? rhs[f][t] : 0.0;
  }
}

} // namespace

template <int R, int NW>
void launch_solve_tile(int kind, const SolveArgs &a, const double *coef,
                       hipStream_t s) {
  solve_tile_kernel<R, false, NW><<<dim3((unsigned)a.count),
                                    dim3(TileShape<R, NW>::NT), 0, s>>>(
      a, coef, kind);
  PACMAN_HIP_CHECK(hipGetLastError());
}

template <int R>
void launch_solve_tile_adjoint(int kind, const SolveArgs &a, hipStream_t s) {
  solve_tile_kernel<R, true, 8><<<dim3((unsigned)a.count),
                                  dim3(TileShape<R, 8>::NT), 0, s>>>(
      a, nullptr, kind);
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace pacman_amd
