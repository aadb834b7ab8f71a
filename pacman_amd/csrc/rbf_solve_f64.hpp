// FP64 per-patch solve kernel, instantiated once per Wendland kind by the
// rbf_solve_c*.hip translation units (one kind per unit keeps every unit's
// compile time bounded and lets the units build in parallel).
//
// One workgroup of 16 x GC threads per patch:
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
#pragma once

#include "rbf_solve_common.hpp"

namespace pacman_amd {
namespace {

using namespace solve_detail;

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

// Waves per SIMD the register budget is sized for. The solve is latency
// bound, so patches in flight per CU count most: with 512-thread groups,
// R = 8 and R = 10 are held to 128 VGPRs (two groups per CU; R = 10 spills a
// few dozen bytes per lane and still runs 1.4x faster than one group per CU
// without spills); R = 12 keeps one group per CU.
constexpr int solve_min_waves(int R, int GC) {
  return GC == 16 ? 2 : (R >= 12 ? 2 : 4);
}

template <int R, int GC, int KIND>
__global__ __launch_bounds__(16 * GC, solve_min_waves(R, GC)) void
solve_f64_kernel(SolveArgs a, const double *coef) {
  using LY = Layout<R, GC>;
  constexpr int NT = 16 * GC;
  constexpr int NMAX = LY::NMAX;
  constexpr int RB = LY::RB;

  __shared__ double px[NMAX], py[NMAX], pz[NMAX];
  __shared__ double rhs[kFMax][NMAX];
  __shared__ double yv[kFMax][NMAX];
  __shared__ double gcol[2][KGRP][NMAX];
  __shared__ double dinvs[NMAX];
  __shared__ double dblk[16][17];
  __shared__ double spart[kFMax][16];
  __shared__ double scoef[kFMax][4];
  __shared__ double epart[NT / 64][kFMax][64];
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
    // lane; the solved value of row jj is broadcast from lane jj through a
    // scalar register (v_readlane).
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
          const double al = readlane_f64(sv * dk, jj);
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
  // Tiles of TP pairs; thread = (pair, slice): slice s of ES sums sources
  // [s q, s q + q), q = ceil(n / ES), two at a time into two partial sums;
  // partials and slices are added in a fixed order (deterministic). The next
  // tile's targets are loaded before the current tile is summed.
  {
    constexpr int TP = 64;
    constexpr int ES = NT / TP;  // slices
    const int64_t p0 = a.pairOff[patch], p1 = a.pairOff[patch + 1];
    const int mall = (int)(p1 - p0);
    const int lane = tid % TP, sl = tid / TP;
    const int qn = (n + ES - 1) / ES;
    const int l0 = min(n, sl * qn), l1 = min(n, l0 + qn);
    int e = 0;
    double yx = 0.0, yy = 0.0, yz = 0.0;
    if (lane < mall) {
      e = a.pairTgt[p0 + lane];
      const int tp = e & 0x7fffffff;
      yx = a.tx[tp];
      yy = a.ty[tp];
      yz = a.tz[tp];
    }
    for (int tile = 0; tile < mall; tile += TP) {
      const int pi = tile + lane;
      const bool mem = pi < mall && ((unsigned)e >> 31) != 0;
      const double cxv = yx, cyv = yy, czv = yz;
      // prefetch the next tile's target
      const int pn = pi + TP;
      if (pn < mall) {
        e = a.pairTgt[p0 + pn];
        const int tp = e & 0x7fffffff;
        yx = a.tx[tp];
        yy = a.ty[tp];
        yz = a.tz[tp];
      }
      const int ecur = mem ? 1 : 0;
      double acc0[kFMax], acc1[kFMax];
#pragma unroll
      for (int f = 0; f < kFMax; ++f) {
        acc0[f] = 0.0;
        acc1[f] = 0.0;
      }
      if (ecur) {
        int l = l0;
        for (; l + 1 < l1; l += 2) {
          const double phi0 = wendland<KIND, double, true>(
              dist_fast(cxv, cyv, czv, px[l], py[l], pz[l]), a.dinv);
          const double phi1 = wendland<KIND, double, true>(
              dist_fast(cxv, cyv, czv, px[l + 1], py[l + 1], pz[l + 1]),
              a.dinv);
#pragma unroll
          for (int f = 0; f < kFMax; ++f)
            if (f < F) {
              acc0[f] = fma(phi0, rhs[f][l], acc0[f]);
              acc1[f] = fma(phi1, rhs[f][l + 1], acc1[f]);
            }
        }
        if (l < l1) {
          const double phi0 = wendland<KIND, double, true>(
              dist_fast(cxv, cyv, czv, px[l], py[l], pz[l]), a.dinv);
#pragma unroll
          for (int f = 0; f < kFMax; ++f)
            if (f < F)
              acc0[f] = fma(phi0, rhs[f][l], acc0[f]);
        }
      }
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          epart[sl][f][lane] = acc0[f] + acc1[f];
      __syncthreads();
      if (sl == 0 && ecur) {
        const int slot = a.pairSlot[p0 + pi];
#pragma unroll
        for (int f = 0; f < kFMax; ++f)
          if (f < F) {
            double t = epart[0][f][lane];
#pragma unroll
            for (int q = 1; q < ES; ++q)
              t += epart[q][f][lane];
            // V beta (solver.hpp:272), around the centre
            const double vb = scoef[f][0] + scoef[f][1] * (cxv - cx) +
                              scoef[f][2] * (cyv - cy) +
                              scoef[f][3] * (czv - cz);
            a.contrib[(int64_t)slot * F + f] = t + vb;
          }
      }
      __syncthreads();
    }
  }
}

} // namespace

template <int KIND>
void launch_solve_kind(int R, const SolveArgs &a, const double *coef,
                       hipStream_t s) {
  const dim3 grid((unsigned)a.count);
  switch (R) {
  case 2:
    solve_f64_kernel<2, 16, KIND><<<grid, dim3(256), 0, s>>>(a, coef);
    break;
  case 4:
    solve_f64_kernel<4, 16, KIND><<<grid, dim3(256), 0, s>>>(a, coef);
    break;
  case 6:
    solve_f64_kernel<6, 16, KIND><<<grid, dim3(256), 0, s>>>(a, coef);
    break;
  case 8:
    solve_f64_kernel<8, 32, KIND><<<grid, dim3(512), 0, s>>>(a, coef);
    break;
  case 10:
    solve_f64_kernel<10, 32, KIND><<<grid, dim3(512), 0, s>>>(a, coef);
    break;
  case 12:
    solve_f64_kernel<12, 32, KIND><<<grid, dim3(512), 0, s>>>(a, coef);
    break;
  default:
    throw std::runtime_error("solve: unsupported register tile R");
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace pacman_amd
