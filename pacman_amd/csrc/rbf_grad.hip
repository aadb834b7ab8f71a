// Gradient of the RBF-PUM transfer with respect to the target coordinates.
//
// For a fixed plan (centres, radius, membership) the transfer is
//   u_f(y) = sum_k w_k(y) t_kf(y),  w_k = psi_k / sum_j psi_j,
//   t_kf(y) = sum_i alpha_kif phi(|y - x_i| / delta) + p_kf(y)
// (solver.hpp:128-169 weights, :265-285 evaluation and blend; psi is
// Wendland C2 over the patch radius, interpolator.hpp:86,91; the sum over j
// runs over every centre within r, the sum over k over the patches that hold
// y as a member). Its derivative uses the analytic Wendland derivatives
// phi'(p) = p * dfac(p) (dfac below), so grad_y phi = dfac(p) delta^-2 (y - x),
// and the saved alpha and polynomial coefficients of the forward apply.
//   d u_f / d y = sum_k [(grad psi_k - w_k grad S) / S t_kf + w_k grad t_kf]
// with S = sum_j psi_j. Both kernels work on one upstream contraction
// sum_f g_f(y) (...) so one slot holds 4 numbers whatever the field count.

#include <hip/hip_runtime.h>

#include "rbf_solve_common.hpp"

namespace pacman_amd {

namespace {

using solve_detail::dist_fast;
using solve_detail::kFMax;
constexpr int kGradBlock = 128;

// phi'(p) / p for the Wendland kernels of wendland<KIND> (0 at p >= 1).
// C0 has a kink at p = 0; its derivative is taken as 0 there.
template <int KIND> __device__ inline double wendland_dfac(double p) {
  const double q = 1.0 - p;
  const double q2 = q * q;
  double v;
  if constexpr (KIND == 0) {
    v = p > 0.0 ? -2.0 * q / p : 0.0;
  } else if constexpr (KIND == 1) {
    v = -20.0 * q2 * q;
  } else if constexpr (KIND == 2) {
    v = -56.0 * (q2 * q2 * q) * (5.0 * p + 1.0);
  } else if constexpr (KIND == 3) {
    const double q4 = q2 * q2;
    v = -22.0 * (q4 * q2 * q) * ((16.0 * p + 7.0) * p + 1.0);
  } else {
    const double q4 = q2 * q2;
    v = -78.0 * (q4 * q4 * q) * (((231.0 * p + 159.0) * p + 45.0) * p + 5.0);
  }
  return p < 1.0 ? v : 0.0;
}

template <int KIND>
__global__ __launch_bounds__(kGradBlock) void target_grad_eval_kernel(
    SolveArgs a, int K, const double *alpha, const double *coef, double *gslot,
    int accumulate) {
  __shared__ double px[kMaxPatch], py[kMaxPatch], pz[kMaxPatch];
  __shared__ double sal[kFMax][kMaxPatch];
  __shared__ double scoef[kFMax][4];
  const int patch = xcd_remap(blockIdx.x, K);
  const int64_t s0 = a.srcOff[patch];
  const int n = (int)(a.srcOff[patch + 1] - s0);
  const int F = a.F;
  const int tid = threadIdx.x;
  for (int t = tid; t < n; t += kGradBlock) {
    const int p = a.srcIdx[s0 + t];
    px[t] = a.sx[p];
    py[t] = a.sy[p];
    pz[t] = a.sz[p];
    for (int f = 0; f < F; ++f)
      sal[f][t] = alpha[(s0 + t) * F + f];
  }
  if (tid < F * 4)
    scoef[tid >> 2][tid & 3] = coef[(int64_t)patch * F * 4 + tid];
  __syncthreads();
  const double cx = a.cx[patch], cy = a.cy[patch], cz = a.cz[patch];
  const double dinv = a.dinv, dinv2 = dinv * dinv;
  const int64_t p0 = a.pairOff[patch], p1 = a.pairOff[patch + 1];
  for (int64_t e = p0 + tid; e < p1; e += kGradBlock) {
    const int pe = a.pairTgt[e];
    if (((unsigned)pe >> 31) == 0)
      continue;  // weight-only pair: no local value
    const int tp = pe & 0x7fffffff;
    const double yx = a.tx[tp], yy = a.ty[tp], yz = a.tz[tp];
    double g[kFMax];
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      g[f] = f < F ? a.gSorted[(int64_t)f * a.M + tp] : 0.0;
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i = 0; i < n; ++i) {
      const double dx = yx - px[i], dy = yy - py[i], dz = yz - pz[i];
      const double d = solve_detail::fast_sqrt(dx * dx + dy * dy + dz * dz);
      const double p = d * dinv;
      if (!(p < 1.0))
        continue;
      double ag = 0.0;
#pragma unroll
      for (int f = 0; f < kFMax; ++f)
        if (f < F)
          ag = fma(g[f], sal[f][i], ag);
      const double phi = wendland<KIND, double, true>(d, dinv);
      const double c = ag * wendland_dfac<KIND>(p);
      v = fma(ag, phi, v);
      gx = fma(c, dx, gx);
      gy = fma(c, dy, gy);
      gz = fma(c, dz, gz);
    }
    gx *= dinv2;
    gy *= dinv2;
    gz *= dinv2;
#pragma unroll
    for (int f = 0; f < kFMax; ++f)
      if (f < F) {
        // V beta around the centre (solver.hpp:272) and its gradient
        v = fma(g[f], scoef[f][0] + scoef[f][1] * (yx - cx) +
                          scoef[f][2] * (yy - cy) + scoef[f][3] * (yz - cz),
                v);
        gx = fma(g[f], scoef[f][1], gx);
        gy = fma(g[f], scoef[f][2], gy);
        gz = fma(g[f], scoef[f][3], gz);
      }
    double *o = gslot + 4 * (int64_t)a.pairSlot[e];
    if (accumulate) {
      o[0] += v;
      o[1] += gx;
      o[2] += gy;
      o[3] += gz;
    } else {
      o[0] = v;
      o[1] = gx;
      o[2] = gy;
      o[3] = gz;
    }
  }
}

__global__ __launch_bounds__(256) void target_grad_blend_kernel(
    BlendArgs a, const double *gslot, int dim, double *coordGrad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M)
    return;
  const int64_t s0 = a.toff[t], s1 = a.toff[t + 1];
  const double y[3] = {a.tx[t], a.ty[t], a.tz[t]};
  const double rinv2 = a.rinv * a.rinv;
  double S = 0.0, dS[3] = {0.0, 0.0, 0.0};
  for (int64_t s = s0; s < s1; ++s) {
    const int k = a.slotPatch[s];
    const double c[3] = {a.cx[k], a.cy[k], a.cz[k]};
    const double d = dist_fast(y[0], y[1], y[2], c[0], c[1], c[2]);
    S += wendland<1, double, true>(d, a.rinv);
    const double f = wendland_dfac<1>(d * a.rinv) * rinv2;
    for (int ax = 0; ax < 3; ++ax)
      dS[ax] = fma(f, y[ax] - c[ax], dS[ax]);
  }
  double gr[3] = {0.0, 0.0, 0.0};
  const double Sinv = 1.0 / S;
  for (int64_t s = s0; s < s1; ++s) {
    if (!a.slotMember[s])
      continue;
    const int k = a.slotPatch[s];
    const double c[3] = {a.cx[k], a.cy[k], a.cz[k]};
    const double d = dist_fast(y[0], y[1], y[2], c[0], c[1], c[2]);
    const double w = wendland<1, double, true>(d, a.rinv) * Sinv;
    const double f = wendland_dfac<1>(d * a.rinv) * rinv2;
    const double *gs = gslot + 4 * s;
    for (int ax = 0; ax < 3; ++ax) {
      const double dw = (f * (y[ax] - c[ax]) - w * dS[ax]) * Sinv;
      gr[ax] = fma(dw, gs[0], fma(w, gs[1 + ax], gr[ax]));
    }
  }
  const int64_t o = a.tperm[t];
  for (int ax = 0; ax < dim; ++ax)
    coordGrad[o * dim + ax] = s1 > s0 ? gr[ax] : 0.0;
}

} // namespace

void launch_target_grad_eval(const SolveArgs &a, int64_t K, int kind,
                             const double *alpha, const double *coef,
                             double *gslot, bool accumulate, hipStream_t s) {
  if (K <= 0)
    return;
  if (a.F < 1 || a.F > kFMax)
    throw std::invalid_argument("target gradient: 1..8 fields per launch");
  const dim3 grid((unsigned)K), block(kGradBlock);
  const int acc = accumulate ? 1 : 0;
  switch (kind) {
  case 0:
    target_grad_eval_kernel<0><<<grid, block, 0, s>>>(a, (int)K, alpha, coef,
                                                       gslot, acc);
    break;
  case 1:
    target_grad_eval_kernel<1><<<grid, block, 0, s>>>(a, (int)K, alpha, coef,
                                                       gslot, acc);
    break;
  case 2:
    target_grad_eval_kernel<2><<<grid, block, 0, s>>>(a, (int)K, alpha, coef,
                                                       gslot, acc);
    break;
  case 3:
    target_grad_eval_kernel<3><<<grid, block, 0, s>>>(a, (int)K, alpha, coef,
                                                       gslot, acc);
    break;
  case 4:
    target_grad_eval_kernel<4><<<grid, block, 0, s>>>(a, (int)K, alpha, coef,
                                                       gslot, acc);
    break;
  default:
    throw std::invalid_argument("unknown Wendland kind");
  }
  PACMAN_HIP_CHECK(hipGetLastError());
}

void launch_target_grad_blend(const BlendArgs &a, const double *gslot,
                              int dim, double *coordGrad, hipStream_t s) {
  if (a.M <= 0)
    return;
  target_grad_blend_kernel<<<(unsigned)((a.M + 255) / 256), 256, 0, s>>>(
      a, gslot, dim, coordGrad);
  PACMAN_HIP_CHECK(hipGetLastError());
}

} // namespace pacman_amd
