// Device helpers shared by the per-patch solve translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "rbf_internal.hpp"

namespace pacman_amd {
namespace solve_detail {

constexpr int kFMax = 8;   // fields per launch
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
// used here (0 <= x < 1e300). Membership decisions never use it. x = 0 takes
// the estimate of 1e-300 (a finite reciprocal root), so the refinement
// returns exactly 0 there with no select.
__device__ inline double fast_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(fmax(x, 1e-300));
  double s = x * y;
  double h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  const double d = fma(-s, s, x);
  return fma(d, h, s);
}

// Value of v in lane `lane` of the wave (all lanes active; lane a
// compile-time constant in the callers), through scalar registers.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Value of v in lane K of each 16-lane row, for every lane of that row
// (DPP row_newbcast).
template <int K> __device__ __forceinline__ double row_bcast(double v) {
  static_assert(K >= 0 && K < 16, "row lane");
  return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}

__device__ inline double dist_fast(double ax, double ay, double az, double bx,
                                   double by, double bz) {
  const double dx = bx - ax, dy = by - ay, dz = bz - az;
  return fast_sqrt(dx * dx + dy * dy + dz * dz);
}

} // namespace solve_detail
} // namespace pacman_amd
