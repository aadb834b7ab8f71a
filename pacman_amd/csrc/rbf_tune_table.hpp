// Size-bin table of the per-patch solve, generated by pacman_amd/autotune.py
// on AMD Radeon Graphics (gfx950:sramecc+:xnack-) (2026-12-04), workload: 4096000 Gaussian-mixture sources (seed 4) -> 4096000 uniform targets (seed 5), density-aware patches (leaf 40, 20 to 600 sources), Wendland 0x11;
// median of 5 launches per variant. Times in ms (F fields per launch):
//   F=1: R11 8w 2.53/16w 3.27; R12 8w 3.78/16w 2.83; R13 8w 3.28/16w 2.48; workspace fp64 2.30/mixed 6.62
//   F=2: R11 8w 2.64/16w 3.45; R12 8w 3.99/16w 2.99; R13 8w 3.41/16w 2.57; workspace fp64 2.41/mixed 7.53
//   F=3: R11 8w 2.76/16w 3.61; R12 8w 4.16/16w 3.16; R13 8w 3.60/16w 2.78; workspace fp64 2.52/mixed 8.58
//   F=4: R11 8w 2.80/16w 3.83; R12 8w 4.52/16w 3.45; R13 8w 3.74/16w 3.02; workspace fp64 2.64/mixed 9.46
//   F=5: R11 8w 2.90/16w 4.42; R12 8w 4.57/16w 3.88; R13 8w 3.90/16w 3.36; workspace fp64 2.74/mixed 10.46
//   F=6: R11 8w 2.96/16w 4.65; R12 8w 4.68/16w 4.10; R13 8w 4.07/16w 3.61; workspace fp64 2.85/mixed 11.38
//   F=7: R11 8w 3.03/16w 5.02; R12 8w 4.75/16w 4.26; R13 8w 4.15/16w 3.91; workspace fp64 2.99/mixed 12.38
//   F=8: R11 8w 3.08/16w 5.40; R12 8w 4.86/16w 4.59; R13 8w 4.28/16w 4.15; workspace fp64 3.08/mixed 13.23
#pragma once

namespace pacman_amd {

// waves[F - 1][R]: workgroup waves of the register-resident solve of bin R
// with F fields per launch; wsFp32[F - 1]: precision "auto" factors the
// workspace-path patches in FP32 with FP64 refinement.
struct TuneTable {
  int waves[8][18];
  bool wsFp32[8];
};

constexpr TuneTable kTune = {
    {
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
        {8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 16, 16, 16, 16, 16, 16},
    },
    {false, false, false, false, false, false, false, false},
};

} // namespace pacman_amd
