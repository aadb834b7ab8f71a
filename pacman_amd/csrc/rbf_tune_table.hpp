// Size-bin table of the per-patch solve, generated by pacman_amd/autotune.py
// on AMD Radeon Graphics (gfx950:sramecc+:xnack-) (2026-12-04), workload: 4096000 Gaussian-mixture sources (seed 4) -> 4096000 uniform targets (seed 5), density-aware patches (leaf 40, 20 to 600 sources), Wendland 0x11;
// median of 5 launches per variant. Times in ms (F fields per launch):
//   F=1: R11 8w 2.64/16w 3.28; R12 8w 3.75/16w 2.85; R13 8w 3.19/16w 2.45; workspace fp64 2.29/mixed 6.58
//   F=2: R11 8w 2.77/16w 3.46; R12 8w 3.94/16w 3.01; R13 8w 3.31/16w 2.54; workspace fp64 2.41/mixed 7.60
//   F=3: R11 8w 2.91/16w 3.62; R12 8w 4.13/16w 3.17; R13 8w 3.52/16w 2.74; workspace fp64 2.53/mixed 8.57
//   F=4: R11 8w 2.97/16w 3.82; R12 8w 4.47/16w 3.45; R13 8w 3.65/16w 3.00; workspace fp64 2.63/mixed 9.48
//   F=5: R11 8w 3.08/16w 4.41; R12 8w 4.55/16w 3.88; R13 8w 3.81/16w 3.33; workspace fp64 2.74/mixed 10.48
//   F=6: R11 8w 3.14/16w 4.66; R12 8w 4.63/16w 4.11; R13 8w 3.98/16w 3.58; workspace fp64 2.85/mixed 11.37
//   F=7: R11 8w 3.24/16w 4.99; R12 8w 4.70/16w 4.26; R13 8w 4.08/16w 3.87; workspace fp64 2.98/mixed 12.31
//   F=8: R11 8w 3.29/16w 5.39; R12 8w 4.84/16w 4.60; R13 8w 4.18/16w 4.14; workspace fp64 3.10/mixed 13.26
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
