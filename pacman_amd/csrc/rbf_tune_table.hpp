// Size-bin table of the per-patch solve, generated by pacman_amd/autotune.py
// on AMD Radeon Graphics (2026-12-04), workload: 4096000 Gaussian-mixture sources (seed 4) -> 4096000 uniform targets (seed 5), density-aware patches (leaf 40, 20 to 600 sources), Wendland 0x11;
// median of 5 launches per variant. Times in ms (F fields per launch):
//   F=1: R11 8w 2.63/16w 3.26; R12 8w 3.72/16w 2.83; R13 8w 3.20/16w 2.46; workspace fp64 2.29/mixed 6.62
//   F=2: R11 8w 2.76/16w 3.45; R12 8w 3.91/16w 3.00; R13 8w 3.33/16w 2.56; workspace fp64 2.41/mixed 7.61
//   F=3: R11 8w 2.89/16w 3.61; R12 8w 4.11/16w 3.16; R13 8w 3.53/16w 2.75; workspace fp64 2.53/mixed 8.55
//   F=4: R11 8w 2.94/16w 3.81; R12 8w 4.47/16w 3.47; R13 8w 3.68/16w 3.02; workspace fp64 2.63/mixed 9.44
//   F=5: R11 8w 3.12/16w 4.32; R12 8w 4.68/16w 3.79; R13 8w 3.92/16w 3.29; workspace fp64 2.75/mixed 10.47
//   F=6: R11 8w 3.22/16w 4.58; R12 8w 4.87/16w 4.07; R13 8w 4.16/16w 3.56; workspace fp64 2.86/mixed 11.39
//   F=7: R11 8w 3.37/16w 4.97; R12 8w 5.04/16w 4.26; R13 8w 4.30/16w 3.88; workspace fp64 2.99/mixed 12.34
//   F=8: R11 8w 3.49/16w 5.36; R12 8w 5.28/16w 4.59; R13 8w 4.54/16w 4.15; workspace fp64 3.07/mixed 13.21
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
