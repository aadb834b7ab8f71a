// Blocked FP64 per-patch solve, size bins R = 16 .. 17 (16 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<16, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<17, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
} // namespace pacman_amd
