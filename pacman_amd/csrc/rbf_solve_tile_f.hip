// Blocked FP64 per-patch solve, size bins R = 14 .. 15 (16 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<14, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<15, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
} // namespace pacman_amd
