// Blocked FP64 per-patch solve, size bins R = 11 .. 13 (16 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<11, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<12, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<13, 16>(int, const SolveArgs &,
                                          const double *, hipStream_t);
} // namespace pacman_amd
