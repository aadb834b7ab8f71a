// Blocked FP64 per-patch solve, size bins R = 12 .. 13 (8 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<12, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<13, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile_adjoint<12>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<13>(int, const SolveArgs &,
                                             hipStream_t);
} // namespace pacman_amd
