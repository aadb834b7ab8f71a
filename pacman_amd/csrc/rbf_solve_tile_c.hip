// Blocked FP64 per-patch solve, size bins R = 10 .. 11 (8 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<10, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<11, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile_adjoint<10>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<11>(int, const SolveArgs &,
                                             hipStream_t);
} // namespace pacman_amd
