// Blocked FP64 per-patch solve, size bins R = 7 .. 9 (8 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<7, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<8, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<9, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile_adjoint<7>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<8>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<9>(int, const SolveArgs &,
                                             hipStream_t);
} // namespace pacman_amd
