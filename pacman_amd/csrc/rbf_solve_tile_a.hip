// Blocked FP64 per-patch solve, size bins R = 1 .. 6 (8 waves).
#include "rbf_solve_tile.hpp"

namespace pacman_amd {
template void launch_solve_tile<1, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<2, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<3, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<4, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<5, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile<6, 8>(int, const SolveArgs &,
                                          const double *, hipStream_t);
template void launch_solve_tile_adjoint<1>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<2>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<3>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<4>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<5>(int, const SolveArgs &,
                                             hipStream_t);
template void launch_solve_tile_adjoint<6>(int, const SolveArgs &,
                                             hipStream_t);
} // namespace pacman_amd
