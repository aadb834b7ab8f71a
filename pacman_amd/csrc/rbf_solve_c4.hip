// FP64 per-patch solve for the Wendland C4 basis (kind 2).
#include "rbf_solve_f64.hpp"

namespace pacman_amd {
template void launch_solve_kind<2>(int, const SolveArgs &, const double *,
                                    hipStream_t);
} // namespace pacman_amd
