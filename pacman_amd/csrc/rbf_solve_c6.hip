// FP64 per-patch solve for the Wendland C6 basis (kind 3).
#include "rbf_solve_f64.hpp"

namespace pacman_amd {
template void launch_solve_kind<3>(int, const SolveArgs &, const double *,
                                    hipStream_t);
} // namespace pacman_amd
