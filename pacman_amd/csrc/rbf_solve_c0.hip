// FP64 per-patch solve for the Wendland C0 basis (kind 0).
#include "rbf_solve_f64.hpp"

namespace pacman_amd {
template void launch_solve_kind<0>(int, const SolveArgs &, const double *,
                                    hipStream_t);
} // namespace pacman_amd
