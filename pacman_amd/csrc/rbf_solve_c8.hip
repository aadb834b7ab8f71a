// FP64 per-patch solve for the Wendland C8 basis (kind 4).
#include "rbf_solve_f64.hpp"

namespace pacman_amd {
template void launch_solve_kind<4>(int, const SolveArgs &, const double *,
                                    hipStream_t);
} // namespace pacman_amd
