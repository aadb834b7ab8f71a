// FP64 per-patch solve for the Wendland C2 basis (kind 1).
#include "rbf_solve_f64.hpp"

namespace pacman_amd {
template void launch_solve_kind<1>(int, const SolveArgs &, const double *,
                                    hipStream_t);
} // namespace pacman_amd
