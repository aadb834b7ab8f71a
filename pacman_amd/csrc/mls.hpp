// Moving-least-squares transfer plan: k nearest sources and weights of every
// target computed once (setup), values gathered for any number of fields
// (apply). Kernels in mls.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbf_internal.hpp"

namespace pacman_amd {

class MlsPlan {
public:
  // Device pointers, points row-major [n x dim]; synchronous on s.
  MlsPlan(int dim, const double *dSrc, int64_t N, const double *dTgt,
          int64_t M, hipStream_t s);
  // dVals: F x N (field-major), dOut: F x M; asynchronous on s.
  void apply(const double *dVals, int F, double *dOut, hipStream_t s) const;
  // Host copies: idx[t * k + j] (original source index, ascending distance),
  // coef[t * k + j]; value(t) = sum_j coef * vals[idx].
  void neighbours(int *idx, double *coef, hipStream_t s) const;
  int dim() const { return dim_; }
  int k() const { return k_; }
  int64_t numSources() const { return N_; }
  int64_t numTargets() const { return M_; }

private:
  int dim_;
  int64_t N_, M_;
  int k_;
  DevBuf<int> idx_;
  DevBuf<double> coef_;
};

} // namespace pacman_amd
