// C ABI and C++ drop-in entry points of libpacman_rbf.so.
//
// pacman_rbf_interpolate_c / PACMAN::rbf_interpolate replace the reference's
// src/fortran_interface.cpp:63-91 and src/rbf_interface.cpp:101-130: same
// validation order and messages, same error codes, but an execution space
// this build does not provide is an error, not exit(1).
// pacman_fe_interpolate_c / PACMAN::fe_interpolate replace
// src/fortran_interface.cpp and src/interface.cpp:93-125 for the method
// NEAREST_NEAREST. pacman_mls_interpolate_c / PACMAN::MLS_interpolate
// replace src/fortran_interface.cpp:135-164 and src/mls_interface.cpp:77-106.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <exception>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "../../include/fortran_interface.h"
#include "../../include/interface.hpp"
#include "../../include/pacman_mls_plan.h"
#include "../../include/pacman_rbf_plan.h"
#include "mls.hpp"
#include "rbf_plan.hpp"

using pacman_amd::DevBuf;
using pacman_amd::PlanParams;
using pacman_amd::RbfPumPlan;

struct pacman_rbf_plan_s {
  std::unique_ptr<RbfPumPlan> plan;
};

struct pacman_mls_plan_s {
  std::unique_ptr<pacman_amd::MlsPlan> plan;
};

// Scratch of pacman_rbf_plan_apply_noalloc: the sorted source values, the
// (target, patch) contributions and the polynomial coefficients of one call.
// Every device the library runs on holds its own copy from the load of this
// code object until the process ends. In host code `scratch` is a stand-in,
// not a device address: host code uses only the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch()).
// Every element a call reads was written earlier in the same call, so
// nothing has to start at zero; calls that overlap on one device share the
// array and race.
constexpr int64_t kNoAllocScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / 8];

namespace {

thread_local std::string g_lastError;

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double *g_scratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
double *noalloc_scratch() {
  int dev = -1;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    throw std::runtime_error("device index " + std::to_string(dev) +
                             " out of range");
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw std::runtime_error(
          std::string("hipGetSymbolAddress(scratch) failed: ") +
          hipGetErrorString(e));
    g_scratchPtr[dev] = static_cast<double *>(p);
  }
  return g_scratchPtr[dev];
}

void set_error(const char *where, const std::string &msg) {
  g_lastError = std::string(where) + ": " + msg;
}

int kind_of(int rbfFunction) {
  if (rbfFunction >= 0x10 && rbfFunction <= 0x14)
    return rbfFunction - 0x10;
  return -1;
}

const char *space_name(int e) {
  switch (e) {
  case 0:
    return "SERIAL";
  case 1:
    return "OPENMP";
  case 2:
    return "THREADS";
  case 3:
    return "CUDA";
  case 4:
    return "HIP";
  case 5:
    return "SYCL";
  default:
    return nullptr;
  }
}

bool device_present() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

// Runs the full transfer on host arrays with the reference's fixed
// parameters (interpolator.hxx:39-41).
void run_host_transfer(int dim, int kind, const double *src, int64_t ns,
                       const double *vals, const double *tgt, int64_t nt,
                       double *out) {
  pacman_amd::TraceRange tr("pacman::rbf_interpolate (host arrays)");
  hipStream_t s = nullptr;
  DevBuf<double> dsrc((size_t)ns * dim), dvals(ns), dtgt((size_t)nt * dim),
      dout(std::max<int64_t>(nt, 1));
  pacman_amd::copy_to_device(dsrc.get(), src, 8 * (size_t)ns * dim, s);
  pacman_amd::copy_to_device(dtgt.get(), tgt, 8 * (size_t)nt * dim, s);
  // the values are first read by the apply: their copy runs beside the setup
  int dev = 0;
  PACMAN_HIP_CHECK(hipGetDevice(&dev));
  std::exception_ptr valsErr;
  std::thread valsCopy([&] {
    try {
      PACMAN_HIP_CHECK(hipSetDevice(dev));
      pacman_amd::copy_to_device(dvals.get(), vals, 8 * (size_t)ns, s,
                                 /*afterStream=*/false);
    } catch (...) {
      valsErr = std::current_exception();
    }
  });
  PlanParams p;
  p.dim = dim;
  p.kind = kind;
  std::unique_ptr<RbfPumPlan> plan;
  try {
    plan = std::make_unique<RbfPumPlan>(p, dsrc.get(), ns, dtgt.get(), nt, s);
  } catch (...) {
    valsCopy.join();
    throw;
  }
  valsCopy.join();
  if (valsErr)
    std::rethrow_exception(valsErr);
  plan->apply(dvals.get(), 1, dout.get(), s);
  pacman_amd::copy_to_host(out, dout.get(), 8 * (size_t)nt, s);
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
  // The reference's unpivoted LU returns inf / NaN silently on duplicate
  // sources (serial_solve.hpp:76-84); here such a patch is an error.
  int c[3];
  plan->applyCounters(c, s);
  if (c[0] > 0)
    throw std::runtime_error(
        std::to_string(c[0]) +
        " patch(es) are numerically singular (a pivot of the local RBF "
        "matrix vanished: duplicate or near-duplicate source points)");
}

// Nearest-source transfer on host arrays.
void run_host_nearest(int dim, const double *src, int64_t ns,
                      const double *vals, const double *tgt, int64_t nt,
                      double *out, int *status) {
  hipStream_t s = nullptr;
  DevBuf<double> dsrc(std::max<int64_t>(ns * dim, 1)),
      dvals(std::max<int64_t>(ns, 1)), dtgt(std::max<int64_t>(nt * dim, 1)),
      dout(std::max<int64_t>(nt, 1));
  DevBuf<int> dst(std::max<int64_t>(nt, 1));
  if (ns > 0) {
    pacman_amd::copy_to_device(dsrc.get(), src, 8 * (size_t)ns * dim, s);
    pacman_amd::copy_to_device(dvals.get(), vals, 8 * (size_t)ns, s);
  }
  pacman_amd::copy_to_device(dtgt.get(), tgt, 8 * (size_t)nt * dim, s);
  pacman_amd::fe_nearest_transfer(dim, dsrc.get(), ns, dvals.get(), dtgt.get(),
                                  nt, dout.get(), dst.get(), s);
  pacman_amd::copy_to_host(out, dout.get(), 8 * (size_t)nt, s);
  pacman_amd::copy_to_host(status, dst.get(), 4 * (size_t)nt, s);
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

// MLS transfer on host arrays.
void run_host_mls(int dim, const double *src, int64_t ns, const double *vals,
                  const double *tgt, int64_t nt, double *out) {
  hipStream_t s = nullptr;
  DevBuf<double> dsrc((size_t)std::max<int64_t>(ns * dim, 1)),
      dvals((size_t)std::max<int64_t>(ns, 1)),
      dtgt((size_t)std::max<int64_t>(nt * dim, 1)),
      dout((size_t)std::max<int64_t>(nt, 1));
  if (ns > 0) {
    pacman_amd::copy_to_device(dsrc.get(), src, 8 * (size_t)ns * dim, s);
    pacman_amd::copy_to_device(dvals.get(), vals, 8 * (size_t)ns, s);
  }
  pacman_amd::copy_to_device(dtgt.get(), tgt, 8 * (size_t)nt * dim, s);
  pacman_amd::MlsPlan plan(dim, dsrc.get(), ns, dtgt.get(), nt, s);
  plan.apply(dvals.get(), 1, dout.get(), s);
  pacman_amd::copy_to_host(out, dout.get(), 8 * (size_t)nt, s);
  PACMAN_HIP_CHECK(hipStreamSynchronize(s));
}

} // namespace

// ---------------------------------------------------------------------------
// C++ drop-in
// ---------------------------------------------------------------------------
namespace PACMAN {

// Argument checks of the RBF drop-in (src/rbf_interface.cpp:109-114);
// returns the Wendland kind.
static int check_rbf_args(int_t spaceDimension, unsigned char execSpace,
                          unsigned char rbfFunction,
                          const coordinates_t *sourcePoints,
                          int_t nSourcePoints, const fp_t *sourceValues,
                          const coordinates_t *targetPoints,
                          int_t nTargetPoints) {
  if (sourcePoints == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourcePoints is null but nSourcePoints > 0");
  if (targetPoints == nullptr && nTargetPoints > 0)
    throw std::invalid_argument("targetPoints is null but nTargetPoints > 0");
  if (sourceValues == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourceValues is null but nSourcePoints > 0");
  if (spaceDimension < 1 || spaceDimension > 3)
    throw std::runtime_error(
        "The dimension of the points can only be: 1, 2 or 3.\n");
  const char *name = space_name(execSpace);
  if (name == nullptr)
    throw std::invalid_argument(
        "execspace: Serial, OpenMP, Threads, Cuda, HIP, SYCL");
  if (execSpace != ExecSpaces::HIP)
    throw std::invalid_argument(std::string("execution space ") + name +
                                " is not enabled in this build (available: HIP)");
  const int kind = kind_of(rbfFunction);
  if (kind < 0)
    throw std::invalid_argument("Unsupported RBF function selector: " +
                                std::to_string(rbfFunction));
  if (nSourcePoints < 0 || nTargetPoints < 0)
    throw std::invalid_argument("negative number of points");
  return kind;
}

RbfInterpolateResult rbf_interpolate(int_t spaceDimension,
                                     unsigned char execSpace,
                                     unsigned char rbfFunction,
                                     coordinates_t *sourcePoints,
                                     int_t nSourcePoints, fp_t *sourceValues,
                                     coordinates_t *targetPoints,
                                     int_t nTargetPoints) {
  const int kind =
      check_rbf_args(spaceDimension, execSpace, rbfFunction, sourcePoints,
                     nSourcePoints, sourceValues, targetPoints, nTargetPoints);
  RbfInterpolateResult result;
  result.targetValues.assign((size_t)nTargetPoints, 0.0);
  if (nTargetPoints == 0)
    return result;
  run_host_transfer(spaceDimension, kind, sourcePoints, nSourcePoints,
                    sourceValues, targetPoints, nTargetPoints,
                    result.targetValues.data());
  return result;
}

// Argument checks of the FE drop-in (src/interface.cpp:101-107).
static void check_fe_args(int_t spaceDimension, unsigned char execSpace,
                          method_t method,
                          const coordinates_t *sourcePoints,
                          int_t nSourcePoints, const fp_t *sourceValues,
                          const coordinates_t *targetPoints,
                          int_t nTargetPoints) {
  if (sourcePoints == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourcePoints is null but nSourcePoints > 0");
  if (targetPoints == nullptr && nTargetPoints > 0)
    throw std::invalid_argument("targetPoints is null but nTargetPoints > 0");
  if (sourceValues == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourceValues is null but nSourcePoints > 0");
  if (spaceDimension < 1 || spaceDimension > 3)
    throw std::runtime_error(
        "The dimension of the points can only be: 1, 2 or 3.\n");
  const char *name = space_name(execSpace);
  if (name == nullptr)
    throw std::invalid_argument(
        "execspace: Serial, OpenMP, Threads, Cuda, HIP, SYCL");
  if (execSpace != ExecSpaces::HIP)
    throw std::invalid_argument(std::string("execution space ") + name +
                                " is not enabled in this build (available: HIP)");
  if (method != (method_t)TransferMethods::NEAREST_NEAREST)
    throw std::invalid_argument("FE method " + std::to_string(method) +
                                " is not enabled in this build (available: "
                                "NEAREST_NEAREST)");
  if (nSourcePoints < 0 || nTargetPoints < 0)
    throw std::invalid_argument("negative number of points");
}

FeInterpolateResult fe_interpolate(int_t spaceDimension,
                                   unsigned char execSpace, method_t method,
                                   coordinates_t *sourcePoints,
                                   int_t nSourcePoints, fp_t *sourceValues,
                                   int_t *connVal, int_t connValSize,
                                   offset_t *connOff, int_t connOffSize,
                                   cell_t *cellTypes,
                                   coordinates_t *targetPoints,
                                   int_t nTargetPoints, bool fortranIndexing) {
  (void)connVal;
  (void)connValSize;
  (void)connOff;
  (void)connOffSize;
  (void)cellTypes;
  (void)fortranIndexing;
  check_fe_args(spaceDimension, execSpace, method, sourcePoints, nSourcePoints,
                sourceValues, targetPoints, nTargetPoints);
  FeInterpolateResult result;
  result.targetValues.assign((size_t)nTargetPoints, 0.0);
  result.targetStatus.assign((size_t)nTargetPoints, 0);
  if (nTargetPoints == 0)
    return result;
  run_host_nearest(spaceDimension, sourcePoints, nSourcePoints, sourceValues,
                   targetPoints, nTargetPoints, result.targetValues.data(),
                   result.targetStatus.data());
  return result;
}

// Argument checks of the MLS drop-in (src/mls_interface.cpp:83-105).
static void check_mls_args(int_t spaceDimension, unsigned char execSpace,
                           const coordinates_t *sourcePoints,
                           int_t nSourcePoints, const fp_t *sourceValues,
                           const coordinates_t *targetPoints,
                           int_t nTargetPoints) {
  if (sourcePoints == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourcePoints is null but nSourcePoints > 0");
  if (targetPoints == nullptr && nTargetPoints > 0)
    throw std::invalid_argument("targetPoints is null but nTargetPoints > 0");
  if (sourceValues == nullptr && nSourcePoints > 0)
    throw std::invalid_argument("sourceValues is null but nSourcePoints > 0");
  if (spaceDimension < 1 || spaceDimension > 3)
    throw std::runtime_error(
        "The dimension of the points can only be: 1, 2 or 3.\n");
  const char *name = space_name(execSpace);
  if (name == nullptr)
    throw std::invalid_argument(
        "execspace: Serial, OpenMP, Threads, Cuda, HIP, SYCL");
  if (execSpace != ExecSpaces::HIP)
    throw std::invalid_argument(std::string("execution space ") + name +
                                " is not enabled in this build (available: HIP)");
  if (nSourcePoints < 0 || nTargetPoints < 0)
    throw std::invalid_argument("negative number of points");
}

MLSInterpolateResult MLS_interpolate(int_t spaceDimension,
                                     unsigned char execSpace,
                                     coordinates_t *sourcePoints,
                                     int_t nSourcePoints, fp_t *sourceValues,
                                     coordinates_t *targetPoints,
                                     int_t nTargetPoints) {
  check_mls_args(spaceDimension, execSpace, sourcePoints, nSourcePoints,
                 sourceValues, targetPoints, nTargetPoints);
  MLSInterpolateResult result;
  result.targetValues.assign((size_t)nTargetPoints, 0.0);
  if (nTargetPoints == 0)
    return result;
  run_host_mls(spaceDimension, sourcePoints, nSourcePoints, sourceValues,
               targetPoints, nTargetPoints, result.targetValues.data());
  return result;
}

} // namespace PACMAN

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void pacman_kokkos_initialize_c(void) {
  if (device_present()) {
    (void)hipFree(nullptr);  // creates the device context
    try {
      (void)noalloc_scratch();  // the current device's scratch address
    } catch (const std::exception &e) {
      set_error("pacman_kokkos_initialize_c", e.what());
    }
  }
}

void pacman_kokkos_finalize_c(void) {}

int pacman_best_execspace_c(void) {
  return device_present() ? (int)PACMAN_HIP : -1;
}

int pacman_rbf_interpolate_c(int spaceDimension, int execSpace, int rbfFunction,
                             const double *sourcePoints, int nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints, int nTargetPoints,
                             double *targetValues) {
  try {
    if (targetValues == nullptr && nTargetPoints > 0)
      throw std::invalid_argument("targetValues is null but nTargetPoints > 0");
    // straight into the caller's array (no intermediate result vector)
    const int kind = PACMAN::check_rbf_args(
        spaceDimension, (unsigned char)execSpace, (unsigned char)rbfFunction,
        sourcePoints, nSourcePoints, sourceValues, targetPoints,
        nTargetPoints);
    if (nTargetPoints > 0)
      run_host_transfer(spaceDimension, kind, sourcePoints, nSourcePoints,
                        sourceValues, targetPoints, nTargetPoints,
                        targetValues);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_interpolate_c", e.what());
    std::cerr << "pacman_rbf_interpolate_c: " << e.what() << "\n";
    return -1;
  } catch (...) {
    set_error("pacman_rbf_interpolate_c", "unknown exception");
    std::cerr << "pacman_rbf_interpolate_c: unknown exception\n";
    return -2;
  }
}

int pacman_fe_interpolate_c(int spaceDimension, int execSpace, int method,
                            const double *sourcePoints, int nSourcePoints,
                            const double *sourceValues, const int *connVal,
                            int connValSize, const int *connOff,
                            int connOffSize, const int *cellTypes,
                            const double *targetPoints, int nTargetPoints,
                            double *targetValues, int *targetStatus) {
  try {
    if ((targetValues == nullptr || targetStatus == nullptr) &&
        nTargetPoints > 0)
      throw std::invalid_argument(
          "targetValues / targetStatus is null but nTargetPoints > 0");
    // the mesh connectivity is not read by NEAREST_NEAREST
    (void)connVal;
    (void)connValSize;
    (void)connOff;
    (void)connOffSize;
    (void)cellTypes;
    PACMAN::check_fe_args(spaceDimension, (unsigned char)execSpace,
                          (PACMAN::method_t)method, sourcePoints,
                          nSourcePoints, sourceValues, targetPoints,
                          nTargetPoints);
    if (nTargetPoints > 0)
      run_host_nearest(spaceDimension, sourcePoints, nSourcePoints,
                       sourceValues, targetPoints, nTargetPoints,
                       targetValues, targetStatus);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_fe_interpolate_c", e.what());
    std::cerr << "pacman_fe_interpolate_c: " << e.what() << "\n";
    return -1;
  } catch (...) {
    set_error("pacman_fe_interpolate_c", "unknown exception");
    std::cerr << "pacman_fe_interpolate_c: unknown exception\n";
    return -2;
  }
}

int pacman_fe_nearest_device(int spaceDimension, const double *sourcePoints,
                             long long nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints,
                             long long nTargetPoints, double *targetValues,
                             int *targetStatus, void *stream) {
  try {
    if (spaceDimension < 1 || spaceDimension > 3)
      throw std::runtime_error(
          "The dimension of the points can only be: 1, 2 or 3.");
    if (nSourcePoints < 0 || nTargetPoints < 0)
      throw std::invalid_argument("negative number of points");
    if (nSourcePoints > INT_MAX / 2 || nTargetPoints > INT_MAX / 2)
      throw std::invalid_argument("point counts above 2^30 are not supported");
    if ((sourcePoints == nullptr || sourceValues == nullptr) &&
        nSourcePoints > 0)
      throw std::invalid_argument("null source pointer");
    if ((targetPoints == nullptr || targetValues == nullptr ||
         targetStatus == nullptr) &&
        nTargetPoints > 0)
      throw std::invalid_argument("null target pointer");
    pacman_amd::fe_nearest_transfer(spaceDimension, sourcePoints,
                                    nSourcePoints, sourceValues, targetPoints,
                                    nTargetPoints, targetValues, targetStatus,
                                    (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_fe_nearest_device", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_fe_nearest_device", "unknown exception");
    return -2;
  }
}

const char *pacman_rbf_last_error(void) { return g_lastError.c_str(); }

int pacman_rbf_plan_create(pacman_rbf_plan *plan, int spaceDimension,
                           int rbfFunction, int nodesPerCluster,
                           double relativeOverlap, double supportRadius,
                           const double *sourcePoints, long long nSourcePoints,
                           const double *targetPoints, long long nTargetPoints,
                           int onDevice, void *stream) {
  try {
    if (!plan)
      throw std::invalid_argument("plan is null");
    *plan = nullptr;
    if (sourcePoints == nullptr && nSourcePoints > 0)
      throw std::invalid_argument("sourcePoints is null but nSourcePoints > 0");
    if (targetPoints == nullptr && nTargetPoints > 0)
      throw std::invalid_argument("targetPoints is null but nTargetPoints > 0");
    if (spaceDimension < 1 || spaceDimension > 3)
      throw std::runtime_error(
          "The dimension of the points can only be: 1, 2 or 3.");
    const int kind = kind_of(rbfFunction);
    if (kind < 0)
      throw std::invalid_argument("Unsupported RBF function selector: " +
                                  std::to_string(rbfFunction));
    hipStream_t s = (hipStream_t)stream;
    PlanParams p;
    p.dim = spaceDimension;
    p.kind = kind;
    p.npc = nodesPerCluster;
    p.overlap = relativeOverlap;
    p.support = supportRadius;
    auto h = std::make_unique<pacman_rbf_plan_s>();
    if (onDevice) {
      h->plan = std::make_unique<RbfPumPlan>(p, sourcePoints, nSourcePoints,
                                             targetPoints, nTargetPoints, s);
    } else {
      DevBuf<double> ds((size_t)std::max<long long>(nSourcePoints, 0) *
                        spaceDimension),
          dt((size_t)std::max<long long>(nTargetPoints, 0) * spaceDimension);
      if (nSourcePoints > 0)
        pacman_amd::copy_to_device(ds.get(), sourcePoints, 8 * ds.n, s);
      if (nTargetPoints > 0)
        pacman_amd::copy_to_device(dt.get(), targetPoints, 8 * dt.n, s);
      h->plan = std::make_unique<RbfPumPlan>(p, ds.get(), nSourcePoints,
                                             dt.get(), nTargetPoints, s);
    }
    *plan = h.release();
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_create", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_create", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_apply(pacman_rbf_plan plan, const double *sourceValues,
                          int nFields, double *targetValues, int onDevice,
                          void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    hipStream_t s = (hipStream_t)stream;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (sourceValues == nullptr || (targetValues == nullptr && P.numTargets() > 0))
      throw std::invalid_argument("null values pointer");
    if (onDevice) {
      P.apply(sourceValues, nFields, targetValues, s);
    } else {
      DevBuf<double> dv((size_t)nFields * P.numSources()),
          dout((size_t)nFields * std::max<int64_t>(P.numTargets(), 1));
      pacman_amd::copy_to_device(dv.get(), sourceValues, 8 * dv.n, s);
      P.apply(dv.get(), nFields, dout.get(), s);
      if (P.numTargets() > 0)
        pacman_amd::copy_to_host(targetValues, dout.get(),
                                 8 * (size_t)nFields * P.numTargets(), s);
      PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    }
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_apply", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_apply_transpose(pacman_rbf_plan plan,
                                    const double *targetGrad, int nFields,
                                    double *sourceGrad, int onDevice,
                                    void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    hipStream_t s = (hipStream_t)stream;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if ((targetGrad == nullptr && P.numTargets() > 0) ||
        (sourceGrad == nullptr && P.numSources() > 0))
      throw std::invalid_argument("null gradient pointer");
    if (onDevice) {
      P.applyTranspose(targetGrad, nFields, sourceGrad, s);
    } else {
      DevBuf<double> dg((size_t)nFields * std::max<int64_t>(P.numTargets(), 1)),
          dout((size_t)nFields * std::max<int64_t>(P.numSources(), 1));
      if (P.numTargets() > 0)
        pacman_amd::copy_to_device(dg.get(), targetGrad,
                                   8 * (size_t)nFields * P.numTargets(), s);
      P.applyTranspose(dg.get(), nFields, dout.get(), s);
      if (P.numSources() > 0)
        pacman_amd::copy_to_host(sourceGrad, dout.get(),
                                 8 * (size_t)nFields * P.numSources(), s);
      PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    }
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply_transpose", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_apply_transpose", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_apply_graph(pacman_rbf_plan plan,
                                const double *sourceValues, int nFields,
                                double *targetValues, void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (sourceValues == nullptr || (targetValues == nullptr && P.numTargets() > 0))
      throw std::invalid_argument("null values pointer");
    P.applyGraph(sourceValues, nFields, targetValues, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply_graph", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_apply_graph", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_apply_noalloc(pacman_rbf_plan plan,
                                  const double *sourceValues, int nFields,
                                  double *targetValues, void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (sourceValues == nullptr || (targetValues == nullptr && P.numTargets() > 0))
      throw std::invalid_argument("null values pointer");
    double *sp = noalloc_scratch();
    P.applyNoAlloc(sourceValues, nFields, targetValues, (hipStream_t)stream,
                   sp, kNoAllocScratchBytes / 8);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply_noalloc", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_apply_noalloc", "unknown exception");
    return -2;
  }
}

long long pacman_rbf_plan_state_size(pacman_rbf_plan plan, int nFields) {
  if (!plan || !plan->plan || nFields < 1) {
    set_error("pacman_rbf_plan_state_size", "null plan or nFields < 1");
    return -1;
  }
  return plan->plan->stateSize(nFields);
}

int pacman_rbf_plan_apply_state(pacman_rbf_plan plan,
                                const double *sourceValues, int nFields,
                                double *targetValues, double *state,
                                void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (sourceValues == nullptr ||
        (targetValues == nullptr && P.numTargets() > 0))
      throw std::invalid_argument("null values pointer");
    P.applyState(sourceValues, nFields, targetValues, state,
                 (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply_state", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_apply_state", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_target_coord_grad(pacman_rbf_plan plan,
                                      const double *state, int nFields,
                                      const double *targetGrad,
                                      double *coordGrad, void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    RbfPumPlan &P = *plan->plan;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (P.numTargets() > 0 &&
        (state == nullptr || targetGrad == nullptr || coordGrad == nullptr))
      throw std::invalid_argument("null state or gradient pointer");
    P.targetCoordGrad(state, nFields, targetGrad, coordGrad,
                      (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_target_coord_grad", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_target_coord_grad", "unknown exception");
    return -2;
  }
}

int pacman_mls_interpolate_c(int spaceDimension, int execSpace,
                             const double *sourcePoints, int nSourcePoints,
                             const double *sourceValues,
                             const double *targetPoints, int nTargetPoints,
                             double *targetValues) {
  try {
    if (targetValues == nullptr && nTargetPoints > 0)
      throw std::invalid_argument("targetValues is null but nTargetPoints > 0");
    PACMAN::check_mls_args(spaceDimension, (unsigned char)execSpace,
                           sourcePoints, nSourcePoints, sourceValues,
                           targetPoints, nTargetPoints);
    if (nTargetPoints > 0)
      run_host_mls(spaceDimension, sourcePoints, nSourcePoints, sourceValues,
                   targetPoints, nTargetPoints, targetValues);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_mls_interpolate_c", e.what());
    std::cerr << "pacman_mls_interpolate_c: " << e.what() << "\n";
    return -1;
  } catch (...) {
    set_error("pacman_mls_interpolate_c", "unknown exception");
    std::cerr << "pacman_mls_interpolate_c: unknown exception\n";
    return -2;
  }
}

int pacman_mls_plan_create(pacman_mls_plan *plan, int spaceDimension,
                           const double *sourcePoints, long long nSourcePoints,
                           const double *targetPoints, long long nTargetPoints,
                           int onDevice, void *stream) {
  try {
    if (plan == nullptr)
      throw std::invalid_argument("plan is null");
    *plan = nullptr;
    if (spaceDimension < 1 || spaceDimension > 3)
      throw std::runtime_error(
          "The dimension of the points can only be: 1, 2 or 3.\n");
    if (nSourcePoints < 0 || nTargetPoints < 0)
      throw std::invalid_argument("negative number of points");
    if ((sourcePoints == nullptr && nSourcePoints > 0) ||
        (targetPoints == nullptr && nTargetPoints > 0))
      throw std::invalid_argument("null points pointer");
    hipStream_t s = (hipStream_t)stream;
    auto h = std::make_unique<pacman_mls_plan_s>();
    if (onDevice) {
      h->plan = std::make_unique<pacman_amd::MlsPlan>(
          spaceDimension, sourcePoints, nSourcePoints, targetPoints,
          nTargetPoints, s);
    } else {
      DevBuf<double> ds((size_t)std::max<long long>(nSourcePoints, 1) *
                        spaceDimension),
          dt((size_t)std::max<long long>(nTargetPoints, 1) * spaceDimension);
      if (nSourcePoints > 0)
        pacman_amd::copy_to_device(
            ds.get(), sourcePoints,
            8 * (size_t)nSourcePoints * spaceDimension, s);
      if (nTargetPoints > 0)
        pacman_amd::copy_to_device(
            dt.get(), targetPoints,
            8 * (size_t)nTargetPoints * spaceDimension, s);
      h->plan = std::make_unique<pacman_amd::MlsPlan>(
          spaceDimension, ds.get(), nSourcePoints, dt.get(), nTargetPoints, s);
    }
    *plan = h.release();
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_mls_plan_create", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_mls_plan_create", "unknown exception");
    return -2;
  }
}

int pacman_mls_plan_apply(pacman_mls_plan plan, const double *sourceValues,
                          int nFields, double *targetValues, int onDevice,
                          void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    const pacman_amd::MlsPlan &P = *plan->plan;
    hipStream_t s = (hipStream_t)stream;
    if (nFields < 1)
      throw std::invalid_argument("nFields must be >= 1");
    if (sourceValues == nullptr ||
        (targetValues == nullptr && P.numTargets() > 0))
      throw std::invalid_argument("null values pointer");
    if (onDevice) {
      P.apply(sourceValues, nFields, targetValues, s);
    } else {
      DevBuf<double> dv((size_t)nFields * P.numSources()),
          dout((size_t)nFields * std::max<int64_t>(P.numTargets(), 1));
      pacman_amd::copy_to_device(dv.get(), sourceValues, 8 * dv.n, s);
      P.apply(dv.get(), nFields, dout.get(), s);
      if (P.numTargets() > 0)
        pacman_amd::copy_to_host(targetValues, dout.get(),
                                 8 * (size_t)nFields * P.numTargets(), s);
      PACMAN_HIP_CHECK(hipStreamSynchronize(s));
    }
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_mls_plan_apply", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_mls_plan_apply", "unknown exception");
    return -2;
  }
}

int pacman_mls_plan_k(pacman_mls_plan plan) {
  return plan && plan->plan ? plan->plan->k() : -1;
}

int pacman_mls_plan_neighbours(pacman_mls_plan plan, int *idx, double *coef) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    if ((idx == nullptr || coef == nullptr) && plan->plan->numTargets() > 0)
      throw std::invalid_argument("null output pointer");
    plan->plan->neighbours(idx, coef, nullptr);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_mls_plan_neighbours", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_mls_plan_neighbours", "unknown exception");
    return -2;
  }
}

int pacman_mls_plan_destroy(pacman_mls_plan plan) {
  delete plan;
  return 0;
}

int pacman_rbf_plan_build_operator(pacman_rbf_plan plan, void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    plan->plan->buildOperator((hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_build_operator", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_build_operator", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_build_operator_fp32(pacman_rbf_plan plan, void *stream) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    plan->plan->buildOperator((hipStream_t)stream, true);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_build_operator_fp32", e.what());
    return -1;
  } catch (...) {
    set_error("pacman_rbf_plan_build_operator_fp32", "unknown exception");
    return -2;
  }
}

int pacman_rbf_plan_drop_operator(pacman_rbf_plan plan) {
  if (!plan || !plan->plan) {
    set_error("pacman_rbf_plan_drop_operator", "plan is null");
    return -1;
  }
  plan->plan->dropOperator();
  return 0;
}

long long pacman_rbf_plan_operator_bytes(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->operatorBytes() : -1;
}

int pacman_release_device_cache(void) {
  pacman_amd::dev_release_cache();
  return 0;
}

long long pacman_device_cache_bytes(void) {
  return (long long)pacman_amd::dev_cached_bytes();
}

int pacman_rbf_plan_destroy(pacman_rbf_plan plan) {
  delete plan;
  return 0;
}

double pacman_rbf_plan_radius(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->radius() : -1.0;
}

double pacman_rbf_plan_support(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->support() : -1.0;
}

long long pacman_rbf_plan_num_patches(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->numPatches() : -1;
}

long long pacman_rbf_plan_num_pairs(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->numPairs() : -1;
}

int pacman_rbf_plan_max_patch_size(pacman_rbf_plan plan) {
  return plan && plan->plan ? plan->plan->maxPatchSize() : -1;
}

int pacman_rbf_plan_create_adaptive(
    pacman_rbf_plan *plan, int spaceDimension, int rbfFunction,
    int leafPoints, int minPoints, int maxPoints, double relativeOverlap,
    double supportScale, const double *sourcePoints, long long nSourcePoints,
    const double *targetPoints, long long nTargetPoints, int onDevice,
    void *stream) {
  try {
    if (!plan)
      throw std::invalid_argument("plan is null");
    *plan = nullptr;
    if (sourcePoints == nullptr && nSourcePoints > 0)
      throw std::invalid_argument("sourcePoints is null but nSourcePoints > 0");
    if (targetPoints == nullptr && nTargetPoints > 0)
      throw std::invalid_argument("targetPoints is null but nTargetPoints > 0");
    if (spaceDimension < 1 || spaceDimension > 3)
      throw std::runtime_error(
          "The dimension of the points can only be: 1, 2 or 3.");
    const int kind = kind_of(rbfFunction);
    if (kind < 0)
      throw std::invalid_argument("Unsupported RBF function selector: " +
                                  std::to_string(rbfFunction));
    hipStream_t s = (hipStream_t)stream;
    PlanParams p;
    p.dim = spaceDimension;
    p.kind = kind;
    p.overlap = relativeOverlap;
    p.adaptive = true;
    p.ap.leaf = leafPoints;
    p.ap.nmin = minPoints;
    p.ap.nmax = maxPoints;
    p.ap.overlap = relativeOverlap;
    p.ap.supportScale = supportScale;
    auto h = std::make_unique<pacman_rbf_plan_s>();
    if (onDevice) {
      h->plan = std::make_unique<RbfPumPlan>(p, sourcePoints, nSourcePoints,
                                             targetPoints, nTargetPoints, s);
    } else {
      DevBuf<double> ds((size_t)std::max<long long>(nSourcePoints, 0) *
                        spaceDimension),
          dt((size_t)std::max<long long>(nTargetPoints, 0) * spaceDimension);
      if (nSourcePoints > 0)
        pacman_amd::copy_to_device(ds.get(), sourcePoints, 8 * ds.n, s);
      if (nTargetPoints > 0)
        pacman_amd::copy_to_device(dt.get(), targetPoints, 8 * dt.n, s);
      h->plan = std::make_unique<RbfPumPlan>(p, ds.get(), nSourcePoints,
                                             dt.get(), nTargetPoints, s);
    }
    *plan = h.release();
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_create_adaptive", e.what());
    return -1;
  }
}

int pacman_rbf_plan_patch_radii(pacman_rbf_plan plan, double *radii) {
  try {
    if (!plan || !plan->plan || !radii)
      throw std::invalid_argument("plan or output is null");
    plan->plan->patchRadii(radii, nullptr);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_patch_radii", e.what());
    return -1;
  }
}

long long pacman_rbf_plan_num_large_patches(pacman_rbf_plan plan) {
  return plan && plan->plan ? (long long)plan->plan->numLargePatches() : -1;
}

int pacman_rbf_plan_set_precision(pacman_rbf_plan plan, int precision,
                                  int maxRefinements) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    plan->plan->setPrecision(precision, maxRefinements);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_set_precision", e.what());
    return -1;
  }
}

int pacman_rbf_plan_time_solve(pacman_rbf_plan plan, const double *values,
                               int F, int R, int waves, int wsPrecision,
                               int reps, double *ms, void *stream) {
  try {
    if (!plan || !plan->plan || !values || !ms)
      throw std::invalid_argument("plan, values or output is null");
    if (R < 0)
      throw std::invalid_argument("size bin must be >= 0");
    *ms = plan->plan->timeSolve(values, F, R, waves, wsPrecision, reps,
                                (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_time_solve", e.what());
    return -1;
  }
}

int pacman_rbf_plan_apply_counters(pacman_rbf_plan plan, int *counters3,
                                   void *stream) {
  try {
    if (!plan || !plan->plan || !counters3)
      throw std::invalid_argument("plan or output is null");
    plan->plan->applyCounters(counters3, (hipStream_t)stream);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_apply_counters", e.what());
    return -1;
  }
}

int pacman_rbf_plan_lattice(pacman_rbf_plan plan, long long *shape3,
                            long long *nCandidates) {
  if (!plan || !plan->plan)
    return -1;
  for (int a = 0; a < 3; ++a)
    shape3[a] = plan->plan->latticeShape()[a];
  *nCandidates = plan->plan->numCandidates();
  return 0;
}

int pacman_rbf_plan_centers(pacman_rbf_plan plan, double *centers,
                            long long *candidate) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    plan->plan->centers(centers, reinterpret_cast<int64_t *>(candidate),
                        nullptr);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_centers", e.what());
    return -1;
  }
}

long long pacman_rbf_plan_csr_size(pacman_rbf_plan plan, int which) {
  try {
    if (!plan || !plan->plan || which < 0 || which > 2)
      throw std::invalid_argument("bad plan or list selector");
    return plan->plan->csrSize(which);
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_csr_size", e.what());
    return -1;
  }
}

int pacman_rbf_plan_csr(pacman_rbf_plan plan, int which, long long *offsets,
                        long long *indices) {
  try {
    if (!plan || !plan->plan || which < 0 || which > 2)
      throw std::invalid_argument("bad plan or list selector");
    plan->plan->csr(which, reinterpret_cast<int64_t *>(offsets),
                    reinterpret_cast<int64_t *>(indices), nullptr);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_csr", e.what());
    return -1;
  }
}

int pacman_rbf_plan_patch_info(pacman_rbf_plan plan, int *polyColumns,
                               int *status) {
  try {
    if (!plan || !plan->plan)
      throw std::invalid_argument("plan is null");
    plan->plan->patchInfo(polyColumns, status, nullptr);
    return 0;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_patch_info", e.what());
    return -1;
  }
}

int pacman_rbf_plan_phase_times(pacman_rbf_plan plan, double *ms, int n) {
  try {
    if (!plan || !plan->plan || !ms)
      throw std::invalid_argument("plan or output is null");
    double t[RbfPumPlan::kPhases];
    plan->plan->phaseTimes(t);
    const int c = std::min(n, RbfPumPlan::kPhases);
    for (int i = 0; i < c; ++i)
      ms[i] = t[i];
    return c;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_phase_times", e.what());
    return -1;
  }
}

int pacman_rbf_plan_work(pacman_rbf_plan plan, double *work, int n) {
  try {
    if (!plan || !plan->plan || !work)
      throw std::invalid_argument("plan or output is null");
    double w[RbfPumPlan::kWork];
    plan->plan->work(w);
    const int c = std::min(n, RbfPumPlan::kWork);
    for (int i = 0; i < c; ++i)
      work[i] = w[i];
    return c;
  } catch (const std::exception &e) {
    set_error("pacman_rbf_plan_work", e.what());
    return -1;
  }
}

} // extern "C"
