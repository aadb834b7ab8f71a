// Drop-in C++ API of the RBF-PUM transfer path (MI355X / gfx950 build).
//
// Same namespace, names, constants and result types as the reference's flat
// C++ API for this path (src/interface.hpp:25-40, 76-81, 107-110, 132-136;
// src/common/execspaces.hpp:17-24; src/common/types.hpp:29-45). The RBF-PUM
// and MLS entries are complete; fe_interpolate serves the method
// NEAREST_NEAREST.
#pragma once

#include <cstdint>
#include <vector>

namespace PACMAN {

typedef double fp_t;            // src/common/types.hpp:15
typedef double coordinates_t;   // src/common/types.hpp:16
typedef int32_t int_t;          // src/common/types.hpp:18
typedef int32_t offset_t;       // src/common/types.hpp:21
typedef int32_t method_t;       // src/common/types.hpp:24
typedef int32_t cell_t;         // src/common/types.hpp:25

enum class TransferMethods : method_t {  // src/common/types.hpp:29-37
  NEAREST_NEAREST = 0xF0,
  INTERP_CLAMP = 0xF1,
  INTERP_NEAREST = 0xF2,
  INTERP_ZEROFILL = 0xF3,
  INTERP_EXTRAP = 0xF4,
  RBF_PUM = 0xF5,
  MLS = 0xF6,
};

enum class TransferStatus : int32_t {  // src/common/types.hpp:39-45
  OUTSIDE,
  NEAREST,
  INTER,
  EXTRAP,
  CLAMP,
};

struct ExecSpaces {
  static constexpr const unsigned char SERIAL = 0x00;
  static constexpr const unsigned char OPENMP = 0x01;
  static constexpr const unsigned char THREADS = 0x02;
  static constexpr const unsigned char CUDA = 0x03;
  static constexpr const unsigned char HIP = 0x04;
  static constexpr const unsigned char SYCL = 0x05;
};

struct RbfFunctions {
  static constexpr unsigned char WENDLANDC0 = 0x10;
  static constexpr unsigned char WENDLANDC2 = 0x11;
  static constexpr unsigned char WENDLANDC4 = 0x12;
  static constexpr unsigned char WENDLANDC6 = 0x13;
  static constexpr unsigned char WENDLANDC8 = 0x14;
};

struct FeInterpolateResult {
  std::vector<fp_t> targetValues;   ///< Transferred value per target point.
  std::vector<int_t> targetStatus;  ///< TransferStatus code per target point.
};

struct RbfInterpolateResult {
  std::vector<fp_t> targetValues;  ///< Interpolated value per target point.
};

/// Replaces PACMAN::rbf_interpolate (src/interface.hpp:132-136,
/// src/rbf_interface.cpp:101-130). Row-major host arrays; the pointers are
/// only read. Throws std::invalid_argument for a null pointer with a
/// positive count, an unknown basis, or an execution space this build does
/// not provide (the reference calls exit(1) there), and std::runtime_error
/// for a dimension outside {1, 2, 3} or a device failure.
RbfInterpolateResult rbf_interpolate(int_t spaceDimension,
                                     unsigned char execSpace,
                                     unsigned char rbfFunction,
                                     coordinates_t *sourcePoints,
                                     int_t nSourcePoints, fp_t *sourceValues,
                                     coordinates_t *targetPoints,
                                     int_t nTargetPoints);

/// Replaces PACMAN::fe_interpolate (src/interface.hpp:76-81,
/// src/interface.cpp:93-125) for method NEAREST_NEAREST (0xF0): each target
/// takes the value of its nearest source (ties: lowest source index), status
/// OUTSIDE; connectivity arguments are not read. Same null-pointer and
/// dimension checks as the reference; another method, or an execution space
/// this build does not provide, throws std::invalid_argument.
FeInterpolateResult fe_interpolate(int_t spaceDimension,
                                   unsigned char execSpace, method_t method,
                                   coordinates_t *sourcePoints,
                                   int_t nSourcePoints, fp_t *sourceValues,
                                   int_t *connVal, int_t connValSize,
                                   offset_t *connOff, int_t connOffSize,
                                   cell_t *cellTypes,
                                   coordinates_t *targetPoints,
                                   int_t nTargetPoints,
                                   bool fortranIndexing = false);

/// Replaces PACMAN::MLSInterpolateResult (src/interface.hpp:143-146).
struct MLSInterpolateResult {
  std::vector<fp_t> targetValues;  ///< Interpolated value per target point.
};

/// Replaces PACMAN::MLS_interpolate (src/interface.hpp:166-169,
/// src/mls_interface.cpp:77-106): moving least squares with the reference's
/// defaults (degree-2 basis, the basis-size nearest sources, Wendland C0
/// weights). Same null-pointer and dimension errors; an execution space other
/// than HIP throws std::invalid_argument instead of calling exit(1).
MLSInterpolateResult MLS_interpolate(int_t spaceDimension,
                                     unsigned char execSpace,
                                     coordinates_t *sourcePoints,
                                     int_t nSourcePoints, fp_t *sourceValues,
                                     coordinates_t *targetPoints,
                                     int_t nTargetPoints);

} // namespace PACMAN
