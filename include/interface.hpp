// Drop-in C++ API of the RBF-PUM transfer path (MI355X / gfx950 build).
//
// Same namespace, names, constants and result type as the reference's flat
// C++ API for this path (src/interface.hpp:25-31, 107-110, 132-136;
// src/common/execspaces.hpp:17-24). Only the RBF-PUM entry exists here.
#pragma once

#include <cstdint>
#include <vector>

namespace PACMAN {

typedef double fp_t;            // src/common/types.hpp:15
typedef double coordinates_t;   // src/common/types.hpp:16
typedef int32_t int_t;          // src/common/types.hpp:18

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

} // namespace PACMAN
