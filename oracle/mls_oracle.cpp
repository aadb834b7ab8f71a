// Oracle: CPU FP64 restatement of PACMAN's MLS transfer.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only to check or time against.
//
// The reference (src/mls/interpolator.hpp:47-75, src/mls_interface.cpp:
// 29-107) hands the points to ArborX::Interpolation::MovingLeastSquares with
// its defaults (ArborX >= 2.1, third party, not in the reference tree) and
// interpolates one field. The published algorithm of that class, restated:
//   * neighbours: the k nearest sources of each target, k = the size of the
//     polynomial basis (3, 6, 10 for dimension 1, 2, 3);
//   * coordinates relative to the target, s_j = x_j - y;
//   * radius rho = 1.1 * max(eps, max_j |s_j|) (so no neighbour gets weight 0);
//   * weights phi_j = Wendland<0>(|s_j| / rho) = (1 - |s_j|/rho)^2;
//   * P = [p(s_j)] the Vandermonde matrix of the degree-2 monomials;
//   * A = P^T Phi P (moment matrix), A+ its pseudo-inverse from a symmetric
//     SVD;
//   * coefficients c_j = sum_k A+[0][k] p_k(s_j) phi_j (p(0) = e_0);
//   * value(y) = sum_j c_j f(x_j).
// Choices this restatement makes where only running ArborX could tell:
//   * the monomials are evaluated at s_j / rho instead of s_j. The value is
//     the same whenever A has full rank (a change of basis); with unscaled
//     offsets the spread of A's eigenvalues grows like rho^-4 with the point
//     spacing, so any fixed relative cutoff of the pseudo-inverse would drop
//     the quadratic terms of dense clouds. Scaled, the cutoff removes only
//     genuinely rank-deficient directions (on regular grids some monomial
//     columns are exactly zero);
//   * ties in the k-nearest search go to the lowest source index;
//   * A+ P^T Phi = (Phi^1/2 P)+ Phi^1/2: the pseudo-inverse is taken from
//     the SVD of B = Phi^1/2 P (its singular values are the square roots of
//     A's eigenvalues, so the pseudo-inverse is the same, with the condition
//     number of B instead of its square), by one-sided Jacobi; singular
//     values at or below k * eps * max are dropped.
// Distances are evaluated with -ffp-contract=off, exactly as written.

#include <stdint.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = int64_t;

int basisSize(int dim) { return (dim + 1) * (dim + 2) / 2; }

// Degree <= 2 monomials of s (dim components): 1, s_a, then s_a s_b (a <= b).
void basis(int dim, const double *s, double *p) {
  int c = 0;
  p[c++] = 1.0;
  for (int a = 0; a < dim; ++a)
    p[c++] = s[a];
  for (int a = 0; a < dim; ++a)
    for (int b = a; b < dim; ++b)
      p[c++] = s[a] * s[b];
}

double sqdist(int dim, const double *a, const double *b) {
  double acc = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double t = b[d] - a[d];
    acc = acc + t * t;
  }
  return acc;
}

// First row of the pseudo-inverse of the k x k matrix B (k <= 10), by a
// one-sided Jacobi SVD: B V = U S (columns of B V orthogonal, norms s_i),
// B+ = V S+ U^T, so row 0 of B+ is r_j = sum_i V[0][i] (BV)[j][i] / s_i^2
// over the singular values above k * eps * max s (the others are dropped).
void pinvRow0(int k, double B[10][10], double *r) {
  double V[10][10];
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j)
      V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        double a = 0.0, b = 0.0, g = 0.0;
        for (int j = 0; j < k; ++j) {
          a += B[j][p] * B[j][p];
          b += B[j][q] * B[j][q];
          g += B[j][p] * B[j][q];
        }
        if (g == 0.0 || std::fabs(g) <= DBL_EPSILON * std::sqrt(a * b))
          continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                         (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int j = 0; j < k; ++j) {
          const double bp = B[j][p], bq = B[j][q];
          B[j][p] = c * bp - sn * bq;
          B[j][q] = sn * bp + c * bq;
          const double vp = V[j][p], vq = V[j][q];
          V[j][p] = c * vp - sn * vq;
          V[j][q] = sn * vp + c * vq;
        }
      }
    if (!rotated)
      break;
  }
  double s2[10], s2max = 0.0;
  for (int i = 0; i < k; ++i) {
    s2[i] = 0.0;
    for (int j = 0; j < k; ++j)
      s2[i] += B[j][i] * B[j][i];
    s2max = std::max(s2max, s2[i]);
  }
  const double cut = (double)k * DBL_EPSILON;
  for (int j = 0; j < k; ++j) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i)
      if (s2[i] > cut * cut * s2max)
        acc += V[0][i] * B[j][i] / s2[i];
    r[j] = acc;
  }
}

} // namespace

extern "C" {

// k nearest sources of every target (ties: lowest index) and the MLS
// coefficients: idx[t * k + j] (ascending distance), coef[t * k + j];
// value(t) = sum_j coef[t*k+j] vals[idx[t*k+j]]. Returns k, or -1 when
// there are fewer sources than k or dim is not 1..3.
int oracle_mls_coefficients(int dim, const double *src, int64_t ns,
                            const double *tgt, int64_t nt, int64_t *idx,
                            double *coef) {
  if (dim < 1 || dim > 3)
    return -1;
  const int k = basisSize(dim);
  if (ns < k)
    return -1;
#pragma omp parallel for schedule(dynamic, 64)
  for (i64 t = 0; t < nt; ++t) {
    const double *y = tgt + t * dim;
    // brute-force k nearest (d2, index) with lexicographic order
    std::vector<std::pair<double, i64>> best;
    best.reserve(k + 1);
    for (i64 i = 0; i < ns; ++i) {
      const std::pair<double, i64> c(sqdist(dim, y, src + i * dim), i);
      if ((int)best.size() < k || c < best.back()) {
        best.insert(std::upper_bound(best.begin(), best.end(), c), c);
        if ((int)best.size() > k)
          best.pop_back();
      }
    }
    double s[10][3], rho = std::numeric_limits<double>::epsilon();
    for (int j = 0; j < k; ++j) {
      double n2 = 0.0;
      for (int a = 0; a < dim; ++a) {
        s[j][a] = src[best[j].second * dim + a] - y[a];
        n2 += s[j][a] * s[j][a];
      }
      rho = std::max(rho, std::sqrt(n2));
    }
    rho *= 1.1;
    const double rinv = 1.0 / rho;
    double B[10][10], sw[10], r[10];
    for (int j = 0; j < k; ++j) {
      double n2 = 0.0;
      for (int a = 0; a < dim; ++a)
        n2 += s[j][a] * s[j][a];
      const double x = std::sqrt(n2) / rho;
      sw[j] = x < 1.0 ? 1.0 - x : 0.0;  // phi_j^1/2, phi = (1 - x)^2
      double sc[3];
      for (int a = 0; a < dim; ++a)
        sc[a] = s[j][a] * rinv;
      basis(dim, sc, B[j]);
      for (int u = 0; u < k; ++u)
        B[j][u] = sw[j] * B[j][u];
    }
    // c_0 = row 0 of (Phi^1/2 P)+ applied to Phi^1/2 f
    pinvRow0(k, B, r);
    for (int j = 0; j < k; ++j) {
      idx[t * k + j] = best[j].second;
      coef[t * k + j] = r[j] * sw[j];
    }
  }
  return k;
}

} // extern "C"
