// Oracle: CPU FP64 restatement of PACMAN's RBF-PUM transfer.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only to check or time against.
//
// It restates, step by step, the reference algorithm:
//   FindRadius        src/rbf_pum/clustering/clusters_radius.hpp:22-125
//   CreateClusters    src/rbf_pum/clustering/clustering.hpp:127-271
//   callbacks         src/rbf_pum/callbacks.hpp:53-114
//   Interpolate       src/rbf_pum/solver/solver.hpp:34-289
//   Wendland kernels  src/rbf_pum/solver/rbf_functions.hpp:22-150
//   per-patch ops     src/rbf_pum/solver/utils/serial_ops.hpp:28-185,
//                     src/rbf_pum/solver/utils/serial_solve.hpp:46-110
//   distances         src/rbf_pum/utils/operators.hpp:125-185
//   constants         src/common/types.hpp:15-21,98
// Deliberate differences (documented in the repository's design notes):
//   * int64 offsets (the reference's int32 offset_t overflows at 4M points);
//   * members of a patch are listed in ascending point index (the reference
//     lists them in ArborX traversal order, which only changes rounding);
//   * nearest-neighbour ties are broken towards the lowest point index;
//   * the partition-of-unity blend sums each target's contributions in
//     ascending patch order (the reference uses atomics in arbitrary order);
//   * the rank-deficient branch refits Q on exactly the active columns
//     (serial_ops.hpp:124-138 reads a stale column after an axis drop);
//   * KokkosKernels' SerialQR / SerialLU / SerialSVD (third party, not in the
//     reference tree, KokkosKernels >= 5.0) are restated from the textbook
//     algorithms they implement: Householder QR, unpivoted Doolittle LU,
//     and singular values by one-sided Jacobi.
// Build with -ffp-contract=off so the geometric decisions (distances,
// lattice positions) are evaluated exactly as written.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using i64 = int64_t;

constexpr double kEps = 1.0e-14;        // src/common/types.hpp:98
constexpr double kCondMax = 1.0e5;      // serial_ops.hpp:130
constexpr double kPruneFactor = 0.4;    // clustering.hpp:228

// ---------------------------------------------------------------------------
// Distances (src/rbf_pum/utils/operators.hpp:125-185). Coordinates are padded
// to 3 with zeros for dim < 3; the padding adds exact zeros to the sums.
// ---------------------------------------------------------------------------
inline double sqdiff(const double *lhs, const double *rhs) {
  double acc = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double d = rhs[i] - lhs[i];
    acc += d * d;
  }
  return acc;
}
inline double dist(const double *lhs, const double *rhs) {
  return std::sqrt(sqdiff(lhs, rhs));
}

// ---------------------------------------------------------------------------
// Wendland kernels, src/rbf_pum/solver/rbf_functions.hpp:22-150.
// kind: 0=C0, 1=C2, 2=C4, 3=C6, 4=C8. Not normalised: phi(0) = 1,1,3,1,15.
// ---------------------------------------------------------------------------
inline double wendland(int kind, double r, double rinv) {
  const double p = r * rinv;
  const double mask = (r >= 0.0 && p < 1.0 && p >= 0.0) ? 1.0 : 0.0;
  const double q = 1.0 - p;
  double v = 0.0;
  switch (kind) {
  case 0:
    v = q * q;
    break;
  case 1:
    v = std::pow(q, 4) * std::fma(4.0, p, 1.0);
    break;
  case 2:
    v = std::pow(q, 6) * (35.0 * std::pow(p, 2) + std::fma(18.0, p, 3.0));
    break;
  case 3:
    v = std::pow(q, 8) * (32.0 * std::pow(p, 3) + 25.0 * std::pow(p, 2) +
                          std::fma(8.0, p, 1.0));
    break;
  case 4:
    v = std::pow(q, 10) *
        (1287.0 * std::pow(p, 4) + 1350.0 * std::pow(p, 3) +
         630.0 * std::pow(p, 2) + std::fma(150.0, p, 15.0));
    break;
  default:
    v = std::numeric_limits<double>::quiet_NaN();
  }
  return mask * v;
}

// d phi / d p of wendland(kind, p, 1) on [0, 1) by the product rule on the
// closed forms above (0 outside); the kink of C0 at p = 0 is given slope 0.
// Used only by the target-coordinate gradient, an extension the reference
// does not have.
inline double wendlandDp(int kind, double p) {
  if (!(p >= 0.0 && p < 1.0))
    return 0.0;
  const double q = 1.0 - p;
  switch (kind) {
  case 0:
    return p > 0.0 ? -2.0 * q : 0.0;
  case 1:
    return -4.0 * std::pow(q, 3) * (4.0 * p + 1.0) + 4.0 * std::pow(q, 4);
  case 2:
    return -6.0 * std::pow(q, 5) * (35.0 * p * p + 18.0 * p + 3.0) +
           std::pow(q, 6) * (70.0 * p + 18.0);
  case 3:
    return -8.0 * std::pow(q, 7) *
               (32.0 * p * p * p + 25.0 * p * p + 8.0 * p + 1.0) +
           std::pow(q, 8) * (96.0 * p * p + 50.0 * p + 8.0);
  case 4:
    return -10.0 * std::pow(q, 9) *
               (1287.0 * std::pow(p, 4) + 1350.0 * std::pow(p, 3) +
                630.0 * p * p + 150.0 * p + 15.0) +
           std::pow(q, 10) *
               (5148.0 * std::pow(p, 3) + 4050.0 * p * p + 1260.0 * p + 150.0);
  default:
    return std::numeric_limits<double>::quiet_NaN();
  }
}

// grad_y of wendland(kind, |y - x|, rinv): phi'(p) rinv (y - x) / |y - x|.
inline void wendlandGrad(int kind, const double *y, const double *x,
                         double rinv, double *g) {
  const double d = dist(y, x);
  const double c = d > 0.0 ? wendlandDp(kind, d * rinv) * rinv / d : 0.0;
  for (int a = 0; a < 3; ++a)
    g[a] = c * (y[a] - x[a]);
}

// ---------------------------------------------------------------------------
// Uniform cell grid used only to accelerate exact sphere / nearest queries.
// Every query visits all points whose cells overlap the query cube (padded),
// then applies the exact predicate, so results equal brute force.
// ---------------------------------------------------------------------------
struct Grid {
  double lo[3] = {0, 0, 0};
  double h = 1.0;
  i64 dims[3] = {1, 1, 1};
  std::vector<i64> start;  // ncell + 1
  std::vector<i64> items;  // point indices grouped by cell, ascending inside
  const double *P = nullptr;
  bool brute = false;
  i64 n = 0;

  void build(const double *pts, i64 np, double cell, bool bruteForce) {
    P = pts;
    n = np;
    brute = bruteForce || np == 0;
    if (brute)
      return;
    double hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<double>::max();
      hi[a] = -std::numeric_limits<double>::max();
    }
    for (i64 i = 0; i < np; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], pts[3 * i + a]);
        hi[a] = std::max(hi[a], pts[3 * i + a]);
      }
    h = cell > 0 ? cell : 1.0;
    for (;;) {
      i64 tot = 1;
      for (int a = 0; a < 3; ++a) {
        dims[a] = (i64)std::floor((hi[a] - lo[a]) / h) + 1;
        tot *= dims[a];
      }
      if (tot <= (i64)1 << 26)
        break;
      h *= 1.5;
    }
    const i64 nc = dims[0] * dims[1] * dims[2];
    std::vector<i64> key(np);
    start.assign(nc + 1, 0);
    for (i64 i = 0; i < np; ++i) {
      key[i] = cellOf(pts + 3 * i);
      start[key[i] + 1]++;
    }
    for (i64 c = 0; c < nc; ++c)
      start[c + 1] += start[c];
    items.resize(np);
    std::vector<i64> fill(start.begin(), start.end() - 1);
    for (i64 i = 0; i < np; ++i)
      items[fill[key[i]]++] = i;  // ascending i inside each cell
  }
  i64 axisCell(int a, double x) const {
    double t = std::floor((x - lo[a]) / h);
    if (t < 0)
      t = 0;
    if (t > (double)(dims[a] - 1))
      t = (double)(dims[a] - 1);
    return (i64)t;
  }
  i64 cellOf(const double *x) const {
    return (axisCell(0, x[0]) * dims[1] + axisCell(1, x[1])) * dims[2] +
           axisCell(2, x[2]);
  }
  // Calls f(i) for every candidate i in cells overlapping [c-rad, c+rad].
  template <class F> void candidates(const double *c, double rad, F &&f) const {
    if (brute) {
      for (i64 i = 0; i < n; ++i)
        f(i);
      return;
    }
    const double pad = rad * 1e-9 + 1e-300;
    i64 c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
      c0[a] = axisCell(a, c[a] - rad - pad);
      c1[a] = axisCell(a, c[a] + rad + pad);
    }
    for (i64 x = c0[0]; x <= c1[0]; ++x)
      for (i64 y = c0[1]; y <= c1[1]; ++y) {
        const i64 b = (x * dims[1] + y) * dims[2];
        for (i64 k = start[b + c0[2]]; k < start[b + c1[2] + 1]; ++k)
          f(items[k]);
      }
  }
};

// Nearest point to c among P (by squared difference; ties -> lowest index),
// restricted to candidates within rad when a grid is used. Returns -1 when no
// point lies within rad (grid) or when P is empty.
i64 nearestWithin(const Grid &g, const double *c, double rad) {
  i64 best = -1;
  double bd = std::numeric_limits<double>::infinity();
  const double r2 = rad * rad;
  g.candidates(c, rad, [&](i64 i) {
    const double d2 = sqdiff(c, g.P + 3 * i);
    if (!g.brute && d2 > r2)
      return;
    if (d2 < bd || (d2 == bd && i < best)) {
      bd = d2;
      best = i;
    }
  });
  return best;
}

i64 nearestBrute(const double *P, i64 n, const double *c) {
  i64 best = -1;
  double bd = std::numeric_limits<double>::infinity();
  for (i64 i = 0; i < n; ++i) {
    const double d2 = sqdiff(c, P + 3 * i);
    if (d2 < bd) {
      bd = d2;
      best = i;
    }
  }
  return best;
}

// ---------------------------------------------------------------------------
// Dense per-patch linear algebra (textbook restatements of what the
// reference calls in KokkosKernels; row-major storage).
// ---------------------------------------------------------------------------

// Singular values of an n x p matrix (p <= 4) by one-sided Jacobi
// (Hestenes). Returns them in decreasing order.
void singularValues(const std::vector<double> &Q, int n, int p, double *S) {
  std::vector<double> U(Q.begin(), Q.begin() + (size_t)n * p);
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int a = 0; a < p; ++a)
      for (int b = a + 1; b < p; ++b) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < n; ++i) {
          const double ua = U[(size_t)i * p + a], ub = U[(size_t)i * p + b];
          alpha += ua * ua;
          beta += ub * ub;
          gamma += ua * ub;
        }
        if (gamma == 0.0)
          continue;
        const double rel = std::fabs(gamma) / std::sqrt(alpha * beta);
        off = std::max(off, rel);
        if (rel < 1e-17)
          continue;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < n; ++i) {
          const double ua = U[(size_t)i * p + a], ub = U[(size_t)i * p + b];
          U[(size_t)i * p + a] = cs * ua - sn * ub;
          U[(size_t)i * p + b] = sn * ua + cs * ub;
        }
      }
    if (off < 1e-16)
      break;
  }
  for (int a = 0; a < p; ++a) {
    double s = 0;
    for (int i = 0; i < n; ++i)
      s += U[(size_t)i * p + a] * U[(size_t)i * p + a];
    S[a] = std::sqrt(s);
  }
  std::sort(S, S + p, [](double x, double y) { return x > y; });
}

// Least squares min ||Q beta - b|| by Householder QR (n x p, p <= 4).
// Mirrors SerialSolveQR: QR(Q), X = Q^T b, solve R beta = X[0:p]
// (serial_solve.hpp:46-67).
void householderLstsq(std::vector<double> Q, int n, int p, const double *b,
                      double *beta) {
  std::vector<double> x(b, b + n);
  for (int j = 0; j < p && j < n; ++j) {
    double nrm2 = 0.0;
    for (int i = j; i < n; ++i)
      nrm2 += Q[(size_t)i * p + j] * Q[(size_t)i * p + j];
    const double nrm = std::sqrt(nrm2);
    if (nrm == 0.0)
      continue;
    const double x0 = Q[(size_t)j * p + j];
    const double alpha = (x0 >= 0 ? -nrm : nrm);
    // v = column - alpha e_j ; H = I - 2 v v^T / (v^T v)
    std::vector<double> v(n - j);
    for (int i = j; i < n; ++i)
      v[i - j] = Q[(size_t)i * p + j];
    v[0] -= alpha;
    double vv = 0.0;
    for (double e : v)
      vv += e * e;
    if (vv == 0.0)
      continue;
    for (int c = j; c < p; ++c) {
      double s = 0.0;
      for (int i = j; i < n; ++i)
        s += v[i - j] * Q[(size_t)i * p + c];
      const double f = 2.0 * s / vv;
      for (int i = j; i < n; ++i)
        Q[(size_t)i * p + c] -= f * v[i - j];
    }
    double s = 0.0;
    for (int i = j; i < n; ++i)
      s += v[i - j] * x[i];
    const double f = 2.0 * s / vv;
    for (int i = j; i < n; ++i)
      x[i] -= f * v[i - j];
  }
  // back substitution R beta = x[0:p] (rows beyond n do not exist: a patch
  // with fewer points than columns keeps those coefficients at zero)
  for (int j = p - 1; j >= n; --j)
    beta[j] = 0.0;
  for (int j = std::min(p, n) - 1; j >= 0; --j) {
    double s = x[j];
    for (int c = j + 1; c < p; ++c)
      s -= Q[(size_t)j * p + c] * beta[c];
    beta[j] = s / Q[(size_t)j * p + j];
  }
}

// Unpivoted Doolittle LU in place + solve (SerialSolveLU,
// serial_solve.hpp:76-84). A is n x n row-major, b overwritten by the
// solution.
void luSolveInPlace(std::vector<double> &A, int n, double *b) {
  for (int k = 0; k < n; ++k) {
    const double piv = A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = A[(size_t)i * n + k] / piv;
      A[(size_t)i * n + k] = l;
      if (l == 0.0)
        continue;
      double *ri = &A[(size_t)i * n];
      const double *rk = &A[(size_t)k * n];
      for (int j = k + 1; j < n; ++j)
        ri[j] -= l * rk[j];
    }
  }
  for (int i = 0; i < n; ++i) {  // L y = b (unit lower)
    double s = b[i];
    for (int j = 0; j < i; ++j)
      s -= A[(size_t)i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {  // U x = y
    double s = b[i];
    for (int j = i + 1; j < n; ++j)
      s -= A[(size_t)i * n + j] * b[j];
    b[i] = s / A[(size_t)i * n + i];
  }
}

// ---------------------------------------------------------------------------
// Pipeline state
// ---------------------------------------------------------------------------
struct Params {
  int dim = 3;
  int kind = 1;          // Wendland C0..C8 -> 0..4
  int npc = 50;          // interpolator.hxx:39 default
  double overlap = 0.15; // interpolator.hxx:40 default
  double support = 0.1;  // interpolator.hxx:41 default; <= 0 means "= r"
  int brute = 0;         // 1: brute-force queries (test cross-check)
};

struct Setup {
  double lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
  double radius = 0.0;
  double samples[7][3];       // projected FindRadius probes
  double sampleMax[7];        // max squared distance per probe
  i64 shape[3] = {1, 1, 1};   // lattice points per axis (after ++)
  double spacing[3] = {0, 0, 0};
  i64 nbCandidates = 0;
  std::vector<double> centers;        // K x 3 (padded)
  std::vector<i64> centerCandidate;   // lattice index of each kept centre
  std::vector<i64> srcOff, srcIdx;    // CSR: sources within r of centre k
  std::vector<i64> tgtOff, tgtIdx;    // CSR: targets within r - eps
  std::vector<i64> wOff, wIdx;        // CSR per target: centres within r
  std::vector<double> w;              // normalised PU weights
  double minMargin = std::numeric_limits<double>::infinity();
  // adaptive patches only: radius and support of every patch
  std::vector<double> patchR, patchDelta;
  std::vector<int> patchDepth;        // tree depth of the patch's cell
};

std::vector<double> pad3(const double *p, i64 n, int dim) {
  std::vector<double> out((size_t)n * 3, 0.0);
  for (i64 i = 0; i < n; ++i)
    for (int a = 0; a < dim; ++a)
      out[3 * i + a] = p[(size_t)i * dim + a];
  return out;
}

void bbox(const std::vector<double> &P, i64 n, int dim, double *lo,
          double *hi) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = 0.0;
    hi[a] = 0.0;
  }
  for (int a = 0; a < dim; ++a) {
    lo[a] = std::numeric_limits<double>::max();
    hi[a] = -std::numeric_limits<double>::max();
    for (i64 i = 0; i < n; ++i) {
      lo[a] = std::min(lo[a], P[3 * i + a]);
      hi[a] = std::max(hi[a], P[3 * i + a]);
    }
  }
}

// FindRadius, clusters_radius.hpp:22-125.
void findRadius(const Params &prm, const std::vector<double> &S, i64 ns,
                Setup &st) {
  const int D = prm.dim;
  double center[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a)
    center[a] = (st.lower[a] + st.upper[a]) / 2.;
  const int ns7 = 2 * D + 1;
  double smp[7][3];
  for (int a = 0; a < 3; ++a)
    smp[0][a] = center[a];
  for (int ax = 0; ax < D; ++ax) {
    double s[3] = {center[0], center[1], center[2]};
    const double len = std::fabs(st.upper[ax] - st.lower[ax]);
    s[ax] -= len * 0.25;
    for (int a = 0; a < 3; ++a)
      smp[ax + 1][a] = s[a];
    s[ax] += len * 0.5;
    for (int a = 0; a < 3; ++a)
      smp[D + ax + 1][a] = s[a];
  }
  std::vector<double> d2(ns);
  const i64 k = std::min<i64>(prm.npc, ns);
  for (int q = 0; q < ns7; ++q) {
    const i64 nn = nearestBrute(S.data(), ns, smp[q]);
    for (int a = 0; a < 3; ++a)
      st.samples[q][a] = S[3 * nn + a];
    for (i64 i = 0; i < ns; ++i)
      d2[i] = sqdiff(st.samples[q], &S[3 * i]);
    std::nth_element(d2.begin(), d2.begin() + (k - 1), d2.end());
    st.sampleMax[q] = d2[k - 1];
  }
  std::vector<double> m(st.sampleMax, st.sampleMax + ns7);
  std::sort(m.begin(), m.end());
  st.radius = std::sqrt(m[D]);
}

// Row-major lattice neighbour offsets of ZCurveNeighborOffsets<D>
// (clustering.hpp:34-116): all 3^D - 1 unit offsets; only the set matters.
std::vector<i64> neighbourOffsets(int D, const i64 *shape) {
  std::vector<i64> out;
  const int total = (D == 1 ? 3 : D == 2 ? 9 : 27);
  for (int t = 0; t < total; ++t) {
    int o[3] = {0, 0, 0}, tt = t;
    bool zero = true;
    for (int d = 0; d < D; ++d) {
      o[d] = tt % 3 - 1;
      tt /= 3;
      if (o[d])
        zero = false;
    }
    if (zero)
      continue;
    i64 idx = 0, stride = 1;
    for (int d = D - 1; d >= 0; --d) {
      idx += o[d] * stride;
      stride *= shape[d];
    }
    out.push_back(idx);
  }
  return out;
}

// CreateClusters, clustering.hpp:127-271. `sequentialPrune` selects the
// literal sequential loop (:230-254); otherwise the equivalent parallel rule.
void createClusters(const Params &prm, const std::vector<double> &S, i64 ns,
                    const std::vector<double> &T, i64 nt, const Grid &gs,
                    const Grid &gt, bool sequentialPrune, Setup &st) {
  const int D = prm.dim;
  const double r = st.radius;
  const double spacing = std::sqrt(4.0 / D) * r * (1.0 - prm.overlap);
  double distances[3] = {0, 0, 0};
  double minDistance = std::numeric_limits<double>::max();
  i64 nb = 1;
  for (int a = 0; a < D; ++a) {
    double edge = st.upper[a] - st.lower[a];
    edge = (edge < 0) ? -edge : edge;
    const int shp = (int)std::ceil(std::max(1.0, edge / spacing));
    distances[a] = edge / shp;
    if (distances[a] < minDistance)
      minDistance = distances[a];
    st.shape[a] = shp + 1;
    nb *= st.shape[a];
  }
  for (int a = 0; a < 3; ++a)
    st.spacing[a] = distances[a];
  st.nbCandidates = nb;
  std::vector<double> cand((size_t)nb * 3, 0.0);
  std::vector<char> alive(nb, 1);
  const double r2 = r * r;
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 L = 0; L < nb; ++L) {
    i64 rem = L, coords[3] = {0, 0, 0};
    for (int a = D - 1; a >= 0; --a) {
      coords[a] = rem % st.shape[a];
      rem /= st.shape[a];
    }
    double *c = &cand[3 * L];
    for (int a = 0; a < D; ++a)
      c[a] = st.lower[a] + distances[a] * (double)coords[a];
    // 6. empty of sources -> dead (nearest source farther than r)
    const i64 ns_ = nearestWithin(gs, c, r);
    if (ns_ < 0 || sqdiff(c, &S[3 * ns_]) > r2) {
      alive[L] = 0;
      continue;
    }
    // 7. empty of targets -> dead
    const i64 nt_ = nearestWithin(gt, c, r);
    if (nt_ < 0 || sqdiff(c, &T[3 * nt_]) > r2) {
      alive[L] = 0;
      continue;
    }
    // 8. snap to nearest source
    for (int a = 0; a < 3; ++a)
      c[a] = S[3 * ns_ + a];
  }
  // 9. prune centres closer than 0.4 * min spacing
  double thr = kPruneFactor * minDistance;
  thr *= thr;
  const std::vector<i64> offs = neighbourOffsets(D, st.shape);
  if (sequentialPrune) {
    for (i64 i = 0; i < nb; ++i) {
      if (!alive[i])
        continue;
      for (i64 off : offs) {
        const i64 j = i + off;
        if (j >= 0 && j < nb && alive[j] &&
            sqdiff(&cand[3 * i], &cand[3 * j]) < thr) {
          alive[i] = 0;
          break;
        }
      }
    }
  } else {
    std::vector<char> before(alive);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < nb; ++i) {
      if (!before[i])
        continue;
      for (i64 off : offs) {
        if (off <= 0)
          continue;
        const i64 j = i + off;
        if (j < nb && before[j] && sqdiff(&cand[3 * i], &cand[3 * j]) < thr) {
          alive[i] = 0;
          break;
        }
      }
    }
  }
  // 10. re-tag against targets from the snapped position
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 L = 0; L < nb; ++L) {
    if (!alive[L])
      continue;
    const i64 nt_ = nearestWithin(gt, &cand[3 * L], r);
    if (nt_ < 0 || sqdiff(&cand[3 * L], &T[3 * nt_]) > r2)
      alive[L] = 0;
  }
  // 11. stable compaction
  st.centers.clear();
  st.centerCandidate.clear();
  for (i64 L = 0; L < nb; ++L)
    if (alive[L]) {
      for (int a = 0; a < 3; ++a)
        st.centers.push_back(cand[3 * L + a]);
      st.centerCandidate.push_back(L);
    }
  (void)ns;
  (void)nt;
}

// Sphere membership CSR (solver.hpp:44-67): points with dist(c, x) <= rad.
void sphereCSR(const std::vector<double> &C, i64 K, const Grid &g, double rad,
               std::vector<i64> &off, std::vector<i64> &idx, double r,
               double *margin) {
  std::vector<std::vector<i64>> lists(K);
  double mg = std::numeric_limits<double>::infinity();
#pragma omp parallel for schedule(dynamic, 64) reduction(min : mg)
  for (i64 k = 0; k < K; ++k) {
    const double *c = &C[3 * k];
    auto &lst = lists[k];
    g.candidates(c, rad, [&](i64 i) {
      const double d = dist(c, g.P + 3 * i);
      const double m = std::fabs(d - rad) / r;
      if (m < mg)
        mg = m;
      if (d <= rad)
        lst.push_back(i);
    });
    std::sort(lst.begin(), lst.end());
  }
  off.assign(K + 1, 0);
  for (i64 k = 0; k < K; ++k)
    off[k + 1] = off[k] + (i64)lists[k].size();
  idx.resize(off[K]);
  for (i64 k = 0; k < K; ++k)
    std::copy(lists[k].begin(), lists[k].end(), idx.begin() + off[k]);
  if (margin)
    *margin = std::min(*margin, mg);
}

void runSetup(const Params &prm, const std::vector<double> &S, i64 ns,
              const std::vector<double> &T, i64 nt, bool sequentialPrune,
              Setup &st) {
  const int D = prm.dim;
  bbox(S, ns, D, st.lower, st.upper);
  findRadius(prm, S, ns, st);
  const double r = st.radius;
  Grid gs, gt;
  gs.build(S.data(), ns, r, prm.brute != 0);
  gt.build(T.data(), nt, r, prm.brute != 0);
  createClusters(prm, S, ns, T, nt, gs, gt, sequentialPrune, st);
  const i64 K = (i64)st.centers.size() / 3;
  st.minMargin = std::numeric_limits<double>::infinity();
  sphereCSR(st.centers, K, gs, r, st.srcOff, st.srcIdx, r, &st.minMargin);
  sphereCSR(st.centers, K, gt, r - kEps, st.tgtOff, st.tgtIdx, r,
            &st.minMargin);
  // PU weights (solver.hpp:128-169): per target, centres within r.
  Grid gc;
  gc.build(st.centers.data(), K, r, prm.brute != 0);
  std::vector<std::vector<i64>> lists(nt);
  const double rinv = 1.0 / r;
  std::vector<std::vector<double>> wl(nt);
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 t = 0; t < nt; ++t) {
    const double *y = &T[3 * t];
    auto &lst = lists[t];
    gc.candidates(y, r, [&](i64 k) {
      if (dist(y, &st.centers[3 * k]) <= r)
        lst.push_back(k);
    });
    std::sort(lst.begin(), lst.end());
    double sum = 0.0;
    auto &ww = wl[t];
    for (i64 k : lst) {
      const double w = wendland(1, dist(y, &st.centers[3 * k]), rinv);
      sum += w;
      ww.push_back(w);
    }
    for (double &w : ww)
      w /= sum;
  }
  st.wOff.assign(nt + 1, 0);
  for (i64 t = 0; t < nt; ++t)
    st.wOff[t + 1] = st.wOff[t] + (i64)lists[t].size();
  st.wIdx.resize(st.wOff[nt]);
  st.w.resize(st.wOff[nt]);
  for (i64 t = 0; t < nt; ++t) {
    std::copy(lists[t].begin(), lists[t].end(), st.wIdx.begin() + st.wOff[t]);
    std::copy(wl[t].begin(), wl[t].end(), st.w.begin() + st.wOff[t]);
  }
}

// ---------------------------------------------------------------------------
// Adaptive patches. Not in the reference: its single FindRadius radius
// (clusters_radius.hpp:22-125) gives, on strongly clustered clouds, patches
// of one to ~1e5 points and leaves sparse regions uncovered. This builder
// (the GPU one in rbf_adaptive.hip follows it step by step) adapts the
// patches to the source density:
//  1. box = bounding box of sources and targets; each point gets a Morton
//     code of its coordinates quantised to 16 bits per axis,
//     q = min(65535, trunc((x - lo) / ext * 65536)), bits interleaved from
//     the most significant, axis 0 first;
//  2. a binary tree over the code bits: a cell (code prefix of d bits) with
//     more than `leaf` sources is halved along the axis of bit d; cells with
//     neither sources nor targets are dropped; a leaf that holds at least
//     one target becomes a patch;
//  3. patch centre = centre of the cell; radius r = max(rc, dk) with
//     rc = (1 + overlap) * half the cell diagonal (the ball covers the cell,
//     so every target is inside some patch) and dk = distance to the
//     nmin-th nearest source (every patch holds at least nmin sources);
//  4. a patch whose ball holds more than nmax sources is halved again (its
//     children with targets replace it) until none does;
//  5. patches ordered by cell (prefix << (bits - d)); support delta_k =
//     supportScale * r_k; members: sources within r_k, targets within
//     r_k - 1e-14 (solver.hpp:62); PU weights psi(|y - c_k| / r_k) over the
//     patches within r_k of each target, psi = Wendland C2 (as
//     solver.hpp:128-169 with a per-patch radius).
// ---------------------------------------------------------------------------
constexpr int kMortonBits = 16;

struct AdaptiveParams {
  int leaf = 40, nmin = 20, nmax = 600;
  double overlap = 0.15, supportScale = 1.0;
};

inline uint64_t quantize(double x, double lo, double ext) {
  const double t = ((x - lo) / ext) * 65536.0;
  if (!(t > 0.0))
    return 0;
  if (t >= 65535.0)
    return 65535;
  return (uint64_t)t;
}

// Morton prefix of depth d of the cell with integer coordinates j[a] (b_a =
// (d + D - 1 - a) / D bits on axis a).
inline uint64_t cellPrefix(const uint64_t *j, int D, int d) {
  uint64_t p = 0;
  for (int t = 0; t < d; ++t) {
    const int a = t % D;
    const int ba = (d + D - 1 - a) / D;
    const int bit = ba - 1 - t / D;
    p = (p << 1) | ((j[a] >> bit) & 1u);
  }
  return p;
}

inline uint64_t mortonCode(const double *x, int D, const double *lo,
                           const double *ext) {
  uint64_t q[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a)
    q[a] = quantize(x[a], lo[a], ext[a]);
  return cellPrefix(q, D, kMortonBits * D);
}

struct AdaptiveCell {
  uint64_t p;
  int d;
};

// Geometry of a cell: centre (padded to 3) and cover radius.
void cellGeometry(const AdaptiveCell &c, int D, const double *lo,
                  const double *ext, double overlap, double *cen,
                  double *rc) {
  double s = 0.0;
  for (int a = 0; a < 3; ++a)
    cen[a] = 0.0;
  for (int a = 0; a < D; ++a) {
    const int ba = (c.d + D - 1 - a) / D;
    uint64_t ia = 0;  // axis-a bits of the prefix, most significant first
    for (int t = a; t < c.d; t += D)
      ia = (ia << 1) | ((c.p >> (c.d - 1 - t)) & 1u);
    const double h = ext[a] * std::ldexp(1.0, -ba);
    cen[a] = lo[a] + ((double)ia + 0.5) * h;
    s = s + h * h;
  }
  *rc = (1.0 + overlap) * (0.5 * std::sqrt(s));
}

void runAdaptiveSetup(const Params &prm, const AdaptiveParams &ap,
                      const std::vector<double> &S, i64 ns,
                      const std::vector<double> &T, i64 nt, Setup &st) {
  const int D = prm.dim, DB = kMortonBits * D;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) {
    lo[a] = std::numeric_limits<double>::max();
    hi[a] = -std::numeric_limits<double>::max();
    for (i64 i = 0; i < ns; ++i) {
      lo[a] = std::min(lo[a], S[3 * i + a]);
      hi[a] = std::max(hi[a], S[3 * i + a]);
    }
    for (i64 i = 0; i < nt; ++i) {
      lo[a] = std::min(lo[a], T[3 * i + a]);
      hi[a] = std::max(hi[a], T[3 * i + a]);
    }
    ext[a] = hi[a] - lo[a];
    if (!(ext[a] > 0.0))
      ext[a] = 1.0;
    st.lower[a] = lo[a];
    st.upper[a] = hi[a];
  }
  std::vector<uint64_t> sc(ns), tc(nt);
  for (i64 i = 0; i < ns; ++i)
    sc[i] = mortonCode(&S[3 * i], D, lo, ext);
  for (i64 i = 0; i < nt; ++i)
    tc[i] = mortonCode(&T[3 * i], D, lo, ext);
  std::vector<uint64_t> ss(sc), ts(tc);
  std::sort(ss.begin(), ss.end());
  std::sort(ts.begin(), ts.end());
  auto countIn = [&](const std::vector<uint64_t> &v, const AdaptiveCell &c) {
    const uint64_t a = c.p << (DB - c.d), b = (c.p + 1) << (DB - c.d);
    return (i64)(std::lower_bound(v.begin(), v.end(), b) -
                 std::lower_bound(v.begin(), v.end(), a));
  };
  // 2. the tree
  std::vector<AdaptiveCell> leaves, stack{{0, 0}};
  while (!stack.empty()) {
    const AdaptiveCell c = stack.back();
    stack.pop_back();
    const i64 nsrc = countIn(ss, c), ntgt = countIn(ts, c);
    if (nsrc > ap.leaf && c.d < DB) {
      for (int h = 0; h < 2; ++h) {
        const AdaptiveCell ch{2 * c.p + (uint64_t)h, c.d + 1};
        if (countIn(ss, ch) > 0 || countIn(ts, ch) > 0)
          stack.push_back(ch);
      }
    } else if (ntgt > 0) {
      leaves.push_back(c);
    }
  }
  // 3-4. radii; halve patches holding more than nmax sources
  const i64 kq = std::min<i64>(ap.nmin, ns);
  struct Patch {
    AdaptiveCell c;
    double cen[3], r;
    i64 n;
  };
  std::vector<Patch> done;
  std::vector<AdaptiveCell> todo = leaves;
  while (!todo.empty()) {
    std::vector<Patch> cur(todo.size());
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 i = 0; i < (i64)todo.size(); ++i) {
      Patch &P = cur[i];
      P.c = todo[i];
      double rc;
      cellGeometry(P.c, D, lo, ext, ap.overlap, P.cen, &rc);
      std::vector<double> d2(ns);
      for (i64 j = 0; j < ns; ++j)
        d2[j] = sqdiff(P.cen, &S[3 * j]);
      std::nth_element(d2.begin(), d2.begin() + (kq - 1), d2.end());
      const double dk = std::sqrt(d2[kq - 1]);
      P.r = std::max(rc, dk);
      i64 n = 0;
      for (i64 j = 0; j < ns; ++j)
        n += dist(P.cen, &S[3 * j]) <= P.r ? 1 : 0;
      P.n = n;
    }
    todo.clear();
    for (const Patch &P : cur) {
      if (P.n > ap.nmax && P.c.d < DB) {
        for (int h = 0; h < 2; ++h) {
          const AdaptiveCell ch{2 * P.c.p + (uint64_t)h, P.c.d + 1};
          if (countIn(ts, ch) > 0)
            todo.push_back(ch);
        }
      } else {
        done.push_back(P);
      }
    }
  }
  std::sort(done.begin(), done.end(), [&](const Patch &x, const Patch &y) {
    return (x.c.p << (DB - x.c.d)) < (y.c.p << (DB - y.c.d));
  });
  const i64 K = (i64)done.size();
  st.centers.assign((size_t)K * 3, 0.0);
  st.patchR.resize(K);
  st.patchDelta.resize(K);
  st.patchDepth.resize(K);
  st.centerCandidate.resize(K);
  for (i64 k = 0; k < K; ++k) {
    for (int a = 0; a < 3; ++a)
      st.centers[3 * k + a] = done[k].cen[a];
    st.patchR[k] = done[k].r;
    st.patchDelta[k] = ap.supportScale * done[k].r;
    st.patchDepth[k] = done[k].c.d;
    st.centerCandidate[k] = (i64)(done[k].c.p << (DB - done[k].c.d));
  }
  // 5. membership and weights, brute force
  std::vector<std::vector<i64>> sl(K), tl(K);
#pragma omp parallel for schedule(dynamic, 16)
  for (i64 k = 0; k < K; ++k) {
    const double *c = &st.centers[3 * k];
    const double r = st.patchR[k];
    for (i64 j = 0; j < ns; ++j)
      if (dist(c, &S[3 * j]) <= r)
        sl[k].push_back(j);
    for (i64 j = 0; j < nt; ++j)
      if (dist(c, &T[3 * j]) <= r - kEps)
        tl[k].push_back(j);
  }
  st.srcOff.assign(K + 1, 0);
  st.tgtOff.assign(K + 1, 0);
  for (i64 k = 0; k < K; ++k) {
    st.srcOff[k + 1] = st.srcOff[k] + (i64)sl[k].size();
    st.tgtOff[k + 1] = st.tgtOff[k] + (i64)tl[k].size();
  }
  st.srcIdx.clear();
  st.tgtIdx.clear();
  for (i64 k = 0; k < K; ++k) {
    st.srcIdx.insert(st.srcIdx.end(), sl[k].begin(), sl[k].end());
    st.tgtIdx.insert(st.tgtIdx.end(), tl[k].begin(), tl[k].end());
  }
  std::vector<std::vector<i64>> wl(nt);
  std::vector<std::vector<double>> wv(nt);
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 t = 0; t < nt; ++t) {
    const double *y = &T[3 * t];
    double sum = 0.0;
    for (i64 k = 0; k < K; ++k) {
      const double d = dist(y, &st.centers[3 * k]);
      if (d <= st.patchR[k]) {
        const double w = wendland(1, d, 1.0 / st.patchR[k]);
        wl[t].push_back(k);
        wv[t].push_back(w);
        sum += w;
      }
    }
    for (double &w : wv[t])
      w /= sum;
  }
  st.wOff.assign(nt + 1, 0);
  for (i64 t = 0; t < nt; ++t)
    st.wOff[t + 1] = st.wOff[t] + (i64)wl[t].size();
  st.wIdx.resize(st.wOff[nt]);
  st.w.resize(st.wOff[nt]);
  for (i64 t = 0; t < nt; ++t) {
    std::copy(wl[t].begin(), wl[t].end(), st.wIdx.begin() + st.wOff[t]);
    std::copy(wv[t].begin(), wv[t].end(), st.w.begin() + st.wOff[t]);
  }
  st.radius = 0.0;
  for (double r : st.patchR)
    st.radius = std::max(st.radius, r);
}

// Per-patch solve (solver.hpp:217-286) for F fields; writes the local
// interpolant at each member target into loc (indexed like tgtIdx).
struct PatchOut {
  std::vector<double> local; // tgtIdx.size() x F
  std::vector<int> pv;       // active polynomial columns per patch
  bool wantGrad = false;     // also the gradient of each local value
  std::vector<double> grad;  // tgtIdx.size() x F x 3 (d/dy of local)
};

void solvePatches(const Params &prm, const std::vector<double> &S,
                  const double *vals, i64 ns, int F,
                  const std::vector<double> &T, const Setup &st, double delta,
                  i64 kBegin, i64 kEnd, PatchOut &po) {
  const int D = prm.dim;
  const i64 K = (i64)st.centers.size() / 3;
  po.local.assign(st.tgtIdx.size() * (size_t)F, 0.0);
  if (po.wantGrad)
    po.grad.assign(st.tgtIdx.size() * (size_t)F * 3, 0.0);
  po.pv.assign(K, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 k = kBegin; k < kEnd; ++k) {
    // one support for all patches, or per patch (adaptive patches)
    const double dinv =
        1.0 / (st.patchDelta.empty() ? delta : st.patchDelta[k]);
    const i64 s0 = st.srcOff[k], n = st.srcOff[k + 1] - s0;
    const i64 t0 = st.tgtOff[k], m = st.tgtOff[k + 1] - t0;
    if (n == 0)
      continue;
    const i64 *src = &st.srcIdx[s0];
    const i64 *tgt = &st.tgtIdx[t0];
    // SerialFillQ (serial_ops.hpp:108-139): conditioning loop on [1, x_a].
    bool active[3] = {D > 0, D > 1, D > 2};
    int pv = D + 1;
    std::vector<double> Q;
    auto fillQ = [&]() {
      int cols = 1;
      for (int a = 0; a < D; ++a)
        cols += active[a] ? 1 : 0;
      Q.assign((size_t)n * cols, 0.0);
      for (i64 i = 0; i < n; ++i) {
        int c = 0;
        Q[(size_t)i * cols + c++] = 1.0;
        for (int a = 0; a < D; ++a)
          if (active[a])
            Q[(size_t)i * cols + c++] = S[3 * src[i] + a];
      }
      return cols;
    };
    do {
      pv = fillQ();
      double sv[4];
      singularValues(Q, (int)n, pv, sv);
      const double cond = sv[0] / std::max(sv[pv - 1], kEps);
      if (cond > kCondMax) {
        // SerialDeactivateOneAxis (serial_ops.hpp:28-57)
        double best = std::numeric_limits<double>::max();
        int bi = 0;
        for (int a = 0; a < D; ++a) {
          double rng = std::numeric_limits<double>::max();
          if (active[a]) {
            double mn = S[3 * src[0] + a], mx = mn;
            for (i64 i = 1; i < n; ++i) {
              mn = std::min(mn, S[3 * src[i] + a]);
              mx = std::max(mx, S[3 * src[i] + a]);
            }
            rng = std::fabs(mx - mn);
          }
          if (rng < best) {
            best = rng;
            bi = a;
          }
        }
        active[bi] = false;
        --pv;
        if (pv <= 1)
          pv = fillQ();
      } else {
        break;
      }
    } while (pv > 1);
    po.pv[k] = pv;
    // A (team_ops.hpp:35-52): phi(|x_i - x_j|), symmetric.
    std::vector<double> A0((size_t)n * n);
    for (i64 i = 0; i < n; ++i)
      for (i64 j = i; j < n; ++j) {
        const double v =
            wendland(prm.kind, dist(&S[3 * src[i]], &S[3 * src[j]]), dinv);
        A0[(size_t)i * n + j] = v;
        A0[(size_t)j * n + i] = v;
      }
    std::vector<double> b(n), beta(pv), A;
    for (int f = 0; f < F; ++f) {
      const double *vf = vals + (size_t)f * ns;
      for (i64 i = 0; i < n; ++i)
        b[i] = vf[src[i]];
      householderLstsq(Q, (int)n, pv, b.data(), beta.data());
      for (i64 i = 0; i < n; ++i) {  // b -= Q beta (solver.hpp:258-263)
        double s = 0.0;
        for (int c = 0; c < pv; ++c)
          s += Q[(size_t)i * pv + c] * beta[c];
        b[i] -= s;
      }
      A = A0;
      luSolveInPlace(A, (int)n, b.data());
      // t = E alpha + V beta (solver.hpp:265-273)
      for (i64 j = 0; j < m; ++j) {
        const double *y = &T[3 * tgt[j]];
        double t = 0.0;
        for (i64 i = 0; i < n; ++i)
          t += wendland(prm.kind, dist(y, &S[3 * src[i]]), dinv) * b[i];
        double pvv = beta[0];
        int c = 1;
        for (int a = 0; a < D; ++a)
          if (active[a])
            pvv += y[a] * beta[c++];
        po.local[(size_t)(t0 + j) * F + f] = t + pvv;
        if (po.wantGrad) {
          // d/dy of E(y) alpha + V(y) beta with the raw basis [1, y_a]
          double gr[3] = {0.0, 0.0, 0.0}, gi[3];
          for (i64 i = 0; i < n; ++i) {
            wendlandGrad(prm.kind, y, &S[3 * src[i]], dinv, gi);
            for (int a = 0; a < 3; ++a)
              gr[a] += gi[a] * b[i];
          }
          int cc = 1;
          for (int a = 0; a < D; ++a)
            if (active[a])
              gr[a] += beta[cc++];
          for (int a = 0; a < 3; ++a)
            po.grad[((size_t)(t0 + j) * F + f) * 3 + a] = gr[a];
        }
      }
    }
  }
}

// PU blend (solver.hpp:275-285), per target in ascending patch order.
void blend(const Setup &st, const PatchOut &po, i64 nt, int F, double *out) {
  const i64 K = (i64)st.centers.size() / 3;
  // For each (patch, member target) find its local value.
  std::vector<std::vector<std::pair<i64, i64>>> byTarget(nt);
  for (i64 k = 0; k < K; ++k)
    for (i64 j = st.tgtOff[k]; j < st.tgtOff[k + 1]; ++j)
      byTarget[st.tgtIdx[j]].push_back({k, j});
#pragma omp parallel for schedule(static)
  for (i64 t = 0; t < nt; ++t) {
    auto &lst = byTarget[t];
    std::sort(lst.begin(), lst.end());
    for (int f = 0; f < F; ++f) {
      double acc = 0.0;
      for (auto &kj : lst) {
        double w = 0.0;
        for (i64 e = st.wOff[t]; e < st.wOff[t + 1]; ++e)
          if (st.wIdx[e] == kj.first)
            w = st.w[e];
        acc += w * po.local[(size_t)kj.second * F + f];
      }
      out[(size_t)f * nt + t] = acc;
    }
  }
}

// Target-coordinate gradient of sum_f sum_t g_f(t) u_f(t) for a fixed
// setup: u(y) = sum_k w_k(y) t_k(y) with w_k = psi_k / S and S the sum of
// psi over every centre within r of y (solver.hpp:128-169, psi = Wendland C2
// over the patch radius), the sum over k over the patches holding y as a
// member (solver.hpp:275-285). out: nt x dim.
void blendGrad(const Setup &st, const PatchOut &po, const std::vector<double> &T,
               i64 nt, int F, const double *g, int dim, double *out) {
  const i64 K = (i64)st.centers.size() / 3;
  const double rinv = 1.0 / st.radius;
  std::vector<std::vector<std::pair<i64, i64>>> byTarget(nt);
  for (i64 k = 0; k < K; ++k)
    for (i64 j = st.tgtOff[k]; j < st.tgtOff[k + 1]; ++j)
      byTarget[st.tgtIdx[j]].push_back({k, j});
#pragma omp parallel for schedule(static)
  for (i64 t = 0; t < nt; ++t) {
    const double *y = &T[3 * t];
    double Ssum = 0.0, dS[3] = {0.0, 0.0, 0.0}, gi[3];
    for (i64 e = st.wOff[t]; e < st.wOff[t + 1]; ++e) {
      const double *c = &st.centers[3 * st.wIdx[e]];
      Ssum += wendland(1, dist(y, c), rinv);
      wendlandGrad(1, y, c, rinv, gi);
      for (int a = 0; a < 3; ++a)
        dS[a] += gi[a];
    }
    auto &lst = byTarget[t];
    std::sort(lst.begin(), lst.end());
    double acc[3] = {0.0, 0.0, 0.0};
    for (auto &kj : lst) {
      const double *c = &st.centers[3 * kj.first];
      const double w = wendland(1, dist(y, c), rinv) / Ssum;
      wendlandGrad(1, y, c, rinv, gi);
      for (int f = 0; f < F; ++f) {
        const double gf = g[(size_t)f * nt + t];
        const double loc = po.local[(size_t)kj.second * F + f];
        for (int a = 0; a < 3; ++a) {
          const double dw = (gi[a] - w * dS[a]) / Ssum;
          acc[a] += gf * (dw * loc +
                          w * po.grad[((size_t)kj.second * F + f) * 3 + a]);
        }
      }
    }
    for (int a = 0; a < dim; ++a)
      out[(size_t)t * dim + a] = acc[a];
  }
}

int kindOf(int rbfFunction) {
  if (rbfFunction >= 0x10 && rbfFunction <= 0x14)
    return rbfFunction - 0x10;
  if (rbfFunction >= 0 && rbfFunction <= 4)
    return rbfFunction;
  return -1;
}

} // namespace

extern "C" {

int oracle_abi_version(void) { return 1; }

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

double oracle_wendland(int rbfFunction, double r, double radiusInv) {
  return wendland(kindOf(rbfFunction), r, radiusInv);
}

int oracle_singular_values(const double *Q, int n, int p, double *S) {
  if (p < 1 || p > 4 || n < 1)
    return -1;
  std::vector<double> q(Q, Q + (size_t)n * p);
  singularValues(q, n, p, S);
  return 0;
}

int oracle_lstsq(const double *Q, int n, int p, const double *b,
                 double *beta) {
  if (p < 1 || p > 4 || n < p)
    return -1;
  householderLstsq(std::vector<double>(Q, Q + (size_t)n * p), n, p, b, beta);
  return 0;
}

int oracle_lu_solve(const double *A, int n, double *b) {
  std::vector<double> a(A, A + (size_t)n * n);
  luSolveInPlace(a, n, b);
  return 0;
}

// Opaque setup handle so tests can inspect each stage.
struct OracleSetup {
  Params prm;
  Setup st;
  std::vector<double> S, T;
  i64 ns = 0, nt = 0;
};

void *oracle_setup_create(int dim, int rbfFunction, int npc, double overlap,
                          double support, const double *src, int64_t ns,
                          const double *tgt, int64_t nt, int brute,
                          int sequentialPrune) {
  if (dim < 1 || dim > 3 || kindOf(rbfFunction) < 0 || ns <= 0 || nt < 0 ||
      npc <= 0 || !(overlap > 0 && overlap < 1))
    return nullptr;
  auto *o = new OracleSetup;
  o->prm.dim = dim;
  o->prm.kind = kindOf(rbfFunction);
  o->prm.npc = npc;
  o->prm.overlap = overlap;
  o->prm.support = support;
  o->prm.brute = brute;
  o->S = pad3(src, ns, dim);
  o->T = pad3(tgt, nt, dim);
  o->ns = ns;
  o->nt = nt;
  runSetup(o->prm, o->S, ns, o->T, nt, sequentialPrune != 0, o->st);
  return o;
}

void oracle_setup_destroy(void *h) { delete static_cast<OracleSetup *>(h); }

double oracle_setup_radius(void *h) {
  return static_cast<OracleSetup *>(h)->st.radius;
}
double oracle_setup_min_margin(void *h) {
  return static_cast<OracleSetup *>(h)->st.minMargin;
}
int64_t oracle_setup_num_candidates(void *h) {
  return static_cast<OracleSetup *>(h)->st.nbCandidates;
}
int64_t oracle_setup_num_patches(void *h) {
  return (int64_t)static_cast<OracleSetup *>(h)->st.centers.size() / 3;
}
void oracle_setup_lattice(void *h, int64_t *shape3, double *spacing3,
                          double *lower3, double *upper3) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  for (int a = 0; a < 3; ++a) {
    shape3[a] = st.shape[a];
    spacing3[a] = st.spacing[a];
    lower3[a] = st.lower[a];
    upper3[a] = st.upper[a];
  }
}
void oracle_setup_probes(void *h, double *samples21, double *maxd2_7) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  for (int q = 0; q < 7; ++q) {
    for (int a = 0; a < 3; ++a)
      samples21[3 * q + a] = st.samples[q][a];
    maxd2_7[q] = st.sampleMax[q];
  }
}
// centres (K x 3, padded) and their lattice indices
void oracle_setup_centers(void *h, double *centers, int64_t *candidate) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  std::copy(st.centers.begin(), st.centers.end(), centers);
  std::copy(st.centerCandidate.begin(), st.centerCandidate.end(), candidate);
}
// which: 0 = source members, 1 = target members, 2 = per-target centres
int64_t oracle_setup_csr_size(void *h, int which) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  return which == 0 ? (int64_t)st.srcIdx.size()
         : which == 1 ? (int64_t)st.tgtIdx.size()
                      : (int64_t)st.wIdx.size();
}
void oracle_setup_csr(void *h, int which, int64_t *off, int64_t *idx) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  const auto &o = which == 0 ? st.srcOff : which == 1 ? st.tgtOff : st.wOff;
  const auto &i = which == 0 ? st.srcIdx : which == 1 ? st.tgtIdx : st.wIdx;
  std::copy(o.begin(), o.end(), off);
  std::copy(i.begin(), i.end(), idx);
}
void oracle_setup_weights(void *h, double *w) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  std::copy(st.w.begin(), st.w.end(), w);
}

// Solve + blend for F fields (vals: F x ns, field-major). Patches
// [kBegin, kEnd) only (kEnd < 0: all); out: F x nt. pvOut (K) optional.
int oracle_setup_apply(void *h, const double *vals, int F, int64_t kBegin,
                       int64_t kEnd, double *out, int *pvOut) {
  auto *o = static_cast<OracleSetup *>(h);
  const i64 K = (i64)o->st.centers.size() / 3;
  if (kEnd < 0 || kEnd > K)
    kEnd = K;
  if (kBegin < 0)
    kBegin = 0;
  const double delta = o->prm.support > 0 ? o->prm.support : o->st.radius;
  PatchOut po;
  solvePatches(o->prm, o->S, vals, o->ns, F, o->T, o->st, delta, kBegin, kEnd,
               po);
  blend(o->st, po, o->nt, F, out);
  if (pvOut)
    std::copy(po.pv.begin(), po.pv.end(), pvOut);
  return 0;
}

// Adaptive setup (the density-aware builder above). support: supportScale
// (delta_k = supportScale * r_k).
void *oracle_adaptive_create(int dim, int rbfFunction, int leaf, int nmin,
                             int nmax, double overlap, double supportScale,
                             const double *src, int64_t ns, const double *tgt,
                             int64_t nt) {
  if (dim < 1 || dim > 3 || kindOf(rbfFunction) < 0 || ns <= 0 || nt < 0 ||
      leaf < 1 || nmin < 1 || nmax < nmin || !(overlap > 0 && overlap < 1) ||
      !(supportScale > 0))
    return nullptr;
  auto *o = new OracleSetup;
  o->prm.dim = dim;
  o->prm.kind = kindOf(rbfFunction);
  o->prm.support = -1.0;
  o->S = pad3(src, ns, dim);
  o->T = pad3(tgt, nt, dim);
  o->ns = ns;
  o->nt = nt;
  AdaptiveParams ap;
  ap.leaf = leaf;
  ap.nmin = nmin;
  ap.nmax = nmax;
  ap.overlap = overlap;
  ap.supportScale = supportScale;
  runAdaptiveSetup(o->prm, ap, o->S, ns, o->T, nt, o->st);
  return o;
}

// Per patch: radius, support, tree depth.
void oracle_adaptive_patches(void *h, double *radius, double *support,
                             int *depth) {
  auto &st = static_cast<OracleSetup *>(h)->st;
  std::copy(st.patchR.begin(), st.patchR.end(), radius);
  std::copy(st.patchDelta.begin(), st.patchDelta.end(), support);
  std::copy(st.patchDepth.begin(), st.patchDepth.end(), depth);
}

uint64_t oracle_morton_code(const double *x, int dim, const double *lo,
                            const double *ext) {
  return mortonCode(x, dim, lo, ext);
}

// Gradient with respect to the target coordinates (nt x dim) of
// sum_f <g_f, u_f>, u = the transfer of vals (F x ns), g: F x nt.
int oracle_setup_target_grad(void *h, const double *vals, int F,
                             const double *g, double *out) {
  auto *o = static_cast<OracleSetup *>(h);
  if (F < 1)
    return -1;
  const i64 K = (i64)o->st.centers.size() / 3;
  const double delta = o->prm.support > 0 ? o->prm.support : o->st.radius;
  PatchOut po;
  po.wantGrad = true;
  solvePatches(o->prm, o->S, vals, o->ns, F, o->T, o->st, delta, 0, K, po);
  blendGrad(o->st, po, o->T, o->nt, F, g, o->prm.dim, out);
  return 0;
}

// FE method NEAREST_NEAREST (src/finite_elements/FTNearest.hpp:28-52): the
// value of each target's nearest source (NearestExtractIndex,
// src/finite_elements/utils/ArborXCallbacks.hpp:172-182; ties -> lowest
// source index), status OUTSIDE = 0 for every target and value 0 without
// sources (src/common/transfer.hxx:100-109). brute != 0: plain O(N M) scan.
int oracle_fe_nearest(int dim, const double *src, int64_t ns,
                      const double *vals, const double *tgt, int64_t nt,
                      int brute, double *out, int *status) {
  if (dim < 1 || dim > 3)
    return -1;
  std::vector<double> S(3 * (size_t)std::max<int64_t>(ns, 0), 0.0);
  for (i64 i = 0; i < ns; ++i)
    for (int a = 0; a < dim; ++a)
      S[3 * i + a] = src[i * dim + a];
  Grid g;
  g.build(S.data(), ns, 0.0, brute != 0);
  if (!g.brute) {
    // cell edge ~ two points per cell over the bounding box
    double lo[3], hi[3], vol = 1.0;
    int dd = 0;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<double>::max();
      hi[a] = -lo[a];
    }
    for (i64 i = 0; i < ns; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], S[3 * i + a]);
        hi[a] = std::max(hi[a], S[3 * i + a]);
      }
    for (int a = 0; a < 3; ++a)
      if (hi[a] > lo[a]) {
        vol *= hi[a] - lo[a];
        ++dd;
      }
    const double cell =
        dd ? std::pow(vol * 2.0 / (double)ns, 1.0 / dd) : 1.0;
    g = Grid();
    g.build(S.data(), ns, cell, false);
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 t = 0; t < nt; ++t) {
    double q[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < dim; ++a)
      q[a] = tgt[t * dim + a];
    status[t] = 0;
    i64 best = -1;
    if (ns > 0) {
      if (g.brute) {
        double bd = std::numeric_limits<double>::infinity();
        for (i64 i = 0; i < ns; ++i) {
          const double d2 = sqdiff(q, &S[3 * i]);
          if (d2 < bd) {
            bd = d2;
            best = i;
          }
        }
      } else {
        // nearestWithin is exact for every point within rad: grow rad until
        // one is found
        for (double rad = g.h; best < 0; rad *= 2.0)
          best = nearestWithin(g, q, rad);
      }
    }
    out[t] = best >= 0 ? vals[best] : 0.0;
  }
  return 0;
}

// One-call restatement of PACMAN::rbf_interpolate with explicit parameters.
int oracle_rbf_interpolate(int dim, int rbfFunction, int npc, double overlap,
                           double support, const double *src, int64_t ns,
                           const double *vals, const double *tgt, int64_t nt,
                           double *out) {
  void *h = oracle_setup_create(dim, rbfFunction, npc, overlap, support, src,
                                ns, tgt, nt, 0, 0);
  if (!h)
    return -1;
  oracle_setup_apply(h, vals, 1, 0, -1, out, nullptr);
  oracle_setup_destroy(h);
  return 0;
}

} // extern "C"
