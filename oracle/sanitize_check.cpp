// Drives every entry point of the oracle on small inputs; built with
// AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
// sanitize` (test infrastructure: tests/test_oracle_sanitizers.py runs it).
// Exits 0 when every call returns and its results are finite.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
int oracle_set_threads(int n);
double oracle_wendland(int rbfFunction, double r, double radiusInv);
int oracle_singular_values(const double *Q, int n, int p, double *S);
int oracle_lstsq(const double *Q, int n, int p, const double *b, double *x);
int oracle_lu_solve(const double *A, int n, double *b);
void *oracle_setup_create(int dim, int rbfFunction, int npc, double overlap,
                          double support, const double *src, int64_t ns,
                          const double *tgt, int64_t nt, int brute,
                          int sequentialPrune);
void oracle_setup_destroy(void *h);
double oracle_setup_radius(void *h);
int64_t oracle_setup_num_patches(void *h);
int64_t oracle_setup_csr_size(void *h, int which);
void oracle_setup_csr(void *h, int which, int64_t *off, int64_t *idx);
void oracle_setup_weights(void *h, double *w);
int oracle_setup_apply(void *h, const double *vals, int F, int64_t kBegin,
                       int64_t kEnd, double *out, int *pv);
int oracle_setup_target_grad(void *h, const double *vals, int F,
                             const double *g, double *out);
void *oracle_adaptive_create(int dim, int rbfFunction, int leaf, int nmin,
                             int nmax, double overlap, double supportScale,
                             const double *src, int64_t ns, const double *tgt,
                             int64_t nt);
int oracle_fe_nearest(int dim, const double *src, int64_t ns,
                      const double *vals, const double *tgt, int64_t nt,
                      int brute, double *out, int *status);
int oracle_rbf_interpolate(int dim, int rbfFunction, int npc, double overlap,
                           double support, const double *src, int64_t ns,
                           const double *vals, const double *tgt, int64_t nt,
                           double *out);
int oracle_mls_coefficients(int dim, const double *src, int64_t ns,
                            const double *tgt, int64_t nt, int64_t *idx,
                            double *coef);
}

namespace {

int failures = 0;

void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "sanitize_check: %s failed\n", what);
    ++failures;
  }
}

bool finite(const std::vector<double> &v) {
  for (double x : v)
    if (!std::isfinite(x))
      return false;
  return true;
}

std::vector<double> cloud(std::mt19937_64 &g, int n, int dim, double lo,
                          double hi) {
  std::uniform_real_distribution<double> u(lo, hi);
  std::vector<double> p((size_t)n * dim);
  for (double &x : p)
    x = u(g);
  return p;
}

} // namespace

int main() {
  oracle_set_threads(2);
  std::mt19937_64 g(7);
  // dense kernels
  for (int f = 0x10; f <= 0x14; ++f)
    check(std::isfinite(oracle_wendland(f, 0.3, 2.0)), "wendland");
  {
    const int n = 12, p = 4;
    std::vector<double> Q = cloud(g, n, p, -1, 1), b = cloud(g, n, 1, -1, 1);
    std::vector<double> S(p), x(p);
    check(oracle_singular_values(Q.data(), n, p, S.data()) == 0 && finite(S),
          "singular values");
    check(oracle_lstsq(Q.data(), n, p, b.data(), x.data()) == 0 && finite(x),
          "lstsq");
    std::vector<double> A((size_t)n * n, 0.0), r = b;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        A[(size_t)i * n + j] = (i == j ? 4.0 : 0.0) + 1.0 / (1 + i + j);
    check(oracle_lu_solve(A.data(), n, r.data()) == 0 && finite(r), "lu");
  }
  // the reference path in 1-D, 2-D and 3-D, every basis
  for (int dim = 1; dim <= 3; ++dim) {
    const int ns = dim == 1 ? 300 : 1500, nt = ns / 2;
    std::vector<double> S = cloud(g, ns, dim, 0, 1), T = cloud(g, nt, dim, 0, 1);
    std::vector<double> V(2 * (size_t)ns);
    for (size_t i = 0; i < V.size(); ++i)
      V[i] = std::sin(3.0 * S[(i % ns) * dim]) + 0.1 * (double)(i / ns);
    for (int f = 0x10; f <= 0x14; ++f) {
      void *h = oracle_setup_create(dim, f, 30, 0.15, -1.0, S.data(), ns,
                                    T.data(), nt, 0, 0);
      check(h != nullptr, "setup");
      if (!h)
        continue;
      check(oracle_setup_radius(h) > 0, "radius");
      const int64_t K = oracle_setup_num_patches(h);
      for (int which = 0; which < 3; ++which) {
        const int64_t rows = which < 2 ? K : nt;
        std::vector<int64_t> off(rows + 1),
            idx(std::max<int64_t>(1, oracle_setup_csr_size(h, which)));
        oracle_setup_csr(h, which, off.data(), idx.data());
        check(off[rows] == oracle_setup_csr_size(h, which), "csr");
      }
      std::vector<double> w(std::max<int64_t>(1, oracle_setup_csr_size(h, 2)));
      oracle_setup_weights(h, w.data());
      check(finite(w), "weights");
      std::vector<double> out(2 * (size_t)nt);
      std::vector<int> pv(std::max<int64_t>(K, 1));
      check(oracle_setup_apply(h, V.data(), 2, 0, -1, out.data(), pv.data()) ==
                    0 &&
                finite(out),
            "apply");
      std::vector<double> gT(2 * (size_t)nt, 1.0), grad((size_t)nt * dim);
      check(oracle_setup_target_grad(h, V.data(), 2, gT.data(), grad.data()) ==
                    0 &&
                finite(grad),
            "target grad");
      oracle_setup_destroy(h);
    }
    // reference defaults through the drop-in restatement
    std::vector<double> out(nt);
    check(oracle_rbf_interpolate(dim, 0x11, 50, 0.15, 0.1, S.data(), ns,
                                 V.data(), T.data(), nt, out.data()) == 0 &&
              finite(out),
          "rbf_interpolate");
    // density-aware patches
    void *a = oracle_adaptive_create(dim, 0x11, 20, 8, 120, 0.15, 1.0,
                                     S.data(), ns, T.data(), nt);
    check(a != nullptr, "adaptive");
    if (a) {
      std::vector<double> o(2 * (size_t)nt);
      std::vector<int> pv(std::max<int64_t>(oracle_setup_num_patches(a), 1));
      check(oracle_setup_apply(a, V.data(), 2, 0, -1, o.data(), pv.data()) ==
                    0 &&
                finite(o),
            "adaptive apply");
      oracle_setup_destroy(a);
    }
    // nearest source and MLS
    std::vector<double> fo(nt);
    std::vector<int> st(nt);
    check(oracle_fe_nearest(dim, S.data(), ns, V.data(), T.data(), nt, 0,
                            fo.data(), st.data()) == 0,
          "fe nearest");
    const int k = dim == 1 ? 3 : dim == 2 ? 6 : 10;
    std::vector<int64_t> mi((size_t)nt * k);
    std::vector<double> mw((size_t)nt * k);
    check(oracle_mls_coefficients(dim, S.data(), ns, T.data(), nt, mi.data(),
                                  mw.data()) == k &&
              finite(mw),
          "mls");
  }
  if (failures) {
    std::fprintf(stderr, "sanitize_check: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("sanitize_check: ok\n");
  return 0;
}
