// CPU baseline: the reference algorithm's solve (kernel/CG_solver.m with the
// kernel/Vcycle.m cycle) in FP64 C++ with OpenMP, on the index arrays and
// element matrices the oracle's set-up builds (oracle/gmgs_oracle.py).
//
// TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg (timed on the
// host cores) and by tests/ (checked against the numpy oracle).  The product
// never links or loads it.
//
// Data follow the reference's own layout (1-based ids converted to 0-based by
// the caller):
//   edof   Ne x 24 DOF ids per element, [x1 y1 z1 ... x8 y8 z8] (SetupSolver.m:13-15)
//   Ks     storing state 0: one 24x24 (Ke0); 1: one per element; 2: the
//          universal matrix then the unique elements' (MeshStruct.m:16-19)
//   tmat   27 x Ne intermediate-node ids of each coarse element's fine nodes
//          (BuildingMeshHierarchy.m:75-78), tcoef their multiplicities (:79-84),
//          c2f the intermediate id of every finer compact node (:85-89)
// AtX scatters in 24 passes, one per local DOF: within a pass the elements'
// targets are distinct (local corner k of two elements is two nodes), so each
// pass is a race-free parallel loop (AtX.m:13-15), as are the 8 and 27
// passes of the transfers (RestrictResidual.m:14-16, InterpolationDeviation.m:11-13).
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

typedef struct cpu_level {
  int64_t ne, nn;
  const int32_t* edof;    // ne x 24, 0-based
  int32_t state;          // 0, 1 or 2
  const double* Ks;       // see above
  const int32_t* kidx;    // state 2: matrix index per element (0 = universal)
  const double* diag;     // 3 nn (smoothed levels)
  const int32_t* fixed;   // 0-based fixed DOFs
  int64_t nfixed;
  const int32_t* enod;    // ne x 8, 0-based compact nodes
  int64_t ninter;         // levels >= 2: intermediate nodes
  const int32_t* tmat;    // 27 x ne, 0-based
  const double* tcoef;    // ninter
  const int32_t* c2f;     // nn of the finer level, 0-based
} cpu_level;

}  // extern "C"

namespace {

struct Solver {
  const cpu_level* lev;
  int L;
  const double* in_node;      // 27 x 8 trilinear weights (Operator4...m:5-31)
  const double* cholL;        // nfree x nfree row-major lower factor of the coarsest K_ff
  const int32_t* cfree;       // coarsest free DOFs
  int64_t nfree;
  double omega;
  std::vector<std::vector<double>> ve;   // per level: Ne x 24 element products

  void atx(int l, const double* x, double* y) {
    const cpu_level& lv = lev[l];
    double* v = ve[l].data();
    const int64_t ne = lv.ne;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      const int32_t* ed = lv.edof + 24 * e;
      const double* K = lv.state == 0 ? lv.Ks
                      : lv.state == 1 ? lv.Ks + 576 * e
                                      : lv.Ks + 576 * (int64_t)lv.kidx[e];
      double xe[24];
      for (int i = 0; i < 24; ++i) xe[i] = x[ed[i]];
      double* out = v + 24 * e;
      for (int j = 0; j < 24; ++j) out[j] = 0.0;
      for (int i = 0; i < 24; ++i) {
        const double xi = xe[i];
        const double* Ki = K + 24 * i;
        for (int j = 0; j < 24; ++j) out[j] += xi * Ki[j];
      }
    }
    const int64_t n = 3 * lv.nn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int k = 0; k < 24; ++k) {
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < ne; ++e) y[lv.edof[24 * e + k]] += v[24 * e + k];
    }
    for (int64_t i = 0; i < lv.nfixed; ++i) y[lv.fixed[i]] = 0.0;     // AtX.m:57
  }

  // JacobiSmoother.m: x += omega r / d; r -= AtX(omega r / d) when asked
  void jacobi(int l, double* x, double* r, bool with_residual, std::vector<double>& dx, std::vector<double>& t) {
    const cpu_level& lv = lev[l];
    const int64_t n = 3 * lv.nn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      dx[i] = omega * r[i] / lv.diag[i];
      x[i] += dx[i];
    }
    if (!with_residual) return;
    atx(l, dx.data(), t.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] -= t[i];
  }

  // RestrictResidual.m: level l (coarse) from level l-1
  void restrict_to(int l, const double* rf, double* rc) {
    const cpu_level& lv = lev[l];
    const cpu_level& fine = lev[l - 1];
    std::vector<double> r1(lv.ninter);
    const int64_t ne = lv.ne;
    std::vector<double> t(8 * ne);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * lv.nn; ++i) rc[i] = 0.0;
    for (int c = 0; c < 3; ++c) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < lv.ninter; ++i) r1[i] = 0.0;
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < fine.nn; ++k) r1[lv.c2f[k]] = rf[3 * k + c];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < lv.ninter; ++i) r1[i] /= lv.tcoef[i];
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < ne; ++e) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 27; ++q) {
          const double v = r1[lv.tmat[q * ne + e]];
          for (int k = 0; k < 8; ++k) acc[k] += v * in_node[8 * q + k];
        }
        for (int k = 0; k < 8; ++k) t[8 * e + k] = acc[k];
      }
      for (int k = 0; k < 8; ++k) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) rc[3 * (int64_t)lv.enod[8 * e + k] + c] += t[8 * e + k];
      }
    }
  }

  // InterpolationDeviation.m: level l-1 (fine) from level l
  void interpolate_from(int l, const double* xc, double* xf_out) {
    const cpu_level& lv = lev[l];
    const cpu_level& fine = lev[l - 1];
    std::vector<double> xf(lv.ninter);
    const int64_t ne = lv.ne;
    std::vector<double> t1(27 * ne);
    for (int c = 0; c < 3; ++c) {
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < ne; ++e) {
        double xe[8];
        for (int k = 0; k < 8; ++k) xe[k] = xc[3 * (int64_t)lv.enod[8 * e + k] + c];
        for (int q = 0; q < 27; ++q) {
          double s = 0.0;
          for (int k = 0; k < 8; ++k) s += in_node[8 * q + k] * xe[k];
          t1[q * ne + e] = s;
        }
      }
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < lv.ninter; ++i) xf[i] = 0.0;
      for (int q = 0; q < 27; ++q) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) xf[lv.tmat[q * ne + e]] += t1[q * ne + e];
      }
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < fine.nn; ++k) xf_out[3 * k + c] = xf[lv.c2f[k]] / lv.tcoef[lv.c2f[k]];
    }
  }

  // Vcycle.m:21-23: x(free) = L^-T L^-1 r(free)
  void coarsest(const double* r, double* x) {
    const cpu_level& lv = lev[L - 1];
    const int64_t n = nfree;
    std::vector<double> y(n);
    for (int64_t i = 0; i < 3 * lv.nn; ++i) x[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double* Li = cholL + i * n;
      double s = r[cfree[i]];
      for (int64_t j = 0; j < i; ++j) s -= Li[j] * y[j];
      y[i] = s / Li[i];
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      const double v = y[i] / cholL[i * n + i];
      y[i] = v;
      // column sweep of L^T: subtract v * L(i, j) from the rows j < i
#pragma omp parallel for schedule(static) if (i > 4096)
      for (int64_t j = 0; j < i; ++j) y[j] -= cholL[i * n + j] * v;
    }
    for (int64_t i = 0; i < n; ++i) x[cfree[i]] = y[i];
  }

  struct Work {
    std::vector<std::vector<double>> x, r, dx, t;
  };

  void vcycle(const double* r, double* z, Work& w) {
    for (int l = 0; l < L; ++l) std::fill(w.x[l].begin(), w.x[l].end(), 0.0);
    std::memcpy(w.r[0].data(), r, sizeof(double) * 3 * lev[0].nn);
    for (int l = 1; l < L; ++l) {
      jacobi(l - 1, w.x[l - 1].data(), w.r[l - 1].data(), true, w.dx[l - 1], w.t[l - 1]);
      restrict_to(l, w.r[l - 1].data(), w.r[l].data());
    }
    coarsest(w.r[L - 1].data(), w.x[L - 1].data());
    for (int l = L - 1; l >= 1; --l) {
      interpolate_from(l, w.x[l].data(), w.t[l - 1].data());
      const int64_t n = 3 * lev[l - 1].nn;
      double* xl = w.x[l - 1].data();
      const double* tl = w.t[l - 1].data();
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) xl[i] += tl[i];
      jacobi(l - 1, xl, w.r[l - 1].data(), false, w.dx[l - 1], w.t[l - 1]);
    }
    std::memcpy(z, w.x[0].data(), sizeof(double) * 3 * lev[0].nn);
    for (int64_t i = 0; i < lev[0].nfixed; ++i) z[lev[0].fixed[i]] = 0.0;   // Vcycle.m:35
  }
};

double dot(const double* a, const double* b, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

extern "C" {

// CG_solver.m:1-56 with the Jacobi step (B1), Fletcher-Reeves beta and the
// reference cycle; returns the iterations (negative: not converged) and
// writes x, the final relres and the solve's wall seconds.
int cpu_cg_solve(const cpu_level* lev, int L, const double* in_node, const double* cholL,
                 const int32_t* cfree, int64_t nfree, double omega, const double* b, double tol,
                 int max_it, double* x, double* relres, double* seconds) {
  Solver S{lev, L, in_node, cholL, cfree, nfree, omega, {}};
  S.ve.resize(L);
  Solver::Work w;
  w.x.resize(L); w.r.resize(L); w.dx.resize(L); w.t.resize(L);
  for (int l = 0; l < L; ++l) {
    S.ve[l].resize(24 * (size_t)lev[l].ne);
    const size_t n = 3 * (size_t)lev[l].nn;
    w.x[l].resize(n); w.r[l].resize(n); w.dx[l].resize(n); w.t[l].resize(n);
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = 3 * lev[0].nn;
  std::vector<double> r1(n), r0(n), z1(n), z2(n), p(n), q(n), jdx(n), jt(n);
  const double normB = std::sqrt(dot(b, b, n));
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  S.atx(0, x, q.data());
  for (int64_t i = 0; i < n; ++i) r1[i] = b[i] - q[i];
  int its = 0;
  bool converged = false;
  double res = INFINITY, rz_prev = 0.0;
  while (its <= max_it) {
    S.jacobi(0, x, r1.data(), true, jdx, jt);        // CG_solver.m:18-19
    S.vcycle(r1.data(), z2.data(), w);               // :21
    ++its;
    const double rz = dot(r1.data(), z2.data(), n);
    if (its == 1) {
      std::memcpy(p.data(), z2.data(), sizeof(double) * n);
    } else {
      const double beta = rz / rz_prev;              // :26 (r0'z1 of the previous iteration)
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) p[i] = z2[i] + beta * p[i];
    }
    S.atx(0, p.data(), q.data());
    const double alpha = rz / dot(p.data(), q.data(), n);
    double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r1[i] -= alpha * q[i];
      rr += r1[i] * r1[i];
    }
    res = std::sqrt(rr) / normB;
    if (res < tol) { converged = true; break; }
    rz_prev = rz;
  }
  *relres = res;
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return converged ? its : -its;
}

int cpu_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
