// Minimum-contrast fit of the RBF covariance C(d) = tau2 exp(-psi |d|^nu)
// to per-offset empirical covariances (estimate_kernel_parameters.h:122-151,
// 241-387).  Data preparation, weights (1/sqrt(multiplicity) of each
// distance), bounds, "fixed" parameters (+-1e-5 boxes) and the optional
// psi <= nu constraint follow the reference.  The reference minimises with
// NLopt's LN_COBYLA, which this build does not ship; the same objective is
// minimised here by a projected Levenberg-Marquardt iteration on the
// weighted residuals (feasible set = box, plus the half-space psi <= nu).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <stdexcept>
#include <vector>

namespace drhost {

struct KernelData {
  std::vector<double> x, y, w;  // distance, covariance, weight
};

inline KernelData rbf_lsq_data(const std::vector<float>& cov, const std::vector<float>& dist,
                               const std::vector<int>& npairs, double maxdist) {
  KernelData k;
  for (size_t i = 1; i < npairs.size(); ++i)
    if (npairs[i] != 0 && dist[i] <= maxdist) {
      k.x.push_back(std::fabs((double)dist[i]));
      k.y.push_back((double)cov[i]);
    }
  std::map<double, int> mult;
  for (double v : k.x) ++mult[v];
  k.w.resize(k.x.size());
  for (size_t i = 0; i < k.x.size(); ++i) k.w[i] = std::sqrt(1.0 / mult[k.x[i]]);
  return k;
}

inline double rbf_objective(const double* th, const KernelData& d) {  // _rbf_least_squares
  const double N = (double)d.x.size();
  double obj = 0.0;
  for (size_t i = 0; i < d.x.size(); ++i) {
    const double r = d.w[i] * (d.y[i] - th[0] * std::exp(-th[1] * std::pow(d.x[i], th[2])));
    obj += r * r / N;
  }
  return obj;
}

namespace detail {
inline void project(double* th, const double* lb, const double* ub, bool constrained) {
  for (int k = 0; k < 3; ++k) th[k] = std::min(std::max(th[k], lb[k]), ub[k]);
  if (constrained && th[1] > th[2]) {
    double m = 0.5 * (th[1] + th[2]);
    double a = std::min(std::max(m, lb[1]), ub[1]);
    double b = std::min(std::max(m, lb[2]), ub[2]);
    if (a > b) a = b = std::min(std::max(std::min(a, b), std::max(lb[1], lb[2])), std::min(ub[1], ub[2]));
    th[1] = a;
    th[2] = b;
  }
}
inline bool solve3(double A[3][3], double b[3], const bool* freep) {
  // Gaussian elimination on the free sub-block
  int idx[3], n = 0;
  for (int k = 0; k < 3; ++k)
    if (freep[k]) idx[n++] = k;
  double M[3][4];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) M[i][j] = A[idx[i]][idx[j]];
    M[i][n] = b[idx[i]];
  }
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    if (M[p][c] == 0.0) return false;
    for (int j = 0; j <= n; ++j) std::swap(M[c][j], M[p][j]);
    for (int r = 0; r < n; ++r)
      if (r != c) {
        const double f = M[r][c] / M[c][c];
        for (int j = c; j <= n; ++j) M[r][j] -= f * M[c][j];
      }
  }
  for (int k = 0; k < 3; ++k) b[k] = 0.0;
  for (int i = 0; i < n; ++i) b[idx[i]] = M[i][n] / M[i][i];
  return true;
}
}  // namespace detail

// compute_rbf_parameters: returns 0 on success, 1 on failure; theta in/out
inline int compute_rbf_parameters(std::vector<double>& theta, const std::vector<float>& cov,
                                  const std::vector<float>& dist, const std::vector<int>& npairs,
                                  bool constrained = true, double maxdist = 1e6, double variance = -1,
                                  double bandwidth = -1, double exponent = -1, double xtol = 1e-5) {
  const double tau_max = (double)cov.at(0), eps0 = 1e-5;
  double lb[3] = {eps0, eps0, eps0};
  double ub[3] = {tau_max, HUGE_VAL, 2.0 - eps0};
  int n_fixed = 0;
  if ((variance - eps0) > 0) {
    lb[0] = std::min(variance - eps0, std::max(tau_max - eps0, eps0));
    ub[0] = std::min(variance + eps0, tau_max);
    theta[0] = variance;
    ++n_fixed;
  }
  if ((exponent - eps0) > 0 && (exponent + eps0) <= 2.0) {
    lb[2] = exponent - eps0;
    ub[2] = exponent + eps0;
    theta[2] = exponent;
    theta[1] = exponent / 2;
    ++n_fixed;
  }
  if ((bandwidth - eps0) > 0) {
    lb[1] = bandwidth - eps0;
    ub[1] = bandwidth + eps0;
    theta[1] = bandwidth;
    ++n_fixed;
  }
  if (constrained && theta[1] > theta[2])
    throw std::domain_error("compute_rbf_parameters: requested constraint with bandwidth > exponent");
  if (n_fixed == 3) return 0;
  const KernelData d = rbf_lsq_data(cov, dist, npairs, maxdist);
  if (d.x.empty()) return 1;
  double th[3] = {theta[0], theta[1], theta[2]};
  detail::project(th, lb, ub, constrained);
  bool freep[3];
  for (int k = 0; k < 3; ++k) freep[k] = ub[k] - lb[k] > 2.5 * eps0;
  double f = rbf_objective(th, d);
  double lam = 1e-3;
  for (int it = 0; it < 2000; ++it) {
    double JtJ[3][3] = {{0}}, Jtr[3] = {0};
    for (size_t i = 0; i < d.x.size(); ++i) {
      const double xn = std::pow(d.x[i], th[2]);
      const double e = std::exp(-th[1] * xn);
      const double r = d.w[i] * (d.y[i] - th[0] * e);
      const double lx = d.x[i] > 0 ? std::log(d.x[i]) : 0.0;
      const double J[3] = {-d.w[i] * e, d.w[i] * th[0] * xn * e, d.w[i] * th[0] * th[1] * xn * lx * e};
      for (int a = 0; a < 3; ++a) {
        Jtr[a] += J[a] * r;
        for (int b = 0; b < 3; ++b) JtJ[a][b] += J[a] * J[b];
      }
    }
    // active set: a parameter on a bound whose descent direction points
    // outward is held for this iteration
    bool act[3];
    for (int k = 0; k < 3; ++k) {
      const double span = std::max(1.0, std::fabs(th[k])) * 1e-12;
      act[k] = freep[k] && !((th[k] <= lb[k] + span && Jtr[k] > 0) || (th[k] >= ub[k] - span && Jtr[k] < 0));
    }
    bool moved = false;
    for (int tries = 0; tries < 30; ++tries) {
      double A[3][3], g[3];
      for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) A[a][b] = JtJ[a][b];
        A[a][a] += lam * std::max(JtJ[a][a], 1e-300);
        g[a] = -Jtr[a];
      }
      if (!detail::solve3(A, g, act)) {
        lam *= 10;
        continue;
      }
      double tn[3] = {th[0] + g[0], th[1] + g[1], th[2] + g[2]};
      detail::project(tn, lb, ub, constrained);
      const double fn = rbf_objective(tn, d);
      if (fn < f) {
        double rel = 0.0;
        for (int k = 0; k < 3; ++k) rel = std::max(rel, std::fabs(tn[k] - th[k]) / (std::fabs(th[k]) + xtol));
        for (int k = 0; k < 3; ++k) th[k] = tn[k];
        f = fn;
        lam = std::max(lam / 3.0, 1e-12);
        moved = rel > xtol;
        break;
      }
      lam *= 4.0;
    }
    if (!moved) break;
  }
  if (!(std::isfinite(th[0]) && std::isfinite(th[1]) && std::isfinite(th[2]))) return 1;
  for (int k = 0; k < 3; ++k) theta[k] = th[k];
  return 0;
}

}  // namespace drhost
