// C entry points of the host layer, for tests and embedding
// (libdualres_host.so): the RBF minimum-contrast fit and NIfTI I/O.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dualres_host.h"
#include "nifti.hpp"
#include "rbf_fit.hpp"
#include "sampler.hpp"

extern "C" {

// compute_rbf_parameters (estimate_kernel_parameters.h:261-387); theta in/out.
// Returns 0 ok, 1 fit failure, 2 domain error (constraint violated).
int drh_rbf_fit(const float* cov, const float* dist, const int32_t* npairs, int64_t n, double* theta, int constrained,
                double maxdist, double variance, double bandwidth, double exponent, double xtol) {
  std::vector<float> c(cov, cov + n), d(dist, dist + n);
  std::vector<int> np(npairs, npairs + n);
  std::vector<double> th(theta, theta + 3);
  try {
    const int r = drhost::compute_rbf_parameters(th, c, d, np, constrained != 0, maxdist, variance, bandwidth, exponent,
                                                 xtol);
    for (int k = 0; k < 3; ++k) theta[k] = th[k];
    return r;
  } catch (const std::domain_error&) {
    return 2;
  }
}

double drh_rbf_objective(const float* cov, const float* dist, const int32_t* npairs, int64_t n, const double* theta,
                         double maxdist) {
  std::vector<float> c(cov, cov + n), d(dist, dist + n);
  std::vector<int> np(npairs, npairs + n);
  return drhost::rbf_objective(theta, drhost::rbf_lsq_data(c, d, np, maxdist));
}

// HMCParameters<double> (HMCParameters.h:30-297), the class the CLIs run,
// driven by a fixed sequence of n Metropolis-Hastings rates.  After update k,
// row k of out (n x 8) holds eps, eps_bar, h_bar, eps_target,
// integrator_steps, eps_start_found, warmup and metropolis_hastings_rate.
// Returns 0, or 2 when the constructor rejects the arguments.
int drh_hmc_trace(int burnin, int n_save, int thin, int integrator_steps, double starting_eps, double mh_target,
                  double gamma, double t0, double kappa, const double* rates, int n, double* out) {
  try {
    drhost::HMCParameters<double> h(burnin, n_save, thin, integrator_steps, starting_eps, mh_target, gamma, t0,
                                    kappa);
    for (int k = 0; k < n; ++k) {
      h.update(rates[k]);
      double* o = out + 8 * k;
      o[0] = h.eps_value();
      o[1] = h.eps_bar();
      o[2] = h.h_bar();
      o[3] = h.eps_target();
      o[4] = h.integrator_steps();
      o[5] = h.eps_start_found() ? 1 : 0;
      o[6] = h.warmup() ? 1 : 0;
      o[7] = h.metropolis_hastings_rate();
    }
    return 0;
  } catch (const std::domain_error&) {
    return 2;
  }
}

}  // extern "C"

namespace {
template <typename T>
void fit_chain(dr_chain* c, int nd, int64_t n_h, int64_t n_s, int64_t n_out, double tau2, int burnin, int nsave,
               int thin, int L, uint64_t seed, int64_t post_iters, int captured, double* out) {
  using namespace drhost;
  rng().seed((std::mt19937::result_type)seed);
  Chain chain(c, nd, (long)n_h, (long)n_out);
  HMCParameters<T> hmc(burnin, nsave, thin, L, (T)(1.0 / L), (T)0.65);
  // _initialize_sigma (MultiResParameters.h:764-778), as fit_model draws it
  std::vector<double> sig(nd);
  std::gamma_distribution<T> G((T)tau2, (T)1);
  std::uniform_real_distribution<T> U(0, 1);
  sig[0] = G(rng());
  for (int i = 1; i < nd; ++i) sig[i] = (T)sig[0] / U(rng());
  chain.set_sigma_sq_inv(sig);
  FitOptions opt;
  opt.quiet = true;
  opt.captured = captured != 0;
  NullStream quiet;
  std::ofstream none;
  std::vector<double> lat;
  std::vector<int> steps;
  McmcSummary<T> summ =
      sample_chain<T>(chain, (long)n_h, (long)n_s, hmc, sig, none, opt, quiet, &lat, (long)post_iters, &steps);
  double grads = 0;
  for (int l : steps) grads += l + 1;
  double wall = 0;
  for (double t : lat) wall += t;
  std::vector<double> sorted = lat;
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double q) {
    if (sorted.empty()) return 0.0;
    const double pos = q * (sorted.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - i;
    return i + 1 < sorted.size() ? sorted[i] * (1 - f) + sorted[i + 1] * f : sorted[i];
  };
  out[0] = wall;
  out[1] = grads;
  out[2] = hmc.metropolis_hastings_rate();
  out[3] = hmc.eps_value();
  out[4] = pct(0.5);
  out[5] = pct(0.99);
  out[6] = (double)lat.size();
}
}  // namespace

extern "C" {

// The fit loop of dualgpm / dualgpmf (fit_model's HMC iterations, Gibbs draws,
// dual averaging and device summaries, gaussian_process_model.h:134-171) on a
// chain created elsewhere with its data set, mu holding the data and its
// output indices set.  precision 0 runs the host scalars in double as
// dualgpm, 1 in float as dualgpmf.  post_iters > 0 stops after that many
// post-burn-in iterations; captured != 0 replays each iteration from a
// captured graph.  out[7]: post-burn-in wall seconds, gradient
// evaluations (the sum of L + 1 over the iterations), MH rate, final eps,
// p50 and p99 iteration seconds, iterations.
int drh_fit_chain(dr_chain* c, int precision, int n_datasets, int64_t n_h, int64_t n_s, int64_t n_out, double tau2,
                  int burnin, int nsave, int thin, int L, uint64_t seed, int64_t post_iters, int captured,
                  double* out) {
  try {
    if (precision == 0)
      fit_chain<double>(c, n_datasets, n_h, n_s, n_out, tau2, burnin, nsave, thin, L, seed, post_iters, captured, out);
    else
      fit_chain<float>(c, n_datasets, n_h, n_s, n_out, tau2, burnin, nsave, thin, L, seed, post_iters, captured, out);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "drh_fit_chain: %s\n", e.what());
    return 1;
  }
}

// read a NIfTI-1 volume: dims[3] and qform[16] out; data (nx*ny*nz) if non-null
int drh_read_nifti(const char* path, int* dims, float* qform, double* data) {
  try {
    const drhost::Image im = drhost::read_nifti(path);
    dims[0] = im.nx;
    dims[1] = im.ny;
    dims[2] = im.nz;
    std::memcpy(qform, im.qto_xyz.data(), 16 * sizeof(float));
    if (data) std::memcpy(data, im.data.data(), im.data.size() * sizeof(double));
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"
