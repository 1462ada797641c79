// C entry points of the host layer, for tests and embedding
// (libdualres_host.so): the RBF minimum-contrast fit and NIfTI I/O.
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

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
