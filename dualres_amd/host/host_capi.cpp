// C entry points of the host layer, for tests and embedding
// (libdualres_host.so): the RBF minimum-contrast fit and NIfTI I/O.
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "nifti.hpp"
#include "rbf_fit.hpp"

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
