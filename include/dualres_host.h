/* C ABI of the host layer, libdualres_host.so (dualres_amd/host/).
 *
 * The engine (dualres_hip.h) replaces MultiResParameters<T>; this library
 * holds what sits above it in the reference, written in C++ against the
 * engine: the fit_model loop with HMCParameters' dual averaging, the
 * minimum-contrast RBF fit, the NIfTI-1 reader, and a server that runs many
 * fits on one device.  Every entry point returns 0 on success; the non-zero
 * codes are given per function. */
#ifndef DUALRES_HOST_H
#define DUALRES_HOST_H

#include <stdint.h>

#include "dualres_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* compute_rbf_parameters (estimate_kernel_parameters.h:261-387): the
 * weighted minimum-contrast fit of theta = (tau^2, psi, nu) to n per-offset
 * covariances; theta holds the start and receives the fit.  variance,
 * bandwidth, exponent > 0 fix that parameter (the reference's +-1e-5 box).
 * Returns 0, 1 when the fit fails, 2 when the psi <= nu constraint is
 * violated (the reference's domain_error). */
int drh_rbf_fit(const float* cov, const float* dist, const int32_t* npairs, int64_t n, double* theta, int constrained,
                double maxdist, double variance, double bandwidth, double exponent, double xtol);
/* _rbf_least_squares (estimate_kernel_parameters.h:122-151) at theta. */
double drh_rbf_objective(const float* cov, const float* dist, const int32_t* npairs, int64_t n, const double* theta,
                         double maxdist);

/* HMCParameters<double> (HMCParameters.h:30-297) driven by n given
 * Metropolis-Hastings rates; row k of out (n x 8) holds, after update k,
 * eps, eps_bar, h_bar, eps_target, integrator_steps, eps_start_found,
 * warmup and metropolis_hastings_rate.  Returns 2 when the constructor
 * rejects the arguments. */
int drh_hmc_trace(int burnin, int n_save, int thin, int integrator_steps, double starting_eps, double mh_target,
                  double gamma, double t0, double kappa, const double* rates, int n, double* out);

/* The post-setup loop of fit_model (gaussian_process_model.h:134-171) on a
 * chain created with dr_create, its data set, mu holding the data and its
 * output indices set: HMC updates, Gibbs draws, dual averaging and device
 * summaries.  precision 0 keeps the host scalars in double (dualgpm), 1 in
 * float (dualgpmf).  post_iters > 0 stops after that many post-burn-in
 * iterations.  captured != 0 runs each iteration through dr_update_captured
 * (the same results bit for bit).  out[7]: post-burn-in wall seconds, gradient evaluations (the
 * sum of L + 1 over those iterations: the first runs one step when burn-in
 * ended before the step-size search did), MH rate, final eps, p50 and p99
 * iteration seconds, iterations.  Returns 1 on any error (printed to
 * stderr). */
int drh_fit_chain(dr_chain* c, int precision, int n_datasets, int64_t n_h, int64_t n_s, int64_t n_out, double tau2,
                  int burnin, int nsave, int thin, int L, uint64_t seed, int64_t post_iters, int captured,
                  double* out);

/* NIfTI-1 read (nifti_manipulation.inl's image loading): dims[3] and the
 * row-major qform[16] out, data (nx*ny*nz, x fastest) when non-null.
 * Returns 1 when the file cannot be read. */
int drh_read_nifti(const char* path, int* dims, float* qform, double* data);

/* --- fit server: many independent fits on one device -------------------
 * No counterpart in the reference, which runs one fit per process. */
typedef struct drh_request {
  int n_datasets;          /* 1: single resolution, 2: with the stdres image */
  const double* hr;        /* highres volume, x fastest; 0 / NaN = missing  */
  int hr_dims[3];
  float q_hr[16];          /* row-major qform                               */
  const double* sr;        /* stdres volume (n_datasets == 2)               */
  int sr_dims[3];
  float q_sr[16];
  double theta[3];         /* tau^2, psi, nu                                */
  double neighborhood;     /* kriging radius in mm; <= 0: 3 x largest voxel */
  int burnin, nsave, thin, L;
  double mh_target;
  uint64_t seed;
  int precision, state_mode; /* dr_precision, dr_state_mode                */
  int embed[3];
} drh_request;

typedef struct drh_result {
  int status;              /* dr_status; DR_ECANCELED after drh_server_cancel */
  double mh_rate, eps;
  double queue_s, run_s, total_s; /* submit -> start -> done               */
  int64_t grad_evals, device_bytes, n_out;
} drh_result;

/* workers > 0: at most that many fits at once; workers <= 0: as many as the
 * device byte budget admits, each on its own thread.  backfill lets a later
 * request that fits start while a larger one waits at the head. */
int drh_server_create(int64_t budget_bytes, int workers, int device, int backfill, void** out);
/* The request's volumes are copied.  A request that can never fit the
 * budget completes at once with DR_ENOMEM. */
int drh_server_submit(void* srv, const drh_request* q, int64_t* id);
/* 0: done, result filled and min(cap, n_out) posterior means and variances
 * at the highres data voxels copied; 1: timeout (the request stays). */
int drh_server_wait(void* srv, int64_t id, double timeout_s, drh_result* res, double* mean, double* var, int64_t cap,
                    char* err, int errlen);
/* 0: dropped while queued; 1: running fit told to stop at its next HMC
 * iteration; 2: unknown id or already finished. */
int drh_server_cancel(void* srv, int64_t id);
int drh_server_destroy(void* srv);
/* Device bytes the request holds at its peak (chain, W, kriging scratch). */
int drh_estimate_request_bytes(const drh_request* q, int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* DUALRES_HOST_H */
