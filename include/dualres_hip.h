/*
 * dualres_hip.h -- C ABI of the MI355X (gfx950) engine for the HMC posterior
 * sampler of the dual-resolution Gaussian-process image model.
 *
 * Every entry point returns a dr_status; on failure a message is available
 * from dr_last_error() (thread-local).  No C++ exception crosses this ABI.
 * Inputs are borrowed host pointers copied during the call; outputs go to
 * caller-allocated host buffers.  A dr_chain is single-threaded: calls on
 * one handle must not overlap.  Citations name the reference function each
 * entry point replaces (paths relative to the asw221/dualres checkout).
 */
#ifndef DUALRES_HIP_H_
#define DUALRES_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; mirror the std:: exception kinds the reference throws. */
enum dr_status {
  DR_OK = 0,
  DR_EDOMAIN = 1,  /* std::domain_error  (e.g. MultiResParameters.h:283-284) */
  DR_ELOGIC = 2,   /* std::logic_error   (e.g. MultiResParameters.h:366-367) */
  DR_ERUNTIME = 3, /* std::runtime_error                                     */
  DR_EHIP = 4,     /* HIP runtime failure                                    */
  DR_ENOMEM = 5,   /* device or host allocation failed                       */
  DR_EBUSY = 6,    /* serving: admission refused / no slot                   */
  DR_ECANCELED = 7 /* serving: the request was cancelled                     */
};

enum dr_precision {
  DR_FP64 = 0,     /* gradient, FFTs and state in float64 (dualgpm)          */
  DR_FP32 = 1      /* FFTs/state/gradient in float32, eigenvalues, energies,
                      accept ratio and kriging solves in float64            */
};

enum dr_state_mode {
  DR_STATE_COMPAT = 0, /* also evolve the data-free imaginary chain so that the
                          energies and accept ratio equal the reference's
                          complex-state ones (MultiResParameters.h:428-476,
                          551-603, 783-815)                                 */
  DR_STATE_REAL = 1    /* real state only (the imaginary chain is dropped)  */
};

enum dr_covfun { DR_COV_RBF = 0, DR_COV_RQ = 1 };

/* Chain configuration.  Replaces the MultiResParameters<T> constructor
 * arguments (MultiResParameters.h:45-53, :172-300). */
typedef struct dr_config {
  int n_datasets;      /* 1 (highres only) or 2 (highres + stdres)          */
  int precision;       /* dr_precision                                      */
  int state_mode;      /* dr_state_mode                                     */
  int covfun;          /* dr_covfun                                         */
  int embed[3];        /* embedding torus (M0,M1,M2); all 0 = reference rule
                          2^ceil(log2(2*max_ijk)) (MultiResParameters.h:191-207);
                          otherwise each must be even and >= 2*max_ijk      */
  double theta[3];     /* covariance parameters (variance tau^2, bandwidth
                          psi, exponent nu), reference order                */
  float qform[16];     /* highres qto_xyz, row-major 4x4 (nifti_manipulation.inl:543-551) */
  uint64_t seed;       /* Philox key for device normals                     */
  int device;          /* HIP device ordinal                                */
} dr_config;

typedef struct dr_chain dr_chain;

/* Result of one HMC iteration (MultiResParameters.h:673-727). */
typedef struct dr_update_result {
  double mh_rate;      /* min(1, R), the value update() returns            */
  double R;            /* exp(H1 - H0), NaN mapped to 0                    */
  double H0, H1;       /* initial / proposed total energies                */
  double log_prior0, log_lik0, e0;  /* pieces of H0                         */
  double log_prior1, log_lik1, k1;  /* pieces of H1                         */
  double sse_h, sse_s; /* ||Yh - mu[I]||^2, ||Ys - W mu[I]||^2 of the state
                          kept after the accept step (Gibbs rates,
                          MultiResParameters.h:826-837)                     */
  int accepted;
} dr_update_result;

const char* dr_last_error(void);
int dr_version(void);
int dr_device_count(int* n);

/* Constructor: spectrum, low-rank set, initial mu (a GP draw / 10).
 * ijk: n_h x 3 row-major data indices relative to the input-mask bounding
 * box (fit_model, gaussian_process_model.h:107-113).
 * Errors: DR_EDOMAIN when >= N/2 eigenvalues are nonpositive
 * (MultiResParameters.h:283-284) or the embedding is invalid. */
int dr_create(const dr_config* cfg, const int32_t* ijk, int64_t n_h, dr_chain** out);
int dr_destroy(dr_chain* c);
/* Upper bound on the device bytes dr_create + dr_set_data allocate for this
 * configuration (W counted in its CSR form), for admission control by a
 * scheduler that shares one device between fits.  No counterpart in the
 * reference. */
int dr_estimate_bytes(const dr_config* cfg, const int32_t* ijk, int64_t n_h, int64_t n_s, int64_t w_nnz,
                      int64_t* bytes);

/* Spectrum report (MultiResParameters.h:258-281): embedding dims, N, the
 * number of eigenvalues with real part <= 0 and lambda[0]/N. */
int dr_spectrum_report(dr_chain* c, int dims[3], int64_t* N, int64_t* n_nonpositive, double* max_eig);
/* Real parts of lambda on the full torus (length N, x fastest). */
int dr_get_lambda(dr_chain* c, double* lambda_out);

/* Data (MultiResData.h:96-127): Yh in nonzero order (length n_h), Ys
 * (length n_s) and W as row-major CSR with columns indexing Yh.  For
 * n_datasets == 1 pass n_s = 0 and NULL W arrays. */
int dr_set_data(dr_chain* c, const double* Yh, int64_t n_h, const double* Ys, int64_t n_s,
                const int64_t* W_rowptr, const int32_t* W_cols, const double* W_vals);

/* Device layout of W.  By default the largest class of rows that are equal
 * up to a translation of the image grid (the interior kriging rows) is stored
 * once as a stencil, and only the other rows keep explicit entries (3D
 * offsets packed in 16 bits); W^T v of the stencil rows runs per residue class
 * of the stencil lattice.  force_csr = 1 keeps plain CSR / transposed CSR.
 * Sums differ from CSR only in association order.  Call before dr_set_data.
 * No counterpart in the reference. */
int dr_w_layout(dr_chain* c, int force_csr);
/* stats[0] = stencil layout active; [1] interior (stencil) rows; [2] stencil
 * entries; [3] boundary rows; [4] boundary entries; [5] nnz(W); [6] data
 * voxels touched by boundary rows; [7] voxel tiles of the interior W^T v pass
 * (tiles holding a data voxel); [8] lattice cells staged per tile; [9] voxels
 * per tile; [10] far rows (boundary rows too wide to pack, counted in [3]).
 * stats must hold 11 entries. */
int dr_w_stats(dr_chain* c, int64_t* stats);

/* mcmc_summary<T> (gaussian_process_model.h:191-224) kept on the device:
 * dr_summary_update adds mu, mu^2 and |mu| / max(max|mu|, 1e-6) of the
 * current state at the output voxels to float64 running sums (the reference
 * accumulates in T on the host).  set_output_indices and dr_summary_reset
 * clear the sums; dr_summary_read copies them out (NULL skips an array). */
int dr_summary_reset(dr_chain* c);
int dr_summary_update(dr_chain* c);
int dr_summary_read(dr_chain* c, double* sum_mu, double* sum_mu2, double* sum_abs, int64_t n_out,
                    int64_t* n_updates);

/* MultiResParameters::add_data_to_mu (:631-641). */
int dr_add_data_to_mu(dr_chain* c);
/* MultiResParameters::set_output_indices (:646-667). */
int dr_set_output_indices(dr_chain* c, const int32_t* ijk, int64_t n_out);
/* MultiResParameters::mu_ (:402-408): Re(mu) at the output indices. */
int dr_get_mu_output(dr_chain* c, double* out, int64_t n_out);
/* MultiResParameters::mu (:393-397): Re(mu) at the data indices. */
int dr_get_mu_data(dr_chain* c, double* out, int64_t n_h);

/* Noise precisions sigma^-2 (length n_datasets).  The reference draws them
 * in _initialize_sigma (:764-778) and _update_sigma_sq_inv (:820-859) from
 * its host RNG; here the host passes the drawn values in. */
int dr_set_sigma_sq_inv(dr_chain* c, const double* s);
int dr_get_sigma_sq_inv(dr_chain* c, double* s);

/* Prior precision tau^-2 (_update_tau_sq_inv, MultiResParameters.h:863-872,
 * compiled out of both reference binaries by DUALRES_DONT_UPDATE_TAU,
 * src/CMakeLists.txt:44,59).  Setting it rescales lambda by
 * tau_sq_inv / (previous tau^-2), the reference's direction; it starts at
 * 1 / theta[0].  The host draws the new value, Gamma(n_h/2 + 1, rate
 * dr_mu_data_sumsq / 2), after the sigma draws.  tau() = sqrt(1 / tau^-2). */
int dr_set_tau_sq_inv(dr_chain* c, double tau_sq_inv);
int dr_get_tau_sq_inv(dr_chain* c, double* tau_sq_inv);
/* _real_mu.squaredNorm() (:866-867): sum of Re(mu)^2 over the data voxels
 * of the current state. */
int dr_mu_data_sumsq(dr_chain* c, double* out);

/* One HMC iteration (MultiResParameters::update, :673-727) with the
 * momentum drawn from the device Philox stream (directly as the spectrum
 * F(z) of a white field z, Hermitian white noise) and the host-drawn accept
 * uniform u_accept in [0,1).  Does not perform the Gibbs draws. */
int dr_update(dr_chain* c, double eps, int L, double u_accept, dr_update_result* res);
/* Same with injected momentum normals z = z_re + i z_im on the full torus
 * (length N each, x fastest); z_im is ignored in DR_STATE_REAL. */
int dr_update_with_noise(dr_chain* c, double eps, int L, double u_accept,
                         const double* z_re, const double* z_im, dr_update_result* res);

/* dr_update through a captured HIP graph: the iteration's device work
 * (momentum draw, the L leapfrog steps, the energy sums) is captured once per
 * leapfrog length and replayed, reading the step size, noise precisions and
 * noise stream of the iteration from device memory.  Same kernels, scalars
 * and results as dr_update, bit for bit.  No counterpart in the reference
 * (its L-step loop is MultiResParameters.h:691-700). */
int dr_update_captured(dr_chain* c, double eps, int L, double u_accept, dr_update_result* res);

/* FFT pass shapes picked by timing (an autotuner in the role of FFTW's plan
 * search and wisdom, sized_fft.h:259-269 and src/preplan_fft.cpp:84-156).
 * params[8]: the tile lines (log2) of the fused z pass, the inverse y and x
 * passes and the forward x (residual) and y passes of a leapfrog step, then
 * the radix order of the x, y and z runtime stages (0 largest radix first,
 * 1 odd radices first, 2 smallest first; lengths with a compile-time stage
 * pipeline keep theirs).  dr_autotune times each choice on the chain's own
 * buffers (a 2-step leapfrog from the current state, which it leaves as it
 * was), unless cache_path names a JSON file that holds an entry for this
 * device architecture, precision, state mode, dataset count, torus and image
 * region; a timed result is added to that file.  The tile choice changes
 * only the blocking of the energy reductions, the radix order the rounding
 * of the transforms.  dr_set_tuning applies params (-1: built-in rule). */
int dr_autotune(dr_chain* c, const char* cache_path, int* params, int* from_cache);
int dr_set_tuning(dr_chain* c, const int* params);
int dr_get_tuning(dr_chain* c, int* params);

/* --- batches: many chains of uneven size advanced together --------------- */
enum dr_batch_mode {
  DR_BATCH_GRAPH = 0,   /* one graph per iteration: every chain's captured
                           iteration a parallel branch, one launch          */
  DR_BATCH_STREAMS = 1, /* each chain's captured iteration on its own stream */
  DR_BATCH_EAGER = 2    /* plain kernel launches on each chain's stream      */
};
typedef struct dr_batch dr_batch;
/* The chains (created, data set, on one device) are borrowed until
 * dr_batch_destroy and must not be updated on their own meanwhile; their
 * other calls (sigma, summaries, outputs) stay available between batch
 * updates.  No counterpart in the reference, which runs one chain per
 * process (gaussian_process_model.h:87-177). */
int dr_batch_create(dr_chain* const* chains, int n, int mode, dr_batch** out);
/* One HMC iteration of every chain i with step size eps[i], leapfrog length
 * L[i] and accept uniform u_accept[i]; res[i] as dr_update's.  A chain's
 * results are the same bit for bit in any batch and mode and through
 * dr_update_captured. */
int dr_batch_update(dr_batch* b, const double* eps, const int* L, const double* u_accept, dr_update_result* res);
int dr_batch_destroy(dr_batch* b);

/* Energies of the current state: log_prior (:428-476), log_likelihood
 * (:488-514), log_posterior (:528-544). */
int dr_log_prior(dr_chain* c, double* out);
int dr_log_likelihood(dr_chain* c, double* out);
int dr_log_posterior(dr_chain* c, double* out);
/* Current ||Yh - mu[I]||^2 and ||Ys - W mu[I]||^2. */
int dr_sse(dr_chain* c, double* sse_h, double* sse_s);

/* --- deterministic test hooks ------------------------------------------ */
/* Full-torus state in real space (length N each, x fastest). */
int dr_set_state(dr_chain* c, const double* mu_re, const double* mu_im,
                 const double* p_re, const double* p_im);
int dr_get_state(dr_chain* c, double* mu_re, double* mu_im, double* p_re, double* p_im);
/* Gradient at the current mu (MultiResParameters::gradient, :382-389),
 * full torus, real and imaginary parts. */
int dr_gradient(dr_chain* c, double* g_re, double* g_im);
/* L leapfrog steps from the current (mu, p) with the mass matrix of the
 * current sigma (MultiResParameters.h:690-700); the end point becomes the
 * current state.  Energies before/after are returned in res (e0 and k1 are
 * the kinetic terms of the initial and final momentum). */
int dr_leapfrog(dr_chain* c, double eps, int L, dr_update_result* res);

/* Timing: average device time (ms) of the last n_events leapfrog steps
 * recorded with HIP events on the chain's stream (0 if none). */
int dr_profile_enable(dr_chain* c, int enable);
int dr_profile_read(dr_chain* c, double* ms_per_grad, int64_t* n_grads,
                    double* ms_fft_passes, double* ms_w);
/* Per-kernel average device time (ms) of one leapfrog step, in launch order:
 * inverse z, y, x passes; [W mu, W^T v]; forward x, y, z passes. */
int dr_profile_kernels(dr_chain* c, double* ms, int max_k, int* n_k, int64_t* n_steps);

/* --- kriging matrix build (kriging_matrix.h:147-343) --------------------- */
typedef struct dr_csr dr_csr;
/* hr/std: dense float64 volumes (nx*ny*nz, x fastest), 0 or NaN = missing,
 * already masked.  theta as in dr_config.  Pattern is the reference's
 * (neighbour order, flat-offset range check); weights are float64 solves
 * normalised by their sum (rows with a NaN sum are left empty).  The CSR
 * is held in the device's memory until dr_csr_free; dr_csr_copy copies it
 * out (rowptr: nrow + 1 entries; cols and vals: nnz, may be null when
 * nnz == 0). */
int dr_kriging_build(const double* hr, const int hr_dims[3], const float q_hr[16],
                     const double* std_img, const int std_dims[3], const float q_std[16],
                     const double theta[3], double radius, int device, dr_csr** out);
int dr_csr_info(dr_csr* m, int64_t* nrow, int64_t* ncol, int64_t* nnz, int64_t* n_unique_solves);
int dr_csr_copy(dr_csr* m, int64_t* rowptr, int32_t* cols, double* vals);
/* Unique systems solved by Cholesky and by the pivoted LU that takes the
 * indefinite ones (tau^2 > 1 with K's unit diagonal).  No counterpart in the
 * reference, whose colPivHouseholderQr takes every system
 * (kriging_matrix.h:309-313); a diagnostic of which solver ran. */
int dr_csr_solver_counts(dr_csr* m, int64_t* n_cholesky, int64_t* n_lu);
int dr_csr_free(dr_csr* m);

/* --- minimum contrast summary (estimate_kernel_parameters.h:53-101) ------- */
/* Number of MCE offsets (perturbation_matrix.h:62-106; 25,348). */
int dr_mce_offsets(int64_t* n, int32_t* offsets_xyz /* may be NULL; else n*3 */);
/* img: dense float64 volume (x fastest), 0/NaN missing. Outputs per offset:
 * covariance (float), distance (float), npairs (int). */
int dr_mce_summary(const double* img, const int dims[3], const float qform[16], int device,
                   float* covariance, float* distance, int32_t* npairs);
/* Allocation-free dr_mce_summary.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  d_img is a float64 volume in DEVICE memory
 * (x fastest, 0/NaN missing); d_covariance, d_distance and d_npairs are
 * caller-allocated DEVICE arrays with one entry per MCE offset
 * (dr_mce_offsets, 25,348), written with dr_mce_summary's values.  All
 * intermediates live in one fixed scratch array of 48 MiB per device, held
 * until the process ends: images whose zero-padded torus needs more (about
 * 70^3 voxels and up) are rejected with DR_EDOMAIN, and no device memory is
 * ever allocated or freed.  Work is enqueued on `stream` (a hipStream_t;
 * NULL = the null stream) and the call returns without waiting.  The first
 * call on a device also looks up the scratch address and uploads the offset
 * table (synchronously), so it must not be made under graph capture.  Calls
 * that overlap on one device race on the shared scratch. */
int dr_mce_summary_noalloc(const double* d_img, const int dims[3], const float qform[16], int device, void* stream,
                           float* d_covariance, float* d_distance, int32_t* d_npairs);
/* Binned empirical covariance: nbins distance bins on [0, maxdist] over
 * all offsets with |offset| <= maxdist (FP64 sums per bin). */
int dr_mce_binned(const double* img, const int dims[3], const float qform[16], int device,
                  int nbins, double maxdist, double* bin_cov, double* bin_dist, int64_t* bin_npairs);

#ifdef __cplusplus
}
#endif
#endif /* DUALRES_HIP_H_ */
