// Host-side sampler control: HMCParameters (include/dualres/HMCParameters.h),
// mcmc_summary and fit_model (include/dualres/gaussian_process_model.h),
// driving the device chain through the C ABI.  The host owns the
// std::mt19937 for the scalar draws the reference takes from its global RNG
// (eps jitter, accept uniform, Gibbs gammas, sigma initialisation), in the
// reference's order; the bulk normals come from the device Philox stream.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <stdexcept>
#include <vector>

#include "engine.hpp"
#include "nifti.hpp"

namespace drhost {

// the reference's global generator (defines.h:78); one per thread, so fits
// running side by side on worker threads draw independent sequences
inline std::mt19937& rng() {
  thread_local std::mt19937 g(42);
  return g;
}

// a stream that discards what is written (quiet fits)
struct NullStream : std::ostream {
  struct Buf : std::streambuf {
    int overflow(int c) override { return c; }
  } buf;
  NullStream() : std::ostream(&buf) {}
};

template <typename RealType>
class HMCParameters {
 public:
  typedef RealType value_type;
  HMCParameters(int burnin = 1000, int n_save = 500, int thin = 1, int integrator_steps = 10,
                value_type starting_eps = 0.1, value_type mh_target = 0.65, value_type gamma = 0.05,
                value_type t0 = 10, value_type kappa = 0.75) {
    if (burnin < 0) throw std::domain_error("HMCParameters: burnin must be >= 0");
    if (n_save < 0) throw std::domain_error("HMCParameters: n_save must be >= 0");
    if (thin < 0) throw std::domain_error("HMCParameters: thin must be >= 0");
    if (integrator_steps <= 0) throw std::domain_error("HMCParameters: integrator_steps must be > 0");
    if (starting_eps <= 0) throw std::domain_error("HMCParameters: starting_eps must be > 0");
    if (mh_target <= 0 || mh_target >= 1) throw std::domain_error("HMCParameters: mh_target should be within (0, 1)");
    if (gamma <= 0 || gamma >= 1) throw std::domain_error("HMCParameters: gamma should be within (0, 1)");
    if (t0 < 0) throw std::domain_error("HMCParameters: t0 should be >= 0");
    if (kappa <= 0 || kappa >= 1) throw std::domain_error("HMCParameters: kappa should be within (0, 1)");
    warmup_ = burnin > 0;
    eps_start_found_ = !warmup_;
    burnin_ = burnin;
    L_ = integrator_steps;
    n_save_ = n_save;
    thin_ = thin == 0 ? 1 : thin;
    eps_ = std::max((double)starting_eps, eps0_);
    gamma_ = gamma;
    mh_target_ = mh_target;
    t0_ = t0;
    kappa_ = kappa;
  }
  bool save_iteration() const { return iteration_ > burnin_ && (iteration_ - burnin_) % thin_ == 0; }
  int integrator_steps() const { return eps_start_found_ ? L_ : 1; }
  int iteration() const { return iteration_; }
  int max_iterations() const { return burnin_ + thin_ * n_save_; }
  int n_save() const { return n_save_; }
  int burnin() const { return burnin_; }
  value_type eps() const {
    std::uniform_real_distribution<> U(0.9, 1.1);
    if (warmup_) return (value_type)eps_;
    return (value_type)std::max(eps_ * U(rng()), eps0_);
  }
  value_type eps_value() const { return (value_type)eps_; }
  // dual-averaging state (HMCParameters.h:76-86), read by the host tests
  double eps_bar() const { return eps_bar_; }
  double h_bar() const { return h_bar_; }
  double eps_target() const { return eps_target_; }
  bool eps_start_found() const { return eps_start_found_; }
  bool warmup() const { return warmup_; }
  value_type metropolis_hastings_rate() const { return (value_type)(mh_sum_ / std::max(iteration_ - burnin_, 1)); }
  void update(const value_type& mh_rate) {  // HMCParameters.h:259-297
    if (iteration_ == burnin_) {
      warmup_ = false;
      eps_start_found_ = true;
      if (eps_bar_ > eps0_) eps_ = eps_bar_;
    }
    if (warmup_) {
      if (!eps_start_found_) {
        if (iteration_ == 0) A_ = 2 * int(mh_rate > 0.5) - 1;
        if (std::pow(mh_rate, A_) > std::pow(2.0, -A_)) eps_ *= std::pow(2.0, A_);
        else {
          eps_start_found_ = true;
          eps_target_ = std::log(10.0 * eps_);
        }
      } else {
        h_bar_ = (1 - 1 / (iteration_ + t0_)) * h_bar_ + (mh_target_ - mh_rate) / (iteration_ + t0_);
        eps_ = std::exp(eps_target_ - std::sqrt(static_cast<value_type>(iteration_)) * h_bar_ / gamma_);
        eps_ = std::max(eps_, eps0_);
        eps_bar_ = std::exp(std::pow(static_cast<value_type>(iteration_), -kappa_) * std::log(eps_) +
                            (1 - std::pow(static_cast<value_type>(iteration_), -kappa_)) * std::log(eps_bar_));
      }
    } else {
      mh_sum_ += mh_rate;
    }
    ++iteration_;
  }

 private:
  bool warmup_ = false, eps_start_found_ = false;
  int burnin_ = 0, L_ = 1, iteration_ = 0, n_save_ = 0, thin_ = 1;
  double A_ = 1, eps_ = 0.1, eps0_ = 1e-5, eps_bar_ = 1, eps_target_ = 0, gamma_ = 0.05, h_bar_ = 0,
         mh_sum_ = 0, mh_target_ = 0.65, t0_ = 10, kappa_ = 0.75;
};

// running moments (gaussian_process_model.h:191-302).  The moments of mu are
// accumulated on the device in float64 (dr_summary_update) and copied out by
// finish(); sigma and the log-likelihood sums stay on the host.
template <typename T>
struct McmcSummary {
  std::vector<double> m1, m1abs, m2;
  std::vector<T> sigma;
  T log_lik = 0, mh_rate = 0, sampling_time = 0;
  int samples_ = 0;
  McmcSummary(long n, int ns) : m1(n, 0), m1abs(n, 0), m2(n, 0), sigma(ns, 0) {}
  void update(Chain& chain, const std::vector<T>& s, T ll) {
    check(dr_summary_update(chain.handle()));
    for (size_t i = 0; i < s.size(); ++i) sigma[i] += s[i];
    log_lik += ll;
    ++samples_;
  }
  void finish(Chain& chain) {
    int64_t n = 0;
    check(dr_summary_read(chain.handle(), m1.data(), m2.data(), m1abs.data(), (int64_t)m1.size(), &n));
  }
  int samples() const { return samples_ <= 0 ? 1 : samples_; }
  std::vector<double> mode_mu() const {
    std::vector<double> o(m1.size());
    for (size_t i = 0; i < o.size(); ++i) o[i] = m1[i] / samples();
    return o;
  }
  std::vector<double> var_mu() const {
    std::vector<double> o(m1.size());
    for (size_t i = 0; i < o.size(); ++i) o[i] = (m2[i] - m1[i] * m1[i] / samples()) / samples();
    return o;
  }
  std::vector<double> activation() const {
    const std::vector<double> v = var_mu();
    std::vector<double> m(m1.size());
    double M = -INFINITY;
    for (size_t i = 0; i < m.size(); ++i) {
      m[i] = m1abs[i] / (samples() * std::sqrt(v[i]));
      M = std::max(M, m[i]);
    }
    for (double& x : m) x /= M;
    return m;
  }
};

struct FitOptions {
  int precision = DR_FP64;
  int state_mode = DR_STATE_COMPAT;
  int covfun = DR_COV_RBF;
  int embed[3] = {0, 0, 0};
  int device = 0;
  uint64_t seed = 1;
  bool monitor = false;
  bool output_samples = false;
  bool quiet = false;  // no progress messages (served fits)
  // Gibbs update of tau^-2 after the sigma draws (MultiResParameters.h:723-725),
  // which both reference binaries compile out (DUALRES_DONT_UPDATE_TAU)
  bool update_tau = false;
  // set by another thread to end the fit: checked before the kriging build
  // and before every HMC iteration, which then throws Cancelled
  const std::atomic<bool>* cancel = nullptr;
  // HMC iterations replayed from captured graphs (dr_update_captured; the
  // results are the same bit for bit)
  bool captured = false;
};

struct Cancelled : std::runtime_error {
  Cancelled() : std::runtime_error("fit cancelled") {}
};

inline void check_cancel(const FitOptions& opt) {
  if (opt.cancel && opt.cancel->load(std::memory_order_relaxed)) throw Cancelled();
}

struct MultiResData {  // MultiResData.h:21-127
  int n_datasets = 1;
  std::vector<double> Yh, Ys;
  Csr W;
  std::vector<double> theta;
};

// The post-setup loop of fit_model (gaussian_process_model.h:134-171) on a
// prepared chain (data, mu with data added, output indices, initial sigma
// draws in sig): HMC update, Gibbs draws (and tau with opt.update_tau), dual
// averaging and saves.  iter_s, when given, receives the wall seconds of
// every post-burn-in iteration and iter_L its leapfrog steps (the first
// post-burn-in iteration runs one step when burn-in ended before the
// step-size search did, as in the reference); stop_after > 0 ends the loop
// after that many post-burn-in iterations.
template <typename T>
McmcSummary<T> sample_chain(Chain& chain, long n_h, long n_s, HMCParameters<T>& hmc, std::vector<double> sig,
                            std::ofstream& samples_out, const FitOptions& opt, std::ostream& out,
                            std::vector<double>* iter_s, long stop_after = -1, std::vector<int>* iter_L = nullptr) {
  const int nd = chain.n_datasets();
  McmcSummary<T> summ(chain.n_out(), nd);
  int save_count = 0;
  out << "\nFitting model with rmHMC";
  if (nd > 1) out << " and projected neighborhood process approximation";
  out << std::endl;
  auto start = std::chrono::high_resolution_clock::now();
  std::uniform_real_distribution<double> U01(0, 1);
  std::vector<T> sigma(nd);
  while (save_count < hmc.n_save() && hmc.iteration() <= hmc.max_iterations()) {
    check_cancel(opt);
    const bool post = hmc.iteration() >= hmc.burnin();
    const auto t_iter = std::chrono::steady_clock::now();
    const T eps = hmc.eps();
    const int L = hmc.integrator_steps();
    const double u = U01(rng());
    const dr_update_result r = opt.captured ? chain.update_captured(eps, L, u) : chain.update(eps, L, u);
    if (opt.monitor)
      out << std::setprecision(4) << "Energy Transition (" << r.H0 << " -> " << r.H1 << ");  step size ("
                << eps << ");  MH = (" << r.R << ")" << std::endl;
    // _update_sigma_sq_inv (MultiResParameters.h:820-859)
    {
      std::gamma_distribution<T> G((T)(0.5 * n_h + 1), (T)(1 / (T)(0.5 * r.sse_h)));
      sig[0] = G(rng());
      if (nd == 2) {
        std::gamma_distribution<T> G2((T)(0.5 * n_s + 1), (T)(1 / (T)(0.5 * r.sse_s)));
        sig[1] = G2(rng());
      }
      chain.set_sigma_sq_inv(sig);
    }
    // _update_tau_sq_inv (MultiResParameters.h:863-872)
    if (opt.update_tau) {
      const T shape = (T)(0.5 * n_h + 1);
      const T rate = (T)(0.5 * chain.mu_data_sumsq());
      std::gamma_distribution<T> G(shape, (T)(1 / rate));
      chain.set_tau_sq_inv(G(rng()));
    }
    hmc.update((T)r.mh_rate);
    if (hmc.save_iteration()) {
      if (save_count == 0) start = std::chrono::high_resolution_clock::now();
      for (int i = 0; i < nd; ++i) sigma[i] = (T)std::sqrt(1.0 / sig[i]);
      double ll = sig[0] * r.sse_h;
      if (nd == 2) ll += sig[1] * r.sse_s;
      const T log_lik = (T)(-0.5 * ll);
      summ.update(chain, sigma, log_lik);
      if (samples_out.is_open() && opt.output_samples) {
        const std::vector<double> mu = chain.mu_();
        for (double v : mu) samples_out << (T)v << "\t";
        for (int i = 0; i < nd; ++i) samples_out << sigma[i] << "\t";
        samples_out << (T)std::sqrt(1.0 / chain.tau_sq_inv()) << "\t" << log_lik;
        if (save_count < hmc.n_save() - 1) samples_out << std::endl;
      }
      ++save_count;
    }
    if (post && iter_s)
      iter_s->push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count());
    if (post && iter_L) iter_L->push_back(L);
    if (post && stop_after > 0 && --stop_after == 0) break;
  }
  const auto stop = std::chrono::high_resolution_clock::now();
  summ.finish(chain);
  summ.sampling_time = (T)(std::chrono::duration_cast<std::chrono::milliseconds>(stop - start).count() / 1e3);
  T rate = hmc.metropolis_hastings_rate();
  if (std::isnan(rate)) rate = 0;
  if (rate >= 0 && rate <= 1) summ.mh_rate = rate;
  return summ;
}

// fit_model (gaussian_process_model.h:87-177)
template <typename T>
McmcSummary<T> fit_model(const Image& hr, const Image& input_mask, const Image& output_mask, const MultiResData& data,
                         HMCParameters<T>& hmc, std::ofstream& samples_out, const FitOptions& opt) {
  if (!same_orientation(hr, input_mask) || !same_orientation(hr, output_mask))
    throw std::domain_error("fit_model: image/mask oreintation error");
  NullStream nullout;
  std::ostream& out = opt.quiet ? static_cast<std::ostream&>(nullout) : std::cout;
  const BBox bb = bounding_box(input_mask);
  dr_config cfg{};
  cfg.n_datasets = data.n_datasets;
  cfg.precision = opt.precision;
  cfg.state_mode = opt.state_mode;
  cfg.covfun = opt.covfun;
  for (int a = 0; a < 3; ++a) {
    cfg.embed[a] = opt.embed[a];
    cfg.theta[a] = data.theta[a];
  }
  for (int i = 0; i < 16; ++i) cfg.qform[i] = hr.qto_xyz[i];
  cfg.seed = opt.seed;
  cfg.device = opt.device;
  out << "Selecting DFT algorithms... " << std::flush;
  Chain chain(cfg, nonzero_indices_bounded_by_box(hr, bb));
  out << "Done!" << std::endl;
  int dims[3];
  int64_t N, nonpos;
  double maxeig;
  chain.spectrum(dims, &N, &nonpos, &maxeig);
  out << "Computing eigen values... Done!\n\t<Maximal eigenvalue = " << maxeig << ">\n"
            << std::setprecision(2) << std::fixed << (1.0 - (double)nonpos / N) * 100
            << "% of eigen values have real-part > 0  (" << (N - nonpos) << "/" << N << ")" << std::endl;
  out.unsetf(std::ios::fixed);
  out << std::setprecision(6);
  if (nonpos > 0) out << "\t<Covariance function is numerically not positive definite>" << std::endl;
  chain.set_data(data.Yh, &data.Ys, &data.W);
  if (data.n_datasets == 2) {
    // which device layout W took: the interior rows as one stencil with packed
    // boundary rows, or plain CSR when the stencil layout does not apply
    int64_t st[11];
    check(dr_w_stats(chain.handle(), st));
    if (st[0])
      out << "W layout: stencil for " << st[1] << " interior rows (" << st[2] << " entries), packed entries for "
          << st[3] << " boundary rows" << std::endl;
    else
      out << "W layout: plain CSR (" << st[5] << " nonzeros; the stencil layout does not apply to this W)"
          << std::endl;
  }
  // _initialize_sigma (MultiResParameters.h:764-778), after the mu draw
  std::vector<double> sig(data.n_datasets);
  {
    std::gamma_distribution<T> G((T)data.theta[0], (T)1);
    std::uniform_real_distribution<T> U(0, 1);
    sig[0] = G(rng());
    for (int i = 1; i < data.n_datasets; ++i) sig[i] = (T)sig[0] / U(rng());
  }
  chain.set_sigma_sq_inv(sig);
  chain.add_data_to_mu();
  chain.set_output_indices(nonzero_indices_bounded_by_box(output_mask, bb));
  return sample_chain<T>(chain, (long)data.Yh.size(), (long)data.Ys.size(), hmc, sig, samples_out, opt, out,
                         nullptr);
}

}  // namespace drhost
