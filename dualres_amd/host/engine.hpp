// C++ view of the C ABI (include/dualres_hip.h): status codes rethrown as the
// std:: exception types the reference throws, RAII handles.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dualres_hip.h"

namespace drhost {

inline void check(int st) {
  if (st == DR_OK) return;
  const std::string m = dr_last_error();
  switch (st) {
    case DR_EDOMAIN: throw std::domain_error(m);
    case DR_ELOGIC: throw std::logic_error(m);
    default: throw std::runtime_error(m);
  }
}

// W as row-major CSR with columns indexing the highres nonzero voxels
struct Csr {
  long nrow = 0, ncol = 0, n_unique = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> cols;
  std::vector<double> vals;
};

inline Csr kriging_build(const std::vector<double>& hr, const int hd[3], const std::array<float, 16>& qh,
                         const std::vector<double>& st, const int sd[3], const std::array<float, 16>& qs,
                         const double theta[3], double radius, int device) {
  dr_csr* h = nullptr;
  check(dr_kriging_build(hr.data(), hd, qh.data(), st.data(), sd, qs.data(), theta, radius, device, &h));
  Csr c;
  int64_t nr, nc, nnz, nu;
  const int s = dr_csr_info(h, &nr, &nc, &nnz, &nu);
  if (s != DR_OK) {
    dr_csr_free(h);
    check(s);
  }
  c.nrow = nr;
  c.ncol = nc;
  c.n_unique = nu;
  c.rp.resize(nr + 1);
  c.cols.resize(nnz);
  c.vals.resize(nnz);
  const int s2 = dr_csr_copy(h, c.rp.data(), c.cols.data(), c.vals.data());
  dr_csr_free(h);
  check(s2);
  return c;
}

class Chain {
 public:
  Chain(const dr_config& cfg, const std::vector<std::array<int, 3>>& ijk) : n_h_((long)ijk.size()) {
    std::vector<int32_t> flat;
    flat.reserve(3 * ijk.size());
    for (const auto& v : ijk) flat.insert(flat.end(), {v[0], v[1], v[2]});
    check(dr_create(&cfg, flat.data(), n_h_, &h_));
    nd_ = cfg.n_datasets;
  }
  // a handle created elsewhere (the caller keeps ownership)
  Chain(dr_chain* borrowed, int n_datasets, long n_h, long n_out)
      : h_(borrowed), n_h_(n_h), n_out_(n_out), nd_(n_datasets), owned_(false) {}
  ~Chain() {
    if (h_ && owned_) dr_destroy(h_);
  }
  Chain(const Chain&) = delete;
  Chain& operator=(const Chain&) = delete;
  dr_chain* handle() { return h_; }
  int n_datasets() const { return nd_; }
  long n_h() const { return n_h_; }
  long n_out() const { return n_out_ < 0 ? n_h_ : n_out_; }

  void spectrum(int dims[3], int64_t* N, int64_t* nonpos, double* max_eig) {
    check(dr_spectrum_report(h_, dims, N, nonpos, max_eig));
  }
  void set_data(const std::vector<double>& Yh, const std::vector<double>* Ys, const Csr* W) {
    if (nd_ == 2) {
      check(dr_set_data(h_, Yh.data(), (int64_t)Yh.size(), Ys->data(), (int64_t)Ys->size(), W->rp.data(),
                        W->cols.data(), W->vals.data()));
    } else {
      check(dr_set_data(h_, Yh.data(), (int64_t)Yh.size(), nullptr, 0, nullptr, nullptr, nullptr));
    }
  }
  void add_data_to_mu() { check(dr_add_data_to_mu(h_)); }
  void set_output_indices(const std::vector<std::array<int, 3>>& ijk) {
    std::vector<int32_t> flat;
    for (const auto& v : ijk) flat.insert(flat.end(), {v[0], v[1], v[2]});
    check(dr_set_output_indices(h_, flat.data(), (int64_t)ijk.size()));
    n_out_ = (long)ijk.size();
  }
  void set_sigma_sq_inv(const std::vector<double>& s) { check(dr_set_sigma_sq_inv(h_, s.data())); }
  void set_tau_sq_inv(double t) { check(dr_set_tau_sq_inv(h_, t)); }
  double tau_sq_inv() {
    double t = 0;
    check(dr_get_tau_sq_inv(h_, &t));
    return t;
  }
  double mu_data_sumsq() {
    double s = 0;
    check(dr_mu_data_sumsq(h_, &s));
    return s;
  }
  dr_update_result update(double eps, int L, double u) {
    dr_update_result r{};
    check(dr_update(h_, eps, L, u, &r));
    return r;
  }
  // the same iteration replayed from a graph captured once per L
  dr_update_result update_captured(double eps, int L, double u) {
    dr_update_result r{};
    check(dr_update_captured(h_, eps, L, u, &r));
    return r;
  }
  std::vector<double> mu_() {
    std::vector<double> out(n_out());
    check(dr_get_mu_output(h_, out.data(), (int64_t)out.size()));
    return out;
  }

 private:
  dr_chain* h_ = nullptr;
  long n_h_ = 0, n_out_ = -1;
  int nd_ = 1;
  bool owned_ = true;
};

}  // namespace drhost
