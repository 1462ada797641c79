// Fit server (see server.hpp) and its C entry points.
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "../../include/dualres_host.h"
#include "engine.hpp"
#include "nifti.hpp"
#include "sampler.hpp"

namespace drhost {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Image make_image(const std::vector<double>& v, const int d[3], const std::array<float, 16>& q) {
  if ((long)v.size() != (long)d[0] * d[1] * d[2]) throw std::domain_error("request volume size does not match its dims");
  Image im;
  im.nx = d[0];
  im.ny = d[1];
  im.nz = d[2];
  im.data = v;
  im.qto_xyz = q;
  im.datatype = 64;
  return im;
}

double default_neighborhood(const ServeRequest& r) {
  if (r.neighborhood > 0) return r.neighborhood;
  const auto vd = voxel_dimensions(r.qh);
  return 3.0 * std::max({(double)vd[0], (double)vd[1], (double)vd[2]});
}

dr_config make_config(const ServeRequest& r, int device) {
  dr_config c{};
  c.n_datasets = r.n_datasets;
  c.precision = r.precision;
  c.state_mode = r.state_mode;
  c.covfun = DR_COV_RBF;
  for (int a = 0; a < 3; ++a) {
    c.embed[a] = r.embed[a];
    c.theta[a] = r.theta[a];
  }
  for (int i = 0; i < 16; ++i) c.qform[i] = r.qh[i];
  c.seed = r.seed;
  c.device = device;
  return c;
}

template <typename T>
ServeResult fit_as(const ServeRequest& r, int device, const std::atomic<bool>* cancel) {
  ServeResult out;
  FitOptions fo;
  fo.cancel = cancel;
  check_cancel(fo);
  rng().seed((std::mt19937::result_type)r.seed);
  const Image hr = make_image(r.hr, r.hd, r.qh);
  MultiResData data;
  data.n_datasets = r.n_datasets;
  data.theta.assign(r.theta, r.theta + 3);
  data.Yh = nonzero_data(hr);
  if (r.n_datasets == 2) {
    const Image sr = make_image(r.sr, r.sd, r.qs);
    data.W = kriging_build(hr.data, r.hd, r.qh, sr.data, r.sd, r.qs, r.theta, default_neighborhood(r), device);
    data.Ys = nonzero_data(sr);
  }
  HMCParameters<T> hmc(r.burnin, r.nsave, r.thin, r.L, (T)(1.0 / r.L), (T)r.mh_target);
  check_cancel(fo);
  fo.precision = r.precision;
  fo.state_mode = r.state_mode;
  fo.covfun = DR_COV_RBF;
  for (int a = 0; a < 3; ++a) fo.embed[a] = r.embed[a];
  fo.device = device;
  fo.seed = r.seed;
  fo.quiet = true;
  std::ofstream none;
  McmcSummary<T> s = fit_model(hr, hr, hr, data, hmc, none, fo);
  out.mean = s.mode_mu();
  out.var = s.var_mu();
  out.mh_rate = s.mh_rate;
  out.eps = hmc.eps_value();
  out.grad_evals = (int64_t)r.thin * r.nsave * (r.L + 1);
  return out;
}

}  // namespace

int64_t estimate_request_bytes(const ServeRequest& r) {
  const Image hr = make_image(r.hr, r.hd, r.qh);
  const auto ijk = nonzero_indices_bounded_by_box(hr, bounding_box(hr));
  std::vector<int32_t> flat;
  flat.reserve(3 * ijk.size());
  for (const auto& v : ijk) flat.insert(flat.end(), {v[0], v[1], v[2]});
  int64_t n_s = 0, P = 0;
  if (r.n_datasets == 2) {
    for (double v : r.sr) n_s += present(v) ? 1 : 0;
    const auto vd = voxel_dimensions(r.qh);
    const double rad = default_neighborhood(r);
    P = 1;
    for (int a = 0; a < 3; ++a) P *= 2 * (int64_t)std::floor(rad / std::max(1e-6f, vd[a])) + 1;
  }
  const dr_config cfg = make_config(r, 0);
  int64_t b = 0;
  check(dr_estimate_bytes(&cfg, flat.data(), (int64_t)ijk.size(), n_s, n_s * P, &b));
  if (r.n_datasets == 2) {  // kriging build: patterns, systems (batched to <= 6e9 bytes), CSR
    const double sys = std::min(6.5e9, 1.1 * (double)n_s * (double)P * (double)P * 8.0);
    b += (int64_t)sys + n_s * P * 24;
  }
  return b;
}

ServeResult run_fit(const ServeRequest& r, int device, const std::atomic<bool>* cancel) {
  ServeResult out;
  try {
    out = r.precision == DR_FP64 ? fit_as<double>(r, device, cancel) : fit_as<float>(r, device, cancel);
    out.status = DR_OK;
  } catch (const Cancelled& e) {
    out.status = DR_ECANCELED;
    out.error = e.what();
  } catch (const std::domain_error& e) {
    out.status = DR_EDOMAIN;
    out.error = e.what();
  } catch (const std::logic_error& e) {
    out.status = DR_ELOGIC;
    out.error = e.what();
  } catch (const std::exception& e) {
    out.status = DR_ERUNTIME;
    out.error = e.what();
  }
  return out;
}

Server::Server(int64_t budget_bytes, int workers, int device, bool backfill)
    : budget_(budget_bytes), device_(device), backfill_(backfill) {
  if (workers <= 0) {
    uncapped_ = true;
    threads_.emplace_back([this] { dispatcher(); });
  } else {
    for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { worker(); });
  }
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : threads_) t.join();
  for (auto& f : fits_) f.first.join();
}

int64_t Server::submit(ServeRequest r) {
  const int64_t bytes = estimate_request_bytes(r);
  std::lock_guard<std::mutex> g(mu_);
  const int64_t id = next_id_++;
  if (bytes > budget_) {  // can never be admitted
    ServeResult res;
    res.status = DR_ENOMEM;
    res.error = "request needs " + std::to_string(bytes) + " device bytes, over the server budget";
    res.device_bytes = bytes;
    done_[id] = std::move(res);
    cv_done_.notify_all();
    return id;
  }
  queue_.push_back(Job{id, std::move(r), bytes, now_s(), std::make_shared<std::atomic<bool>>(false)});
  cv_work_.notify_all();
  return id;
}

bool Server::take(Job* job) {
  auto it = queue_.end();
  for (auto q = queue_.begin(); q != queue_.end(); ++q) {
    if (in_use_ + q->bytes <= budget_) {
      it = q;
      break;
    }
    if (!backfill_) break;  // strict FIFO: wait for the head
  }
  if (it == queue_.end()) return false;
  *job = std::move(*it);
  queue_.erase(it);
  in_use_ += job->bytes;
  ++running_;
  running_flags_[job->id] = job->cancel;
  return true;
}

int Server::cancel(int64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto q = queue_.begin(); q != queue_.end(); ++q) {
    if (q->id != id) continue;
    ServeResult res;
    res.status = DR_ECANCELED;
    res.error = "request cancelled before it started";
    res.device_bytes = q->bytes;
    res.queue_s = res.total_s = now_s() - q->t_submit;
    queue_.erase(q);
    done_[id] = std::move(res);
    cv_done_.notify_all();
    cv_work_.notify_all();  // a smaller request behind it may now be admitted
    return 0;
  }
  const auto f = running_flags_.find(id);
  if (f == running_flags_.end()) return 2;
  f->second->store(true);
  return 1;
}

void Server::run_job(Job job) {
  const double t0 = now_s();
  ServeResult res = run_fit(job.req, device_, job.cancel.get());
  const double t1 = now_s();
  res.queue_s = t0 - job.t_submit;
  res.run_s = t1 - t0;
  res.total_s = t1 - job.t_submit;
  res.device_bytes = job.bytes;
  {
    std::lock_guard<std::mutex> g(mu_);
    in_use_ -= job.bytes;
    --running_;
    running_flags_.erase(job.id);
    done_[job.id] = std::move(res);
  }
  cv_done_.notify_all();
  cv_work_.notify_all();
}

void Server::worker() {
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (stop_ && queue_.empty()) return;
        if (take(&job)) break;
        cv_work_.wait(lk);
      }
    }
    run_job(std::move(job));
  }
}

// uncapped server: admit every job the budget allows, each on a new thread;
// threads of finished fits are joined as the dispatcher goes
void Server::dispatcher() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    for (auto it = fits_.begin(); it != fits_.end();) {
      if (it->second->load()) {
        it->first.join();
        it = fits_.erase(it);
      } else {
        ++it;
      }
    }
    if (stop_ && queue_.empty()) return;
    Job job;
    if (take(&job)) {
      auto flag = std::make_shared<std::atomic<bool>>(false);
      fits_.emplace_back(std::thread([this, flag, j = std::move(job)]() mutable {
                           run_job(std::move(j));
                           flag->store(true);
                         }),
                         flag);
      continue;
    }
    cv_work_.wait(lk);
  }
}

bool Server::wait(int64_t id, ServeResult* out, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
  while (done_.find(id) == done_.end()) {
    if (cv_done_.wait_until(lk, until) == std::cv_status::timeout && done_.find(id) == done_.end()) return false;
  }
  *out = std::move(done_[id]);
  done_.erase(id);
  return true;
}

int64_t Server::bytes_in_use() {
  std::lock_guard<std::mutex> g(mu_);
  return in_use_;
}

int Server::running() {
  std::lock_guard<std::mutex> g(mu_);
  return running_;
}

}  // namespace drhost

// ---------------------------------------------------------------- C entry points
extern "C" {

static drhost::ServeRequest to_request(const drh_request* q) {
  drhost::ServeRequest r;
  r.n_datasets = q->n_datasets;
  const long nh = (long)q->hr_dims[0] * q->hr_dims[1] * q->hr_dims[2];
  r.hr.assign(q->hr, q->hr + nh);
  for (int a = 0; a < 3; ++a) {
    r.hd[a] = q->hr_dims[a];
    r.theta[a] = q->theta[a];
    r.embed[a] = q->embed[a];
  }
  for (int i = 0; i < 16; ++i) r.qh[i] = q->q_hr[i];
  if (q->n_datasets == 2) {
    const long ns = (long)q->sr_dims[0] * q->sr_dims[1] * q->sr_dims[2];
    r.sr.assign(q->sr, q->sr + ns);
    for (int a = 0; a < 3; ++a) r.sd[a] = q->sr_dims[a];
    for (int i = 0; i < 16; ++i) r.qs[i] = q->q_sr[i];
  }
  r.neighborhood = q->neighborhood;
  r.burnin = q->burnin;
  r.nsave = q->nsave;
  r.thin = q->thin;
  r.L = q->L;
  r.mh_target = q->mh_target;
  r.seed = q->seed;
  r.precision = q->precision;
  r.state_mode = q->state_mode;
  return r;
}

int drh_server_create(int64_t budget_bytes, int workers, int device, int backfill, void** out) {
  try {
    *out = new drhost::Server(budget_bytes, workers, device, backfill != 0);
    return 0;
  } catch (...) {
    return 1;
  }
}

int drh_server_submit(void* s, const drh_request* q, int64_t* id) {
  try {
    *id = static_cast<drhost::Server*>(s)->submit(to_request(q));
    return 0;
  } catch (...) {
    return 1;
  }
}

// 0: done (result filled, min(cap, n_out) values copied); 1: timeout
int drh_server_wait(void* s, int64_t id, double timeout_s, drh_result* res, double* mean, double* var, int64_t cap,
                    char* err, int errlen) {
  drhost::ServeResult r;
  if (!static_cast<drhost::Server*>(s)->wait(id, &r, timeout_s)) return 1;
  res->status = r.status;
  res->mh_rate = r.mh_rate;
  res->eps = r.eps;
  res->queue_s = r.queue_s;
  res->run_s = r.run_s;
  res->total_s = r.total_s;
  res->grad_evals = r.grad_evals;
  res->device_bytes = r.device_bytes;
  res->n_out = (int64_t)r.mean.size();
  const int64_t n = std::min<int64_t>(cap, (int64_t)r.mean.size());
  if (mean && n > 0) std::memcpy(mean, r.mean.data(), n * sizeof(double));
  if (var && n > 0) std::memcpy(var, r.var.data(), n * sizeof(double));
  if (err && errlen > 0) {
    std::strncpy(err, r.error.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
  return 0;
}

// 0: dropped from the queue; 1: running fit told to stop; 2: unknown or done
int drh_server_cancel(void* s, int64_t id) { return static_cast<drhost::Server*>(s)->cancel(id); }

int drh_server_destroy(void* s) {
  delete static_cast<drhost::Server*>(s);
  return 0;
}

int drh_estimate_request_bytes(const drh_request* q, int64_t* bytes) {
  try {
    *bytes = drhost::estimate_request_bytes(to_request(q));
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"
