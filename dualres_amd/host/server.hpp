// Fit server: many independent posterior fits share one device.
//
// Requests wait in a FIFO queue.  A request is admitted when the device
// bytes it will need (dr_estimate_bytes for its chain, plus the kriging
// build's scratch for a dual-resolution fit) fit under the budget next to
// the fits already running; with backfill on, a later request that fits may
// start while a larger one at the head waits.  Admitted fits run to
// completion on their own threads, each with its own chain (own HIP stream)
// and its own host generator seeded from the request, so a fit's result does
// not depend on what else runs beside it.  With workers > 0 at most that many
// fits run at once; with workers <= 0 the device-byte budget alone limits
// how many run together (a thread is started per admitted fit).
#pragma once

#include <array>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace drhost {

struct ServeRequest {
  int n_datasets = 1;
  std::vector<double> hr;  // highres volume, x fastest; 0 / NaN = missing
  int hd[3] = {0, 0, 0};
  std::array<float, 16> qh{};
  std::vector<double> sr;  // stdres volume (dual-resolution fits)
  int sd[3] = {0, 0, 0};
  std::array<float, 16> qs{};
  double theta[3] = {1, 0.6, 1.5};
  double neighborhood = -1;  // <= 0: 3 x the largest highres voxel side
  int burnin = 1000, nsave = 1000, thin = 3, L = 25;
  double mh_target = 0.65;
  uint64_t seed = 1;
  int precision = 1;   // dr_precision
  int state_mode = 0;  // dr_state_mode
  int embed[3] = {0, 0, 0};
};

struct ServeResult {
  int status = 0;  // dr_status
  std::string error;
  std::vector<double> mean, var;  // posterior moments at the highres data voxels
  double mh_rate = 0, eps = 0;
  double queue_s = 0, run_s = 0, total_s = 0;  // submit -> start -> done
  int64_t grad_evals = 0, device_bytes = 0;
};

class Server {
 public:
  Server(int64_t budget_bytes, int workers, int device, bool backfill);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;
  int64_t submit(ServeRequest r);  // returns the request id
  // false on timeout; the result is removed from the server when returned
  bool wait(int64_t id, ServeResult* out, double timeout_s);
  // 0: the request was still queued and is dropped (its result, status
  // DR_ECANCELED, is ready); 1: its fit is running and is told to stop at its
  // next HMC iteration (the result then has DR_ECANCELED unless the fit ended
  // first); 2: no such request, or its result is already there
  int cancel(int64_t id);
  int64_t bytes_in_use();
  int running();

 private:
  struct Job {
    int64_t id;
    ServeRequest req;
    int64_t bytes;
    double t_submit;
    std::shared_ptr<std::atomic<bool>> cancel;
  };
  void worker();
  void dispatcher();
  bool take(Job* job);  // under mu_: pop the first admissible job
  void run_job(Job job);
  int64_t budget_;
  bool uncapped_ = false;
  // per-fit threads of an uncapped server, with a flag set when each ends
  std::deque<std::pair<std::thread, std::shared_ptr<std::atomic<bool>>>> fits_;
  int device_;
  bool backfill_;
  bool stop_ = false;
  int64_t next_id_ = 1, in_use_ = 0;
  int running_ = 0;
  std::deque<Job> queue_;
  std::map<int64_t, ServeResult> done_;
  std::map<int64_t, std::shared_ptr<std::atomic<bool>>> running_flags_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::thread> threads_;
};

// device bytes a request will hold at its peak (chain + W + kriging scratch)
int64_t estimate_request_bytes(const ServeRequest& r);
// runs one fit on the calling thread (what a worker does); cancel, when
// given, ends it early with status DR_ECANCELED once set
ServeResult run_fit(const ServeRequest& r, int device, const std::atomic<bool>* cancel = nullptr);

}  // namespace drhost
