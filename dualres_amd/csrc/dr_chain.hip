// HMC chain on one GPU: Fourier-domain real state, pruned 3D FFT passes,
// fused leapfrog updates, CSR kriging products, float64 energies.
//
// State layout in HBM (working precision T):
//   muA, muB  : half spectra of mu (current, proposal)  [M2][M1][M0/2+1]
//   p         : half spectrum of the momentum
//   miA, miB, pim : imaginary chain (DR_STATE_COMPAT only)
//   invlam    : 1/lambda on the half spectrum, 0 on Z
//   T1, T2    : pass scratch (after z / after z,y of the inverse)
//   RA, RB    : mu on the image region in real space, [n2][n1][n0]
//   Yg, mask, WTv : data on the image region
//   W (CSR, columns = image-region grid index), WT (transposed CSR)
// Reference: include/dualres/MultiResParameters.h (whole file).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include <unistd.h>

#include "dr_kernels.hpp"

namespace drh {

// ---------------------------------------------------------------- small utils
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

std::vector<int> factorize(int M) {
  std::vector<int> f;
  int m = M;
  while (m % 16 == 0) { f.push_back(16); m /= 16; }
  while (m % 8 == 0 && m > 8) { f.push_back(8); m /= 8; }
  if (m % 8 == 0) { f.push_back(8); m /= 8; }
  while (m % 4 == 0) { f.push_back(4); m /= 4; }
  while (m % 2 == 0) { f.push_back(2); m /= 2; }
  for (int r : {3, 5, 7}) while (m % r == 0) { f.push_back(r); m /= r; }
  if (m != 1) throw Error(DR_EDOMAIN, "FFT length " + std::to_string(M) + " has a prime factor > 7");
  return f;
}

// Transform lengths whose per-step passes use the compile-time stage
// pipeline (fft_ct); other lengths take the runtime-radix path.
template <class F>
void with_fft_len(int M, F&& f) {
  switch (M) {
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    case 160: f(std::integral_constant<int, 160>{}); break;
    case 256: f(std::integral_constant<int, 256>{}); break;
    case 320: f(std::integral_constant<int, 320>{}); break;
    case 512: f(std::integral_constant<int, 512>{}); break;
    default: f(std::integral_constant<int, 0>{}); break;
  }
}

int pick_tbs(int len, int rows_lds, int elem_bytes, bool padded, int max_lines, int tile_elems) {
  int tbs = 0;
  auto lds = [&](int tb) { return (long)rows_lds * (padded ? tb + 1 : tb) * elem_bytes; };
  while (true) {
    const int tb = 2 << tbs;
    if ((long)tb * len > tile_elems) break;
    if (lds(tb) > 63488) break;
    if ((1 << tbs) >= max_lines) break;
    ++tbs;
  }
  return tbs;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(DR_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  void zero(hipStream_t s) {
    if (p) DR_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
};

__global__ void __launch_bounds__(kThreads) k_finalize(FinalizeArgs a, const double* part, double* sums) {
  __shared__ double red[kThreads];
  const int s = blockIdx.x;
  if (s >= a.nslot) return;
  const int nb = a.nblk[s];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) acc += part[s * a.pstride + i];
  const double r = block_sum(acc, red);
  if (threadIdx.x == 0) sums[s] = nb > 0 ? r : 0.0;
}


// ---------------------------------------------------------------- reduction slots
enum Slot {
  S_E0 = 0, S_E0I, S_PRIOR_A, S_PRIORI_A, S_SSEH_A, S_SSES_A,
  S_PRIOR_B, S_PRIORI_B, S_KIN_B, S_KINI_B, S_SSEH_B, S_SSES_B,
  S_KIN_A, S_KINI_A, S_ZCOUNT, S_TMP, NSLOT
};

// ---------------------------------------------------------------- chain
struct ChainBase {
  virtual ~ChainBase() = default;
  dr_config cfg{};
  int M[3] = {0, 0, 0};
  int H0 = 0;
  long N = 0, Nh = 0;
  int reg[3] = {0, 0, 0};
  long nreg = 0;
  long n_h = 0, n_s = 0;
  std::vector<int32_t> ijk;         // data indices (n_h x 3)
  std::vector<int32_t> out_ijk;     // output indices (n_out x 3)
  std::vector<double> Yh, Ys;       // data in nonzero order
  std::vector<int64_t> W_rp;        // CSR as given (columns index Yh)
  std::vector<int32_t> W_cols;
  std::vector<double> W_vals;
  double sigma[2] = {1.0, 1.0};
  uint64_t iteration = 0;
  long n_nonpos = 0;
  double max_eig = 0.0;
  bool data_set = false;
  bool use_cls = false;    // class-stencil W kernels active
  bool force_csr = false;  // keep the plain CSR kernels (tests)
  int64_t wstats[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // cached energies of the current state A
  bool stats_valid = false;
  double prior_sum = 0.0, sse_h = 0.0, sse_s = 0.0;
  // prior precision tau^-2 (1 / theta[0] until the caller updates it)
  double tau_sq_inv = 0.0;
  // profiling
  bool prof = false;
  double prof_ms = 0.0, prof_fft_ms = 0.0, prof_w_ms = 0.0;
  long prof_n = 0;
  hipStream_t stream = nullptr;

  virtual void build() = 0;
  virtual void upload_data() = 0;
  virtual void add_data_to_mu() = 0;
  virtual void set_region(const int r[3]) = 0;
  virtual void get_mu_region_gather(const std::vector<int32_t>& gidx, double* out) = 0;
  virtual void get_mu_out(double* out, long n) = 0;
  virtual void set_outputs() = 0;
  virtual void summary_reset() = 0;
  virtual void summary_update() = 0;
  virtual void summary_read(double* s1, double* s2, double* sa, long n, int64_t* count) = 0;
  long n_out_dev = 0;  // output voxels held on the device
  virtual void update(double eps, int L, double u, const double* zr, const double* zi, dr_update_result* res) = 0;
  virtual void leapfrog(double eps, int L, dr_update_result* res) = 0;
  virtual void set_state(const double* mr, const double* mi, const double* pr, const double* pi) = 0;
  virtual void get_state(double* mr, double* mi, double* pr, double* pi) = 0;
  virtual void gradient(double* gr, double* gi) = 0;
  virtual void refresh_stats() = 0;
  virtual void set_tau_sq_inv(double t) = 0;
  virtual double mu_data_sumsq() = 0;
  virtual void get_lambda(double* out) = 0;
  virtual void reset_kernel_profile() = 0;
  virtual void kernel_profile(double* ms, int max_k, int* n_k, int64_t* n_steps) = 0;
  // captured iterations (one HIP graph per leapfrog length; see Chain)
  uint64_t layout_gen = 0;  // changes whenever the buffers a captured graph uses move
  virtual void update_graph(double eps, int L, double u, dr_update_result* res) = 0;
  virtual void set_batch_slots(void* dev_scalars, void* host_scalars, double* dev_sums) = 0;
  virtual void stage_scalars(double eps) = 0;
  virtual hipGraph_t capture_iteration(int L) = 0;
  virtual void enqueue_iteration(double eps, int L, hipStream_t s) = 0;
  virtual void enqueue_eager(double eps, int L) = 0;
  // autotuned pass shapes: 5 tile-line exponents, 3 radix orders (see Chain)
  virtual void set_tuning(const int* prm) = 0;
  virtual void get_tuning(int* prm) const = 0;
  virtual void autotune(int* prm) = 0;
  virtual void finish_iteration(const double* h, double u, dr_update_result* res, bool copy, hipStream_t s) = 0;

  long region_index(int x, int y, int z) const { return x + (long)reg[0] * (y + (long)reg[1] * z); }
  double log_lik(double sh, double ss) const {
    double ll = sigma[0] * sh;
    if (cfg.n_datasets == 2) ll += sigma[1] * ss;
    return -0.5 * ll;
  }
  bool compat() const { return cfg.state_mode == DR_STATE_COMPAT; }
};

template <typename T>
struct Chain : public ChainBase {
  // FFT tables and pass descriptors (x: length M0/2 into the M0 table)
  DevBuf<cx<float>> twf[3];
  DevBuf<cx<double>> twd[3];
  AxisFFT<float> adf[3]{};
  AxisFFT<double> add[3]{};
  // state
  DevBuf<cx<T>> muA, muB, p, miA, miB, pim, T1, T2;
  DevBuf<T> invlam;
  DevBuf<double> lam;  // Re lambda on the half spectrum (float64)
  DevBuf<T> invlam0;    // 1/lambda and lambda at theta[0], kept once tau^-2 is updated
  DevBuf<double> lam0;
  DevBuf<T> RA, RB, Yg, WTv, WTvb, Rf, zbuf;
  DevBuf<uint8_t> mask;
  DevBuf<int> Wrp, Wcols, WTrp, WTrows, gI, outg;
  DevBuf<T> Wvals, WTvals, Ysd, vsd, outbuf;
  DevBuf<double> part, sums;
  // posterior summary at the output voxels (mcmc_summary), float64
  DevBuf<double> sm1, sm2, sma;
  DevBuf<unsigned long long> smax;
  int64_t nsum = 0;
  // stencil + boundary W (see build_stencil)
  struct WStencil {
    DevBuf<int> irow, ibase, ilidx, sub_start, sub_ldo, tiles, mpar, wstart, woff;
    DevBuf<T> sub_w, Vl, ww;
    DevBuf<int> brow, brp, bbase, tbase, tg, trp, frow, frp, fcol, tfp, tfidx;
    DevBuf<uint16_t> bd, td;
    DevBuf<T> bw, tw, Vc, fw, tfw, Vf;
    int n_int = 0, nD = 0, n_bnd = 0, n_touch = 0, n_far = 0, n_mslot = 0, ncw = 0;
    // interior W^T v tiles: voxel tile dims, LDS box dims, lattice c + sg Z^3,
    // smallest lattice offset, lattice grid box (lo, lv)
    int n_tiles = 0, ncls = 1, tdim[3] = {1, 1, 1}, ldim[3] = {1, 1, 1}, c[3] = {0, 0, 0}, sg[3] = {1, 1, 1},
        omin[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, lv[3] = {1, 1, 1};
  } ws;
  long pstride = 0;
  int slot_nblk[NSLOT] = {0};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // per-kernel profile of one leapfrog step: events between launches
  // Captured iterations.  The kernels of a captured HMC iteration read their
  // per-iteration scalars from dsc (staged in pinned host memory hsc and
  // copied before each launch), and finalize writes the energy sums to
  // sums_out; a batch points both at its own arrays.  gsc is set while
  // capturing, so the functors built then bind to dsc.
  DevBuf<StepScalars<T>> sc_own;
  StepScalars<T>* dsc = nullptr;
  StepScalars<T>* hsc = nullptr;
  bool hsc_own = false;
  const StepScalars<T>* gsc = nullptr;
  double* sums_out = nullptr;
  std::map<int, hipGraphExec_t> gexec;
  static constexpr int kMarks = 12;
  hipEvent_t pev[kMarks] = {};
  int rec = -1, nrec = 0;
  double prof_k[kMarks] = {0};
  void mark() {
    if (rec >= 0 && rec < kMarks) {
      DR_HIP(hipEventRecord(pev[rec], stream));
      ++rec;
    }
  }

  // Autotuned pass shapes (dr_autotune): tile lines (log2) of the z, inverse
  // y, inverse x, forward x and forward y passes of a leapfrog step, -1 for
  // pick_tbs's rule, capped at what the tile and LDS allow; and the radix
  // order of each axis's runtime stages (make_axis).
  enum { P_Z = 0, P_YI, P_XI, P_XF, P_YF, NPASS };
  int tune_tbs[NPASS] = {-1, -1, -1, -1, -1};
  int tune_order[3] = {0, 0, 0};
  int tuned_tbs(int pid, int rule, int cap) const {
    if (pid < 0 || tune_tbs[pid] < 0) return rule;
    return std::min(tune_tbs[pid], cap);
  }

  ~Chain() override {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    if (hsc_own && hsc) (void)hipHostFree(hsc);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : pev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // ------------------------------------------------------------ launch helpers
  template <typename U>
  const AxisFFT<U>& axis_dev(int a) const {
    if constexpr (std::is_same<U, float>::value) return adf[a];
    else return add[a];
  }

  // transform along z (a=2) or y (a=1) over lines with contiguous batch NB
  template <typename U, bool INV, class Pro, class Epi>
  void strided(int a, int n_outer, int NB, int k_in, int k_out, Pro pro, Epi epi, int slot = -1, int pid = -1) {
    const AxisFFT<U> ad = axis_dev<U>(a);
    DR_REQUIRE(k_in >= 1 && k_in <= ad.M && k_out >= 1 && k_out <= ad.M && NB >= 1 && n_outer >= 1,
               DR_ELOGIC, "strided pass: bad extents");
    const int tbs = tuned_tbs(pid, pick_tbs(ad.M, ad.M, sizeof(cx<U>), false, NB, kTileStrided),
                              pick_tbs(ad.M, ad.M, sizeof(cx<U>), false, 1 << 20, kTileStrided));
    const int TB = 1 << tbs;
    dim3 grid(cdiv(NB, TB), n_outer);
    const size_t lds = (size_t)ad.M * TB * sizeof(cx<U>);
    if (slot >= 0) note_blocks(slot, grid.x * grid.y);
    auto go = [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      hipLaunchKernelGGL((k_strided<U, INV, Pro, Epi, kEptStrided, MC>), grid, dim3(kThreads), lds, stream, ad, NB,
                         k_in, k_out, tbs, pro, epi);
    };
    with_fft_len(ad.M, go);
    DR_HIP(hipGetLastError());
    mark();
  }
  template <typename U, class Pro, class Epi>
  void xr2c(int nrows, Pro pro, Epi epi, int slot = -1, int pid = -1) {
    const AxisFFT<U> ad = axis_dev<U>(0);
    const int Mh = ad.M;
    const int tbs = tuned_tbs(pid, pick_tbs(Mh, Mh + 1, sizeof(cx<U>), true, xlines(nrows, Mh), kTileX),
                              pick_tbs(Mh, Mh + 1, sizeof(cx<U>), true, 1 << 20, kTileX));
    const int TB = 1 << tbs;
    const int grid = cdiv(nrows, TB);
    const size_t lds = (size_t)(Mh + 1) * (TB + 1) * sizeof(cx<U>);
    if (slot >= 0) note_blocks(slot, grid);
    auto go = [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      hipLaunchKernelGGL((k_r2c_x<U, Pro, Epi, kEptX, MC>), dim3(grid), dim3(kThreads), lds, stream, ad, nrows, tbs,
                         pro, epi);
    };
    with_fft_len(Mh, go);
    DR_HIP(hipGetLastError());
    mark();
  }
  template <typename U, class Pro, class Epi>
  void xc2r(int nrows, Pro pro, Epi epi, int slot = -1, int pid = -1) {
    const AxisFFT<U> ad = axis_dev<U>(0);
    const int Mh = ad.M;
    // twice the forward pass's rows per tile: at 1 mm (16,704 rows of 128)
    // 16-row tiles ran the inverse pass in 13.8 us against 16.2 us for 8 rows,
    // while the forward pass with its residual prologue was faster at 8
    const int tbs = tuned_tbs(pid, pick_tbs(Mh, Mh + 1, sizeof(cx<U>), true, 2 * xlines(nrows, Mh), kTileX),
                              pick_tbs(Mh, Mh + 1, sizeof(cx<U>), true, 1 << 20, kTileX));
    const int TB = 1 << tbs;
    const int grid = cdiv(nrows, TB);
    const size_t lds = (size_t)(Mh + 1) * (TB + 1) * sizeof(cx<U>);
    if (slot >= 0) note_blocks(slot, grid);
    auto go = [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      hipLaunchKernelGGL((k_c2r_x<U, Pro, Epi, kEptX, MC>), dim3(grid), dim3(kThreads), lds, stream, ad, nrows, tbs,
                         pro, epi);
    };
    with_fft_len(Mh, go);
    DR_HIP(hipGetLastError());
    mark();
  }
  // lines per x-pass workgroup: enough work per group (>= ~1024 elements) but
  // at least ~2048 workgroups when there are enough rows to fill 256 CUs
  static int xlines(int nrows, int Mh) { return std::max(std::max(1, nrows / 2048), cdiv(1024, std::max(Mh, 1))); }
  void note_blocks(int slot, long nb) {
    DR_REQUIRE(nb <= pstride, DR_ELOGIC, "reduction grid exceeds partial buffer");
    slot_nblk[slot] = (int)nb;
  }
  double* pslot(int s) { return part.p + (long)s * pstride; }

  // ------------------------------------------------------------ build
  void build() override {
    DR_HIP(hipSetDevice(cfg.device));
    DR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) DR_HIP(hipEventCreate(&e));
    for (auto& e : pev) DR_HIP(hipEventCreate(&e));
    H0 = M[0] / 2 + 1;
    N = (long)M[0] * M[1] * M[2];
    Nh = (long)H0 * M[1] * M[2];
    tau_sq_inv = 1.0 / cfg.theta[0];  // MultiResParameters.h:256
    for (int a = 0; a < 3; ++a) {
      const int len = M[a];
      const int flen = a == 0 ? len / 2 : len;
      const int tws = a == 0 ? 2 : 1;
      auto td = twiddles<double>(len);
      twd[a].alloc(len);
      DR_HIP(hipMemcpy(twd[a].p, td.data(), len * sizeof(cx<double>), hipMemcpyHostToDevice));
      add[a] = make_axis<double>(flen, twd[a].p, tws);
      if (std::is_same<T, float>::value) {
        auto tf = twiddles<float>(len);
        twf[a].alloc(len);
        DR_HIP(hipMemcpy(twf[a].p, tf.data(), len * sizeof(cx<float>), hipMemcpyHostToDevice));
        adf[a] = make_axis<float>(flen, twf[a].p, tws);
      }
    }
    // partial buffer: the largest grid any reducing kernel uses
    pstride = std::max<long>(1024, Nh / 8 + 1);
    pstride = std::max<long>(pstride, (long)M[1] * M[2] + 1);
    pstride = std::max<long>(pstride, (long)M[1] * H0 + 1);
    part.alloc((size_t)NSLOT * pstride);
    sums.alloc(NSLOT);
    part.zero(stream);
    sums.zero(stream);
    invlam.alloc(Nh);
    lam.alloc(Nh);
    muA.alloc(Nh);
    muB.alloc(Nh);
    p.alloc(Nh);
    p.zero(stream);
    T1.alloc(Nh);
    T2.alloc(Nh);
    if (compat()) {
      miA.alloc(Nh);
      miB.alloc(Nh);
      pim.alloc(Nh);
      pim.zero(stream);
    }
    int r[3];
    for (int a = 0; a < 3; ++a) {
      int m = 0;
      for (long i = 0; i < n_h; ++i) m = std::max(m, ijk[i * 3 + a]);
      r[a] = m + 1;
    }
    reg[0] = reg[1] = reg[2] = 0;
    set_region(r);
    compute_spectrum();
    init_mu();
  }

  // ------------------------------------------------------------ spectrum
  void compute_spectrum() {
    // float64 scratch for the spectrum transform
    DevBuf<cx<double>> s1, s2;
    s1.alloc(Nh);
    s2.alloc(Nh);
    LoadBase lb{};
    lb.M0 = M[0]; lb.M1 = M[1]; lb.M2 = M[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) lb.q[3 * i + j] = cfg.qform[4 * i + j];
    lb.tau2 = cfg.theta[0]; lb.psi = cfg.theta[1]; lb.nu = cfg.theta[2]; lb.covfun = cfg.covfun;
    const int rows = M[1] * M[2];
    xr2c<double>(rows, lb, StoreRowC<double>{{}, s2.p, H0});
    strided<double, false>(1, M[2], H0, M[1], M[1], LoadS<double>{{}, s2.p, (long)M[1] * H0, H0},
                           StoreS<double>{{}, s1.p, (long)M[1] * H0, H0});
    EpiSpectrum<T> es{lam.p, invlam.p, (long)M[1] * H0, H0, pslot(S_ZCOUNT), 0.0};
    strided<double, false>(2, 1, M[1] * H0, M[2], M[2], LoadS<double>{{}, s1.p, 0, (long)M[1] * H0}, es, S_ZCOUNT);
    finalize();
    double h[NSLOT];
    DR_HIP(hipMemcpyAsync(h, sums.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    double l0 = 0.0;
    DR_HIP(hipMemcpyAsync(&l0, lam.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
    n_nonpos = (long)llround(h[S_ZCOUNT]);
    max_eig = l0 / (double)N;
    if (n_nonpos >= N / 2) throw Error(DR_EDOMAIN, "Too many negative eigen values");
  }

  void get_lambda(double* out) override {
    std::vector<double> h(Nh);
    DR_HIP(hipMemcpyAsync(h.data(), lam.p, Nh * sizeof(double), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
    // expand the Hermitian half spectrum (lambda even and real)
    for (int z = 0; z < M[2]; ++z)
      for (int y = 0; y < M[1]; ++y)
        for (int x = 0; x < M[0]; ++x) {
          int kx = x, ky = y, kz = z;
          if (x >= H0) { kx = M[0] - x; ky = (M[1] - y) % M[1]; kz = (M[2] - z) % M[2]; }
          out[x + (long)M[0] * (y + (long)M[1] * z)] = h[kx + (long)H0 * (ky + (long)M[1] * kz)];
        }
  }

  // ------------------------------------------------------------ generic transforms
  // full real torus (Rf) -> half spectrum X
  void r2c_full(const T* src, cx<T>* X) {
    xr2c<T>(M[1] * M[2], LoadRowR<T>{{}, src, M[0]}, StoreRowC<T>{{}, T2.p, H0});
    strided<T, false>(1, M[2], H0, M[1], M[1], LoadS<T>{{}, T2.p, (long)M[1] * H0, H0},
                      StoreS<T>{{}, T1.p, (long)M[1] * H0, H0});
    strided<T, false>(2, 1, M[1] * H0, M[2], M[2], LoadS<T>{{}, T1.p, 0, (long)M[1] * H0},
                      StoreS<T>{{}, X, 0, (long)M[1] * H0});
  }
  // half spectrum X -> full real torus dst, scaled by 1/N
  void c2r_full(const cx<T>* X, T* dst) {
    strided<T, true>(2, 1, M[1] * H0, M[2], M[2], LoadS<T>{{}, X, 0, (long)M[1] * H0},
                     StoreS<T>{{}, T1.p, 0, (long)M[1] * H0});
    strided<T, true>(1, M[2], H0, M[1], M[1], LoadS<T>{{}, T1.p, (long)M[1] * H0, H0},
                     StoreS<T>{{}, T2.p, (long)M[1] * H0, H0});
    xc2r<T>(M[1] * M[2], LoadRowC<T>{{}, T2.p, H0}, StoreRowR<T>{{}, dst, M[0], (T)(1.0 / (double)N)});
  }
  // inverse transform pruned to the image region; Pro is the z-pass prologue
  template <class Pro>
  void c2r_region(Pro pro, T* R) {
    strided<T, true>(2, 1, M[1] * H0, M[2], reg[2], pro, StoreS<T>{{}, T1.p, 0, (long)M[1] * H0});
    strided<T, true>(1, reg[2], H0, M[1], reg[1], LoadS<T>{{}, T1.p, (long)M[1] * H0, H0},
                     StoreS<T>{{}, T2.p, (long)reg[1] * H0, H0});
    xc2r<T>(reg[1] * reg[2], LoadRowC<T>{{}, T2.p, H0}, StoreRowR<T>{{}, R, reg[0], (T)(1.0 / (double)N)});
  }
  // forward transform from the image region; Pro is the x-pass prologue,
  // Epi the z-pass epilogue
  template <class Pro, class Epi>
  void r2c_region(Pro pro, Epi epi, int pro_slot, int epi_slot) {
    xr2c<T>(reg[1] * reg[2], pro, StoreRowC<T>{{}, T2.p, H0}, pro_slot);
    strided<T, false>(1, reg[2], H0, reg[1], M[1], LoadS<T>{{}, T2.p, (long)reg[1] * H0, H0},
                      StoreS<T>{{}, T1.p, (long)M[1] * H0, H0});
    strided<T, false>(2, 1, M[1] * H0, reg[2], M[2], LoadS<T>{{}, T1.p, 0, (long)M[1] * H0}, epi, epi_slot);
  }
  // full-torus noise field -> spectral momentum (or initial mu)
  template <class Epi>
  // full-torus noise field -> spectral field; the epilogue's sums go to slot
  void noise_r2c(uint64_t stream_id, const T* zinj, int slot, Epi epi) {
    if (zinj) {
      xr2c<T>(M[1] * M[2], LoadNoiseBuf<T>{zinj, M[0], nullptr, 0.0}, StoreRowC<T>{{}, T2.p, H0});
    } else {
      xr2c<T>(M[1] * M[2], LoadNoise<T>{cfg.seed, stream_id, M[0] / 2, nullptr, 0.0, gsc, (int)(stream_id & 1)},
              StoreRowC<T>{{}, T2.p, H0});
    }
    strided<T, false>(1, M[2], H0, M[1], M[1], LoadS<T>{{}, T2.p, (long)M[1] * H0, H0},
                      StoreS<T>{{}, T1.p, (long)M[1] * H0, H0});
    strided<T, false>(2, 1, M[1] * H0, M[2], M[2], LoadS<T>{{}, T1.p, 0, (long)M[1] * H0}, epi, slot);
  }

  // the spectral momentum drawn directly (k_momentum_direct), E0 sums to slot
  void momentum_direct(uint64_t stream_id, cx<T>* dst, int slot) {
    DR_REQUIRE(M[0] % 2 == 0 && M[1] % 2 == 0 && M[2] % 2 == 0, DR_ELOGIC, "momentum draw: odd torus size");
    const long Nh = (long)H0 * M[1] * M[2];
    DR_REQUIRE(Nh < (1L << 31), DR_ELOGIC, "momentum draw: half spectrum too large for 32-bit indices");
    const int g = (int)std::min<long>(cdiv(Nh / 2, kThreads), 4096);  // two modes per thread
    note_blocks(slot, g);
    const double nN = (double)M[0] * M[1] * M[2];
    hipLaunchKernelGGL(k_momentum_direct<T>, dim3(g), dim3(kThreads), 0, stream, Nh, H0, M[1], M[2], cfg.seed,
                       stream_id, (int)(stream_id & 1), (T)sigma[0], invlam.p, dst, pslot(slot), gsc,
                       (T)std::sqrt(0.5 * nN), (T)std::sqrt(nN));
    DR_HIP(hipGetLastError());
  }

  void finalize() {
    FinalizeArgs a{};
    a.nslot = NSLOT;
    for (int s = 0; s < NSLOT; ++s) a.nblk[s] = slot_nblk[s];
    a.pstride = pstride;
    hipLaunchKernelGGL(k_finalize, dim3(NSLOT), dim3(kThreads), 0, stream, a, part.p, sums_out ? sums_out : sums.p);
    DR_HIP(hipGetLastError());
  }
  void read_sums(double* h) {
    DR_HIP(hipMemcpyAsync(h, sums_out ? sums_out : sums.p, NSLOT * sizeof(double), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
  }
  void clear_slots() {
    for (int s = 0; s < NSLOT; ++s) slot_nblk[s] = 0;
  }

  // ------------------------------------------------------------ init
  void init_mu() {
    // mu = B(F(z) sqrt(lambda) / (10 N)) with z complex normal (:743-759)
    noise_r2c(2, nullptr, S_TMP, EpiInitMu<T>{{}, muA.p, lam.p, (long)M[1] * H0});
    if (compat()) noise_r2c(3, nullptr, S_TMP, EpiInitMu<T>{{}, miA.p, lam.p, (long)M[1] * H0});
    stats_valid = false;
  }

  // ------------------------------------------------------------ region / data
  void set_region(const int r[3]) override {
    drop_graphs();
    for (int a = 0; a < 3; ++a)
      DR_REQUIRE(r[a] >= 1 && r[a] <= M[a], DR_EDOMAIN, "image region exceeds the embedding torus");
    if (r[0] == reg[0] && r[1] == reg[1] && r[2] == reg[2]) return;
    for (int a = 0; a < 3; ++a) reg[a] = r[a];
    nreg = (long)reg[0] * reg[1] * reg[2];
    RA.alloc(nreg + 4);  // + 4: the aligned 4-voxel windows of W mu may end past the region
    RB.alloc(nreg + 4);
    RA.zero(stream);
    RB.zero(stream);
    Yg.alloc(nreg);
    mask.alloc(nreg);
    Yg.zero(stream);
    mask.zero(stream);
    std::vector<int32_t> g(n_h);
    for (long i = 0; i < n_h; ++i) g[i] = (int32_t)region_index(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
    gI.alloc(n_h);
    DR_HIP(hipMemcpyAsync(gI.p, g.data(), n_h * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    DR_HIP(hipStreamSynchronize(stream));
    if (data_set) upload_data();
    set_outputs();
    stats_valid = false;
  }

  // output voxels (the data voxels until set_output_indices) on the region
  // grid, and the running summary over them, which this resets
  void set_outputs() override {
    std::vector<int32_t> og;
    if (out_ijk.empty()) {
      og.resize(n_h);
      for (long i = 0; i < n_h; ++i) og[i] = (int32_t)region_index(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
    } else {
      og.resize(out_ijk.size() / 3);
      for (size_t i = 0; i < og.size(); ++i)
        og[i] = (int32_t)region_index(out_ijk[3 * i], out_ijk[3 * i + 1], out_ijk[3 * i + 2]);
    }
    n_out_dev = (long)og.size();
    const size_t no = std::max<size_t>(og.size(), 1);
    outg.alloc(no);
    outbuf.alloc(no);
    sm1.alloc(no);
    sm2.alloc(no);
    sma.alloc(no);
    if (!smax.p) smax.alloc(1);
    if (!og.empty())
      DR_HIP(hipMemcpyAsync(outg.p, og.data(), og.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    summary_reset();
    DR_HIP(hipStreamSynchronize(stream));
  }

  void get_mu_out(double* out, long n) override {
    DR_REQUIRE(n == n_out_dev, DR_ELOGIC, "dr_get_mu_output: wrong output length");
    refresh_stats();
    if (n == 0) return;
    hipLaunchKernelGGL(k_gather<T>, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, stream, n, outg.p, RA.p, outbuf.p);
    DR_HIP(hipGetLastError());
    std::vector<T> h(n);
    DR_HIP(hipMemcpyAsync(h.data(), outbuf.p, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
    for (long i = 0; i < n; ++i) out[i] = (double)h[i];
  }
  void summary_reset() override {
    sm1.zero(stream);
    sm2.zero(stream);
    sma.zero(stream);
    nsum = 0;
  }
  // mcmc_summary<T>::update (gaussian_process_model.h:205-224) of the current
  // state at the output voxels
  void summary_update() override {
    refresh_stats();
    const long n = n_out_dev;
    if (n == 0) return;
    const int g = std::min(cdiv(n, kThreads), 4096);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, smax.p);
    hipLaunchKernelGGL(k_sum_gather<T>, dim3(g), dim3(kThreads), 0, stream, n, outg.p, RA.p, outbuf.p, smax.p);
    hipLaunchKernelGGL(k_sum_acc<T>, dim3(g), dim3(kThreads), 0, stream, n, outbuf.p, smax.p, sm1.p, sm2.p, sma.p);
    DR_HIP(hipGetLastError());
    ++nsum;
  }
  void summary_read(double* s1, double* s2, double* sa, long n, int64_t* count) override {
    DR_REQUIRE(n == n_out_dev, DR_ELOGIC, "dr_summary_read: wrong output length");
    if (n > 0) {
      if (s1) DR_HIP(hipMemcpyAsync(s1, sm1.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (s2) DR_HIP(hipMemcpyAsync(s2, sm2.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (sa) DR_HIP(hipMemcpyAsync(sa, sma.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    DR_HIP(hipStreamSynchronize(stream));
    if (count) *count = nsum;
  }

  // ------------------------------------------------------------ class-stencil W
  // Rows equal up to translation share a class: (entry count, relative grid
  // offsets, weights) hashed, then verified entry by entry.
  static void classify(long nrows, const std::vector<int32_t>& rp, const std::vector<int32_t>& doff,
                       const std::vector<T>& w, std::vector<int32_t>& cls, std::vector<int32_t>& coff,
                       std::vector<int32_t>& cn, std::vector<int32_t>& cdoff, std::vector<T>& cw) {
    struct Key {
      uint64_t a, b;
      int n;
      bool operator==(const Key& o) const { return a == o.a && b == o.b && n == o.n; }
    };
    struct KH {
      size_t operator()(const Key& k) const { return (size_t)(k.a ^ (k.b * 0x9E3779B97F4A7C15ULL) ^ (uint64_t)k.n); }
    };
    auto mix = [](uint64_t x) {
      x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
      return x;
    };
    std::unordered_map<Key, std::vector<int32_t>, KH> seen;  // key -> candidate classes
    std::vector<int32_t> crow;                                // representative row of each class
    cls.assign(nrows, -1);
    for (long i = 0; i < nrows; ++i) {
      const int32_t a = rp[i], n = rp[i + 1] - rp[i];
      if (n == 0) continue;
      uint64_t h1 = 0, h2 = 0;
      for (int32_t t = 0; t < n; ++t) {
        uint64_t wb = 0;
        std::memcpy(&wb, &w[a + t], sizeof(T));
        const uint64_t k = ((uint64_t)(uint32_t)doff[a + t] << 32) ^ wb;
        h1 += mix(k + 0x9E3779B97F4A7C15ULL * (uint64_t)(t + 1));
        h2 += mix((k * 0xD6E8FEB86659FD93ULL) ^ ((uint64_t)(t + 7) * 0xA0761D6478BD642FULL));
      }
      auto& cand = seen[Key{h1, h2, n}];
      int32_t found = -1;
      for (int32_t c : cand) {
        const int32_t r = crow[c], b = rp[r];
        if (std::memcmp(&doff[a], &doff[b], n * sizeof(int32_t)) == 0 &&
            std::memcmp(&w[a], &w[b], n * sizeof(T)) == 0) {
          found = c;
          break;
        }
      }
      if (found < 0) {
        found = (int32_t)crow.size();
        crow.push_back((int32_t)i);
        coff.push_back((int32_t)cdoff.size());
        cn.push_back(n);
        cdoff.insert(cdoff.end(), doff.begin() + a, doff.begin() + a + n);
        cw.insert(cw.end(), w.begin() + a, w.begin() + a + n);
        cand.push_back(found);
      }
      cls[i] = found;
    }
  }

  template <typename V>
  void upload(DevBuf<V>& d, const std::vector<V>& h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) DR_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, stream));
  }

  // Stencil + boundary layout of W (interior and boundary blocks of k_wmu_sb
  // and k_wtv_sb).  cols: image-region grid index of every entry.  Returns
  // false (plain CSR kernels) when the boundary rows touching one voxel lie
  // more than 65535 rows apart, or the stencil is too wide for the LDS table
  // of W mu or the lattice box of a W^T v tile.  Boundary rows wider than 16
  // voxels on an axis become far rows.
  // Entry order for the quad loads of the boundary blocks: a list
  // [start, end) (a multiple of 4) is cut into blocks of `lanes` quads; in a
  // block of nq quads, quad l holds entries l, l + nq, l + 2 nq, l + 3 nq, so
  // the k-th gather of all lanes reads consecutive entries.
  static void interleave_quads(std::vector<uint16_t>& d, std::vector<T>& w, long start, long end, int lanes) {
    std::vector<uint16_t> td;
    std::vector<T> tw;
    for (long b0 = start; b0 < end; b0 += 4L * lanes) {
      const long nb = std::min<long>(4L * lanes, end - b0), nq = nb / 4;
      td.assign(d.begin() + b0, d.begin() + b0 + nb);
      tw.assign(w.begin() + b0, w.begin() + b0 + nb);
      for (long l = 0; l < nq; ++l)
        for (int k = 0; k < 4; ++k) {
          d[b0 + 4 * l + k] = td[l + k * nq];
          w[b0 + 4 * l + k] = tw[l + k * nq];
        }
    }
  }
  static long fdiv(long a, long b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
  static long fmod_(long a, long b) { return a - b * fdiv(a, b); }
  bool build_stencil(const std::vector<int32_t>& rp, const std::vector<int32_t>& cols, const std::vector<T>& vals) {
    const long nnz = rp[n_s];
    const long R0 = reg[0], R01 = (long)reg[0] * reg[1];
    if (R01 >= (1L << 24)) return false;  // packed offsets unpack with 24-bit multiplies
    auto xyz = [&](long r, long* v) {
      v[0] = r % R0;
      v[1] = (r / R0) % reg[1];
      v[2] = r / R01;
    };
    // 1. translation classes of the rows; the largest is the interior stencil
    std::vector<int32_t> rg0(n_s, 0), doff(nnz);
    for (long i = 0; i < n_s; ++i) {
      if (rp[i + 1] == rp[i]) continue;
      rg0[i] = cols[rp[i]];
      for (int32_t t = rp[i]; t < rp[i + 1]; ++t) doff[t] = cols[t] - rg0[i];
    }
    std::vector<int32_t> rcls, coff, cn, cdoff;
    std::vector<T> cw;
    classify(n_s, rp, doff, vals, rcls, coff, cn, cdoff, cw);
    std::vector<long> cnt(cn.size(), 0);
    for (long i = 0; i < n_s; ++i)
      if (rcls[i] >= 0) cnt[rcls[i]]++;
    int cs = -1;
    for (int c = 0; c < (int)cn.size(); ++c)
      if (cs < 0 || cnt[c] > cnt[cs]) cs = c;
    if (cs >= 0 && cnt[cs] < 2) cs = -1;
    std::vector<char> interior(n_s, 0);
    std::vector<int32_t> D;
    std::vector<T> wD;
    std::vector<long> d3;  // nD x 3 displacements from the base
    long dmin[3] = {0, 0, 0}, dmax[3] = {0, 0, 0};
    if (cs >= 0) {
      long r = 0;
      while (rcls[r] != cs) ++r;
      const int nD = cn[cs];
      D.assign(cdoff.begin() + coff[cs], cdoff.begin() + coff[cs] + nD);
      wD.assign(cw.begin() + coff[cs], cw.begin() + coff[cs] + nD);
      d3.resize(3 * (size_t)nD);
      long b[3], e[3];
      xyz(rg0[r], b);
      for (int s = 0; s < nD; ++s) {
        xyz((long)rg0[r] + D[s], e);
        for (int a = 0; a < 3; ++a) {
          d3[3 * s + a] = e[a] - b[a];
          dmin[a] = std::min(dmin[a], e[a] - b[a]);
          dmax[a] = std::max(dmax[a], e[a] - b[a]);
        }
      }
      // a row is interior when its stencil stays inside the region grid on
      // every axis (then flat offsets and 3D displacements agree)
      for (long i = 0; i < n_s; ++i) {
        if (rcls[i] != cs) continue;
        xyz(rg0[i], b);
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = ok && b[a] + dmin[a] >= 0 && b[a] + dmax[a] < reg[a];
        interior[i] = ok;
      }
    }
    std::vector<int32_t> irow, ibase;
    for (long i = 0; i < n_s; ++i)
      if (interior[i]) {
        irow.push_back((int32_t)i);
        ibase.push_back(rg0[i]);
      }
    const long n_int = (long)irow.size();
    // 2. lattice of the interior bases and the residue-class sub-stencils
    long c[3] = {0, 0, 0}, sg[3] = {1, 1, 1};
    if (n_int > 0) {
      xyz(ibase[0], c);
      long gcd_[3] = {0, 0, 0}, b[3];
      for (long k = 0; k < n_int; ++k) {
        xyz(ibase[k], b);
        for (int a = 0; a < 3; ++a) gcd_[a] = std::gcd(gcd_[a], std::labs(b[a] - c[a]));
      }
      for (int a = 0; a < 3; ++a) {
        sg[a] = gcd_[a] > 0 ? std::min<long>(gcd_[a], 64) : 1;
        c[a] = fmod_(c[a], sg[a]);
      }
      for (long k = 0; k < n_int; ++k) {  // gcd capped: re-check membership
        xyz(ibase[k], b);
        for (int a = 0; a < 3; ++a)
          if (fmod_(b[a] - c[a], sg[a]) != 0) return false;
      }
    }
    const int nD = (int)D.size();
    long lo[3], hi[3], L[3], omin[3], omax[3];
    for (int a = 0; a < 3; ++a) {
      const long ubmin = fdiv(-c[a], sg[a]), ubmax = fdiv(reg[a] - 1 - c[a], sg[a]);
      omin[a] = 0;
      omax[a] = 0;
      for (int s = 0; s < nD; ++s) {
        const long o = -fdiv(d3[3 * s + a], sg[a]);
        omin[a] = std::min(omin[a], o);
        omax[a] = std::max(omax[a], o);
      }
      lo[a] = ubmin + omin[a];
      hi[a] = ubmax + omax[a];
      L[a] = hi[a] - lo[a] + 1;
    }
    const long nVl = L[0] * L[1] * L[2];
    DR_REQUIRE(nVl < (1L << 31), DR_EDOMAIN, "W stencil lattice too large");
    auto flatL = [&](const long* u) { return (u[0] - lo[0]) + L[0] * ((u[1] - lo[1]) + L[1] * (u[2] - lo[2])); };
    const int ncls = (int)(sg[0] * sg[1] * sg[2]);
    std::vector<int32_t> sub_start(ncls + 1, 0), sub_o(3 * (size_t)nD);
    std::vector<T> sub_w(nD);
    {
      std::vector<int> scls(nD);
      for (int s = 0; s < nD; ++s) {
        long rho[3];
        for (int a = 0; a < 3; ++a) rho[a] = fmod_(d3[3 * s + a], sg[a]);
        scls[s] = (int)(rho[0] + sg[0] * (rho[1] + sg[1] * rho[2]));
        sub_start[scls[s] + 1]++;
      }
      for (int k = 0; k < ncls; ++k) sub_start[k + 1] += sub_start[k];
      std::vector<int32_t> fill(sub_start.begin(), sub_start.end() - 1);
      for (int s = 0; s < nD; ++s) {  // stencil order within a class
        long o[3];
        for (int a = 0; a < 3; ++a) o[a] = -fdiv(d3[3 * s + a], sg[a]);
        const int k = fill[scls[s]]++;
        for (int a = 0; a < 3; ++a) sub_o[3 * k + a] = (int32_t)o[a];
        sub_w[k] = wD[s];
      }
    }
    std::vector<int32_t> ilidx(n_int);
    for (long k = 0; k < n_int; ++k) {
      long b[3], u[3];
      xyz(ibase[k], b);
      for (int a = 0; a < 3; ++a) u[a] = (b[a] - c[a]) / sg[a];
      ilidx[k] = (int32_t)flatL(u);
    }
    // 2b. interior W mu by aligned 4-voxel windows: a row whose base b has
    //     b mod 4 = p reads its stencil as 16-byte-aligned windows (offsets
    //     off = 4 w - p from b; weight 0 where a window voxel is no stencil
    //     entry), so one vector load serves up to four entries.  Rows are
    //     bucketed by p, one parity per workgroup of kThreads / 8 rows.
    std::vector<int32_t> wstart(5, 0), woff;
    std::vector<T> ww;
    for (int p = 0; p < 4; ++p) {
      std::unordered_map<long, int> at;
      for (int e = 0; e < nD; ++e) {
        const long a4 = p + (long)D[e], wdw = fdiv(a4, 4), off = 4 * wdw - p;
        auto it = at.find(off);
        int idx = 0;
        if (it == at.end()) {
          idx = (int)woff.size();
          at.emplace(off, idx);
          woff.push_back((int32_t)off);
          ww.insert(ww.end(), 4, (T)0);
        } else {
          idx = it->second;
        }
        ww[4 * (size_t)idx + (size_t)(a4 - 4 * wdw)] = wD[e];
      }
      wstart[p + 1] = (int32_t)woff.size();
    }
    long ncw = 0;
    for (int p = 0; p < 4; ++p) ncw = std::max<long>(ncw, wstart[p + 1] - wstart[p]);
    if (ncw * (4 + 4 * (long)sizeof(T)) > 48 * 1024) return false;  // LDS window table of W mu
    const int rpb = kThreads / kWmuLanes;
    std::vector<int32_t> mrow, mbase, mlidx, mpar;
    for (int p = 0; p < 4; ++p) {
      const size_t s0 = mrow.size();
      for (long k = 0; k < n_int; ++k)
        if (ibase[k] % 4 == p) {
          mrow.push_back(irow[k]);
          mbase.push_back(ibase[k]);
          mlidx.push_back(ilidx[k]);
        }
      while (mrow.size() % rpb) {
        mrow.push_back(-1);
        mbase.push_back(0);
        mlidx.push_back(0);
      }
      for (size_t b = s0; b < mrow.size(); b += rpb) mpar.push_back(p);
    }
    // 3. interior W^T v by tiles: voxel tiles T0 x T1 x T2 that hold a data
    //    voxel; each stages the lattice box its voxels reach (LT0 x LT1 x LT2)
    //    in LDS.  T = (32, 4, 2) lattice steps per axis, so one residue class
    //    of a tile is 256 voxels with 32 consecutive lattice x per half-wave;
    //    halved (z, then y, then x) until the box fits 48 KiB of LDS.
    std::vector<int32_t> gidx(n_h);
    for (long d = 0; d < n_h; ++d) gidx[d] = (int32_t)region_index(ijk[3 * d], ijk[3 * d + 1], ijk[3 * d + 2]);
    long Tt[3] = {32 * sg[0], 4 * sg[1], 2 * sg[2]}, LT[3];
    auto box_dims = [&](const long* t, long* lt) {
      for (int a = 0; a < 3; ++a) lt[a] = (t[a] + sg[a] - 1) / sg[a] + 1 + (omax[a] - omin[a]);
    };
    while (true) {
      box_dims(Tt, LT);
      if (LT[0] * LT[1] * LT[2] * (long)sizeof(T) <= 48 * 1024) break;
      const int ax = Tt[2] > sg[2] ? 2 : Tt[1] > sg[1] ? 1 : Tt[0] > sg[0] ? 0 : -1;
      if (ax < 0) return false;  // stencil too wide for one LDS box
      Tt[ax] = std::max(sg[ax], Tt[ax] / 2);
    }
    const long nT[3] = {cdiv(reg[0], Tt[0]), cdiv(reg[1], Tt[1]), cdiv(reg[2], Tt[2])};
    std::vector<char> has(nT[0] * nT[1] * nT[2], 0);
    for (long d = 0; d < n_h; ++d) {
      long x[3];
      xyz(gidx[d], x);
      has[x[0] / Tt[0] + nT[0] * (x[1] / Tt[1] + nT[1] * (x[2] / Tt[2]))] = 1;
    }
    std::vector<int32_t> tiles;
    for (long t = 0; t < (long)has.size(); ++t)
      if (has[t]) {
        tiles.push_back((int32_t)((t % nT[0]) * Tt[0]));
        tiles.push_back((int32_t)(((t / nT[0]) % nT[1]) * Tt[1]));
        tiles.push_back((int32_t)((t / (nT[0] * nT[1])) * Tt[2]));
      }
    std::vector<int32_t> sub_ldo(nD);
    for (int e = 0; e < nD; ++e)
      sub_ldo[e] = (int32_t)(sub_o[3 * e] + LT[0] * (sub_o[3 * e + 1] + LT[1] * sub_o[3 * e + 2]));
    // 4. boundary rows: packed offsets from the row's bounding-box corner; v
    //    of boundary row k goes to Vc[k], a compact array.  Rows that span
    //    more than 16 voxels on an axis (the reference's flat-offset
    //    neighbourhood can wrap across a row end) are far rows: explicit
    //    region indices, v in a second compact array.
    std::vector<int32_t> brow, brp{0}, bbase;
    std::vector<uint16_t> bd;
    std::vector<T> bw;
    std::vector<int32_t> row_bnd(n_s, -1), row_far(n_s, -1);
    std::vector<int32_t> frow, frp{0}, fcol;
    std::vector<T> fwt;
    long nb_real = 0;
    for (long i = 0; i < n_s; ++i) {
      if (interior[i]) continue;
      long mn[3] = {LONG_MAX, LONG_MAX, LONG_MAX}, mx[3] = {LONG_MIN, LONG_MIN, LONG_MIN}, x[3];
      for (int32_t t = rp[i]; t < rp[i + 1]; ++t) {
        xyz(cols[t], x);
        for (int a = 0; a < 3; ++a) {
          mn[a] = std::min(mn[a], x[a]);
          mx[a] = std::max(mx[a], x[a]);
        }
      }
      if (rp[i + 1] > rp[i]) {
        bool wide = false;
        for (int a = 0; a < 3; ++a) wide = wide || mx[a] - mn[a] > 15;
        if (wide) {  // far row
          row_far[i] = (int32_t)frow.size();
          frow.push_back((int32_t)i);
          for (int32_t t = rp[i]; t < rp[i + 1]; ++t) {
            fcol.push_back(cols[t]);
            fwt.push_back(vals[t]);
          }
          frp.push_back((int32_t)fcol.size());
          nb_real += rp[i + 1] - rp[i];
          continue;
        }
      }
      row_bnd[i] = (int32_t)brow.size();
      const long base = rp[i + 1] > rp[i] ? mn[0] + R0 * mn[1] + R01 * mn[2] : 0;
      nb_real += rp[i + 1] - rp[i];
      for (int32_t t = rp[i]; t < rp[i + 1]; ++t) {
        xyz(cols[t], x);
        bd.push_back((uint16_t)((x[0] - mn[0]) | ((x[1] - mn[1]) << 5) | ((x[2] - mn[2]) << 10)));
        bw.push_back(vals[t]);
      }
      while (bw.size() % 4) {  // whole quads: offset 0 (the base), weight 0
        bd.push_back(0);
        bw.push_back((T)0);
      }
      interleave_quads(bd, bw, brp.back(), (long)bw.size(), kWmuBndLanes);
      brow.push_back((int32_t)i);
      bbase.push_back((int32_t)base);
      brp.push_back((int32_t)bw.size());
    }
    // 5. transposed boundary entries per touched data voxel, in row order:
    //    quads of (boundary row - the voxel's first boundary row, weight) as
    //    16-bit deltas into Vc, then (far row, weight) pairs
    std::vector<int32_t> tcount(n_h + 1, 0), fcount(n_h + 1, 0);
    for (long i = 0; i < n_s; ++i)
      if (!interior[i])
        for (long t = W_rp[i]; t < W_rp[i + 1]; ++t) (row_far[i] >= 0 ? fcount : tcount)[W_cols[t] + 1]++;
    std::vector<int32_t> tstart(n_h + 1, 0);
    for (long d = 0; d < n_h; ++d) tstart[d + 1] = tstart[d] + ((tcount[d + 1] + 3) & ~3);
    const long nbt = tstart[n_h];
    std::vector<uint16_t> tdall(nbt, 0);  // pad: delta 0, weight 0
    std::vector<T> twall(nbt, (T)0);
    std::vector<int32_t> vbase(n_h, -1);  // first boundary row touching each voxel
    {
      std::vector<int32_t> fill(tstart.begin(), tstart.end() - 1);
      for (long i = 0; i < n_s; ++i) {
        if (row_bnd[i] < 0) continue;
        for (long t = W_rp[i]; t < W_rp[i + 1]; ++t) {
          const int32_t dv = W_cols[t];
          if (vbase[dv] < 0) vbase[dv] = row_bnd[i];  // rows come in increasing order
          const long delta = (long)row_bnd[i] - vbase[dv];
          if (delta > 65535) return false;  // rows of one voxel too far apart for 16-bit deltas
          tdall[fill[dv]] = (uint16_t)delta;
          twall[fill[dv]] = vals[t];
          fill[dv]++;
        }
      }
    }
    for (long d = 0; d < n_h; ++d) interleave_quads(tdall, twall, tstart[d], tstart[d + 1], kWtvLanes);
    std::vector<int32_t> fstart(n_h + 1, 0);
    for (long d = 0; d < n_h; ++d) fstart[d + 1] = fstart[d] + fcount[d + 1];
    std::vector<int32_t> tfidx(fstart[n_h]);
    std::vector<T> tfw(fstart[n_h]);
    {
      std::vector<int32_t> fill(fstart.begin(), fstart.end() - 1);
      for (long i = 0; i < n_s; ++i) {
        if (row_far[i] < 0) continue;
        for (long t = W_rp[i]; t < W_rp[i + 1]; ++t) {
          const int32_t dv = W_cols[t];
          tfidx[fill[dv]] = row_far[i];
          tfw[fill[dv]] = vals[t];
          fill[dv]++;
        }
      }
    }
    std::vector<int32_t> tg, trp{0}, tfp{0}, tfp_start, tbase;
    for (long d = 0; d < n_h; ++d)
      if (tstart[d + 1] > tstart[d] || fstart[d + 1] > fstart[d]) {
        tg.push_back(gidx[d]);
        tbase.push_back(std::max(vbase[d], 0));
        trp.push_back(tstart[d + 1]);
        tfp_start.push_back(fstart[d]);
        tfp.push_back(fstart[d + 1]);
      }
    // packed ranges of touched voxels are [trp[k], trp[k + 1]) only when the
    // untouched voxels between them have empty lists, which holds: tstart is
    // cumulative over all data voxels and untouched voxels add nothing.
    for (size_t k = 0; k < tfp_start.size(); ++k)
      DR_REQUIRE(tfp_start[k] == tfp[k], DR_ELOGIC, "W stencil: far transposed ranges not contiguous");
    // 5b. touched voxels by decreasing entry count: the kWtvLanes-lane groups
    //     of a wavefront then run loops of about the same length (entry
    //     counts range from 1 to ~140 at 1 mm; in voxel order a wavefront
    //     waits on its longest voxel).  Each voxel keeps its entry list.
    {
      const size_t nt = tg.size();
      std::vector<size_t> ord(nt);
      std::iota(ord.begin(), ord.end(), (size_t)0);
      auto len = [&](size_t k) { return (long)(trp[k + 1] - trp[k]) + (long)(tfp[k + 1] - tfp[k]); };
      std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return len(x) > len(y); });
      std::vector<uint16_t> td2;
      std::vector<T> tw2, tfw2;
      std::vector<int32_t> tg2, tb2, trp2{0}, tfp2{0}, tfi2;
      td2.reserve(tdall.size());
      tw2.reserve(twall.size());
      for (size_t o : ord) {
        tg2.push_back(tg[o]);
        tb2.push_back(tbase[o]);
        td2.insert(td2.end(), tdall.begin() + trp[o], tdall.begin() + trp[o + 1]);
        tw2.insert(tw2.end(), twall.begin() + trp[o], twall.begin() + trp[o + 1]);
        trp2.push_back((int32_t)td2.size());
        tfi2.insert(tfi2.end(), tfidx.begin() + tfp[o], tfidx.begin() + tfp[o + 1]);
        tfw2.insert(tfw2.end(), tfw.begin() + tfp[o], tfw.begin() + tfp[o + 1]);
        tfp2.push_back((int32_t)tfi2.size());
      }
      tg.swap(tg2);
      tbase.swap(tb2);
      tdall.swap(td2);
      twall.swap(tw2);
      trp.swap(trp2);
      tfidx.swap(tfi2);
      tfw.swap(tfw2);
      tfp.swap(tfp2);
    }
    // 5c. the 64 / kWtvLanes voxels of one wavefront share one quad range:
    //     voxel i's quad q (lane q % kWtvLanes, trip q / kWtvLanes) goes to
    //     slot (q / kWtvLanes) * 64 + i * kWtvLanes + q % kWtvLanes, so each
    //     load instruction of the wavefront reads 64 consecutive quads.  The
    //     range lasts as many trips as the group's longest list; the rest is
    //     padding (delta 0, weight 0).  tgs: the groups' entry offsets.
    std::vector<int32_t> tgs{0};
    {
      constexpr int VPW = 64 / kWtvLanes;
      const long nt = (long)tg.size(), ngrp = (nt + VPW - 1) / VPW;
      for (long g = 0; g < ngrp; ++g) {
        long trips = 0;
        for (long k = g * VPW; k < std::min(nt, (g + 1) * VPW); ++k)
          trips = std::max(trips, ((long)(trp[k + 1] - trp[k]) / 4 + kWtvLanes - 1) / kWtvLanes);
        DR_REQUIRE((long)tgs.back() + trips * 256 < (1L << 31), DR_ELOGIC, "W stencil: too many transposed entries");
        tgs.push_back((int32_t)(tgs.back() + trips * 256));
      }
      std::vector<uint16_t> tdg(tgs.back(), 0);
      std::vector<T> twg(tgs.back(), (T)0);
      for (long k = 0; k < nt; ++k) {
        const long g = k / VPW, i = k % VPW, nq = (trp[k + 1] - trp[k]) / 4;
        for (long q = 0; q < nq; ++q) {
          const long dst = 4 * (tgs[g] / 4 + (q / kWtvLanes) * 64 + i * kWtvLanes + q % kWtvLanes);
          for (int e = 0; e < 4; ++e) {
            tdg[dst + e] = tdall[trp[k] + 4 * q + e];
            twg[dst + e] = twall[trp[k] + 4 * q + e];
          }
        }
      }
      tdall.swap(tdg);
      twall.swap(twg);
    }
    // 6. every device access of the four kernels stays inside its buffer
    {
      long fmin = 0, fmax = 0;
      for (int s = 0; s < nD; ++s) {
        fmin = std::min(fmin, (long)D[s]);
        fmax = std::max(fmax, (long)D[s]);
      }
      for (long k = 0; k < n_int; ++k) {
        DR_REQUIRE(ilidx[k] >= 0 && ilidx[k] < nVl, DR_ELOGIC, "W stencil: lattice index out of range");
        DR_REQUIRE((long)ibase[k] + fmin >= 0 && (long)ibase[k] + fmax < nreg, DR_ELOGIC,
                   "W stencil: interior entry outside the region");
      }
      // interior W^T v tiles: every voxel of every class reads LDS cells
      // base + ldo with base = its lattice point less the box origin
      const long ncell = LT[0] * LT[1] * LT[2];
      for (size_t t = 0; t < tiles.size(); t += 3) {
        long X0[3], u0[3];
        for (int a = 0; a < 3; ++a) {
          X0[a] = tiles[t + a];
          DR_REQUIRE(X0[a] >= 0 && X0[a] < reg[a], DR_ELOGIC, "W stencil: tile outside the region");
          u0[a] = fdiv(X0[a] - c[a], sg[a]) + omin[a];
        }
        for (int k = 0; k < ncls; ++k) {
          const long rho[3] = {k % sg[0], (k / sg[0]) % sg[1], k / (sg[0] * sg[1])};
          long bmin = 0, bmax = 0, st = 1;
          bool empty = false;
          for (int a = 0; a < 3; ++a) {
            const long r = fmod_(rho[a] - (X0[a] - c[a]), sg[a]);
            if (r >= Tt[a]) empty = true;
            const long n = r < Tt[a] ? (Tt[a] - r + sg[a] - 1) / sg[a] : 0;
            const long ub = fdiv(X0[a] + r - c[a], sg[a]) - u0[a];
            // per axis: ub + j + o in [0, LT) for j < n and o in [omin, omax]
            DR_REQUIRE(ub + omin[a] >= 0 && ub + n - 1 + omax[a] < LT[a], DR_ELOGIC,
                       "W stencil: tile voxel reach outside its box");
            bmin += st * ub;
            bmax += st * (ub + n - 1);
            st *= LT[a];
          }
          if (empty) continue;
          for (int e = sub_start[k]; e < sub_start[k + 1]; ++e)
            DR_REQUIRE(bmin + sub_ldo[e] >= 0 && bmax + sub_ldo[e] < ncell, DR_ELOGIC,
                       "W stencil: sub-stencil read outside the tile box");
        }
      }
      for (int e = 0; e < nD; ++e)
        for (int a = 0; a < 3; ++a)
          DR_REQUIRE(sub_o[3 * e + a] >= omin[a] && sub_o[3 * e + a] <= omax[a], DR_ELOGIC,
                     "W stencil: sub-stencil offset outside the box margin");
    }
    for (size_t k = 0; k < brow.size(); ++k) {
      for (int32_t t = brp[k]; t < brp[k + 1]; ++t) {
        const long q = bd[t], j = (long)bbase[k] + (q & 31) + R0 * ((q >> 5) & 31) + R01 * (q >> 10);
        DR_REQUIRE(j >= 0 && j < nreg, DR_ELOGIC, "W stencil: boundary entry outside the region");
      }
    }
    for (size_t k = 0; k < tg.size(); ++k)
      DR_REQUIRE(tg[k] >= 0 && tg[k] < nreg, DR_ELOGIC, "W stencil: touched voxel outside the region");
    DR_REQUIRE((long)tgs.size() == ((long)tg.size() + 64 / kWtvLanes - 1) / (64 / kWtvLanes) + 1, DR_ELOGIC,
               "W stencil: transposed group offsets");
    for (size_t g = 0; g + 1 < tgs.size(); ++g)
      for (int32_t t = tgs[g]; t < tgs[g + 1]; ++t) {
        const size_t k = g * (64 / kWtvLanes) + ((t - tgs[g]) / 4 % 64) / kWtvLanes;  // the voxel reading t
        if (k < tg.size())
          DR_REQUIRE((long)tbase[k] + tdall[t] < std::max<long>((long)brow.size(), 1), DR_ELOGIC,
                     "W stencil: transposed entry outside the boundary rows");
      }
    for (size_t k = 0; k < mrow.size(); ++k) {
      if (mrow[k] < 0) continue;
      const int p = mpar[k / rpb];
      DR_REQUIRE(mbase[k] % 4 == p, DR_ELOGIC, "W stencil: row parity");
      for (int c = wstart[p]; c < wstart[p + 1]; ++c)
        DR_REQUIRE((long)mbase[k] + woff[c] >= 0 && (long)mbase[k] + woff[c] + 3 < nreg + 4 &&
                       ((long)mbase[k] + woff[c]) % 4 == 0,
                   DR_ELOGIC, "W stencil: window outside the padded region or unaligned");
    }
    for (size_t t = 0; t < fcol.size(); ++t)
      DR_REQUIRE(fcol[t] >= 0 && fcol[t] < nreg, DR_ELOGIC, "W stencil: far-row entry outside the region");
    for (size_t t = 0; t < tfidx.size(); ++t)
      DR_REQUIRE(tfidx[t] >= 0 && tfidx[t] < (long)frow.size(), DR_ELOGIC, "W stencil: far row index out of range");
    DR_REQUIRE(tfp.size() == tg.size() + 1 && trp.size() == tg.size() + 1, DR_ELOGIC,
               "W stencil: transposed row pointers");
    // 7. upload
    ws.n_int = (int)n_int;
    ws.n_far = (int)frow.size();
    ws.nD = nD;
    ws.n_bnd = (int)brow.size();
    ws.n_touch = (int)tg.size();
    ws.n_tiles = (int)(tiles.size() / 3);
    for (int a = 0; a < 3; ++a) {
      ws.tdim[a] = (int)Tt[a];
      ws.ldim[a] = (int)LT[a];
      ws.c[a] = (int)c[a];
      ws.sg[a] = (int)sg[a];
      ws.omin[a] = (int)omin[a];
      ws.lo[a] = (int)lo[a];
      ws.lv[a] = (int)L[a];
    }
    ws.ncls = ncls;
    upload(ws.irow, mrow);
    upload(ws.ibase, mbase);
    upload(ws.ilidx, mlidx);
    upload(ws.mpar, mpar);
    upload(ws.wstart, wstart);
    upload(ws.woff, woff);
    upload(ws.ww, ww);
    ws.n_mslot = (int)mrow.size();
    ws.ncw = (int)ncw;
    upload(ws.sub_start, sub_start);
    upload(ws.sub_ldo, sub_ldo);
    upload(ws.sub_w, sub_w);
    upload(ws.tiles, tiles);
    upload(ws.brow, brow);
    upload(ws.brp, brp);
    upload(ws.bbase, bbase);
    upload(ws.tbase, tbase);
    upload(ws.bd, bd);
    upload(ws.bw, bw);
    upload(ws.tg, tg);
    upload(ws.trp, tgs);
    upload(ws.td, tdall);
    upload(ws.tw, twall);
    upload(ws.frow, frow);
    upload(ws.frp, frp);
    upload(ws.fcol, fcol);
    upload(ws.fw, fwt);
    upload(ws.tfp, tfp);
    upload(ws.tfidx, tfidx);
    upload(ws.tfw, tfw);
    ws.Vf.alloc(std::max<long>((long)frow.size(), 1));
    ws.Vf.zero(stream);
    ws.Vl.alloc(std::max<long>(nVl, 1));
    ws.Vl.zero(stream);
    ws.Vc.alloc(std::max<long>((long)brow.size(), 1));
    ws.Vc.zero(stream);
    DR_HIP(hipStreamSynchronize(stream));
    wstats[0] = n_int;
    wstats[1] = nD;
    wstats[2] = (int64_t)(brow.size() + frow.size());
    wstats[3] = nb_real;
    wstats[4] = nnz;
    wstats[5] = (int64_t)tg.size();
    wstats[6] = (int64_t)(tiles.size() / 3);
    wstats[7] = LT[0] * LT[1] * LT[2];
    wstats[8] = Tt[0] * Tt[1] * Tt[2];
    wstats[9] = (int64_t)frow.size();
    return true;
  }

  void upload_data() override {
    drop_graphs();
    // data on the region grid
    std::vector<T> yg(nreg, (T)0);
    std::vector<uint8_t> mk_(nreg, 0);
    std::vector<int32_t> g(n_h);
    for (long i = 0; i < n_h; ++i) {
      g[i] = (int32_t)region_index(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
      yg[g[i]] = (T)Yh[i];
      mk_[g[i]] = 1;
    }
    DR_HIP(hipMemcpyAsync(Yg.p, yg.data(), nreg * sizeof(T), hipMemcpyHostToDevice, stream));
    DR_HIP(hipMemcpyAsync(mask.p, mk_.data(), nreg, hipMemcpyHostToDevice, stream));
    if (cfg.n_datasets == 2) {
      const long need = cdiv(n_s, kThreads / 64) + cdiv(n_s * kWmuLanes, kThreads) + 8;  // + parity-bucket padding
      if (need > pstride) {
        pstride = need;
        part.alloc((size_t)NSLOT * pstride);
        part.zero(stream);
      }
      const long nnz = W_rp[n_s];
      DR_REQUIRE(nnz < (1L << 31), DR_EDOMAIN, "W has too many nonzeros for 32-bit indices");
      std::vector<int32_t> rp(n_s + 1), cols(nnz);
      std::vector<T> vals(nnz);
      for (long i = 0; i <= n_s; ++i) rp[i] = (int32_t)W_rp[i];
      for (long j = 0; j < nnz; ++j) {
        const int32_t c = W_cols[j];
        DR_REQUIRE(c >= 0 && c < n_h, DR_EDOMAIN, "W column index out of range");
        cols[j] = g[c];
        vals[j] = (T)W_vals[j];
      }
      Ysd.alloc(n_s);
      vsd.alloc(n_s);
      WTv.alloc(nreg);
      WTv.zero(stream);
      WTvb.alloc(nreg);
      WTvb.zero(stream);
      std::vector<T> ys(n_s);
      for (long i = 0; i < n_s; ++i) ys[i] = (T)Ys[i];
      DR_HIP(hipMemcpyAsync(Ysd.p, ys.data(), n_s * sizeof(T), hipMemcpyHostToDevice, stream));
      use_cls = !force_csr && build_stencil(rp, cols, vals);
      if (!use_cls) {
        // transposed CSR for W^T v, rows ascending within each column
        std::vector<int32_t> trp(n_h + 1, 0), trows(nnz);
        std::vector<T> tvals(nnz);
        for (long j = 0; j < nnz; ++j) trp[W_cols[j] + 1]++;
        for (long d = 0; d < n_h; ++d) trp[d + 1] += trp[d];
        std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
        for (long i = 0; i < n_s; ++i)
          for (long j = W_rp[i]; j < W_rp[i + 1]; ++j) {
            const int32_t c = W_cols[j];
            trows[fill[c]] = (int32_t)i;
            tvals[fill[c]] = (T)W_vals[j];
            fill[c]++;
          }
        Wrp.alloc(n_s + 1);
        Wcols.alloc(std::max<long>(nnz, 1));
        Wvals.alloc(std::max<long>(nnz, 1));
        WTrp.alloc(n_h + 1);
        WTrows.alloc(std::max<long>(nnz, 1));
        WTvals.alloc(std::max<long>(nnz, 1));
        DR_HIP(hipMemcpyAsync(Wrp.p, rp.data(), (n_s + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        if (nnz) {
          DR_HIP(hipMemcpyAsync(Wcols.p, cols.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, stream));
          DR_HIP(hipMemcpyAsync(Wvals.p, vals.data(), nnz * sizeof(T), hipMemcpyHostToDevice, stream));
          DR_HIP(hipMemcpyAsync(WTrows.p, trows.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice, stream));
          DR_HIP(hipMemcpyAsync(WTvals.p, tvals.data(), nnz * sizeof(T), hipMemcpyHostToDevice, stream));
        }
        DR_HIP(hipMemcpyAsync(WTrp.p, trp.data(), (n_h + 1) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
        DR_HIP(hipStreamSynchronize(stream));
      }
    }
    DR_HIP(hipStreamSynchronize(stream));
    data_set = true;
    stats_valid = false;
  }

  void add_data_to_mu() override {
    DR_REQUIRE(data_set, DR_ELOGIC, "add_data_to_mu: data not set");
    r2c_region(LoadRowR<T>{{}, Yg.p, reg[0]}, EpiAddTo<T>{{}, muA.p, (long)M[1] * H0}, -1, -1);
    stats_valid = false;
  }

  void get_mu_region_gather(const std::vector<int32_t>& g, double* out) override {
    refresh_stats();
    const long n = (long)g.size();
    if (n == 0) return;
    DevBuf<int> gd;
    gd.alloc(n);
    outbuf.alloc(n);
    DR_HIP(hipMemcpyAsync(gd.p, g.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_gather<T>, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, stream, n, gd.p, RA.p, outbuf.p);
    DR_HIP(hipGetLastError());
    std::vector<T> h(n);
    DR_HIP(hipMemcpyAsync(h.data(), outbuf.p, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
    for (long i = 0; i < n; ++i) out[i] = (double)h[i];
  }

  // ------------------------------------------------------------ gradient pieces
  // likelihood products at region field R: v = sigma_s (Ys - W R), WTv = W^T v
  void w_products(const T* R, int slot) {
    if (cfg.n_datasets != 2) return;
    const int g1 = cdiv(n_s, kThreads / 64);
    const int g2 = cdiv(n_h, kThreads / 16);
    if (use_cls) {
      const int gi = cdiv((long)ws.n_mslot * kWmuLanes, kThreads), gb = cdiv(ws.n_bnd, kThreads / kWmuBndLanes),
                gf = cdiv(ws.n_far, kThreads / 64);
      if (slot >= 0) note_blocks(slot, gi + gb + gf);
      const int R0 = reg[0], R01 = reg[0] * reg[1];
      if (gi + gb + gf > 0) {
        const WmuInt<T> ai{ws.n_mslot, ws.irow.p, ws.ibase.p, ws.ilidx.p, ws.mpar.p, ws.wstart.p, ws.woff.p,
                           ws.ww.p, ws.ncw, ws.Vl.p};
        const WmuBnd<T> ab{ws.n_bnd, ws.brow.p, ws.brp.p, ws.bbase.p, ws.bd.p, ws.bw.p, R0, R01, ws.Vc.p};
        const WmuFar<T> af{ws.n_far, ws.frow.p, ws.frp.p, ws.fcol.p, ws.fw.p, ws.Vf.p};
        hipLaunchKernelGGL(k_wmu_sb<T>, dim3(gi + gb + gf), dim3(kThreads), (size_t)ws.ncw * (4 + 4 * sizeof(T)), stream,
                           ai, ab, af, gi, gb, R, Ysd.p, (T)sigma[1], slot >= 0 ? pslot(slot) : nullptr, gsc);
        DR_HIP(hipGetLastError());
      }
      mark();
      const int ti = ws.n_tiles, tb = cdiv(ws.n_touch, kThreads / kWtvLanes);
      WtvInt<T> wi{};
      wi.tiles = ws.tiles.p;
      wi.cls_start = ws.sub_start.p;
      wi.ldo = ws.sub_ldo.p;
      wi.w = reinterpret_cast<const typename Bits<T>::type*>(ws.sub_w.p);
      wi.Vl = ws.Vl.p;
      wi.mask = mask.p;
      wi.out = WTv.p;
      wi.ncls = ws.ncls;
      for (int a = 0; a < 3; ++a) {
        wi.TD[a] = ws.tdim[a];
        wi.LT[a] = ws.ldim[a];
        wi.c[a] = ws.c[a];
        wi.sg[a] = ws.sg[a];
        wi.omin[a] = ws.omin[a];
        wi.lo[a] = ws.lo[a];
        wi.lv[a] = ws.lv[a];
        wi.R[a] = reg[a];
      }
      const WtvBnd<T> wb{ws.n_touch, ws.tg.p, ws.trp.p, ws.tbase.p, ws.td.p, ws.tw.p, ws.Vc.p, WTvb.p,
                         ws.tfp.p, ws.tfidx.p, ws.tfw.p, ws.Vf.p};
      const size_t lds_t = (size_t)ws.ldim[0] * ws.ldim[1] * ws.ldim[2] * sizeof(T);
      hipLaunchKernelGGL(k_wtv_sb<T>, dim3(ti + tb), dim3(kThreads), lds_t, stream, wi, wb, ti);
      DR_HIP(hipGetLastError());
      mark();
      return;
    }
    if (slot >= 0) note_blocks(slot, g1);
    hipLaunchKernelGGL(k_wmu<T>, dim3(g1), dim3(kThreads), 0, stream, (int)n_s, Wrp.p, Wcols.p, Wvals.p, R, Ysd.p,
                       (T)sigma[1], vsd.p, slot >= 0 ? pslot(slot) : nullptr, gsc);
    DR_HIP(hipGetLastError());
    mark();
    hipLaunchKernelGGL(k_wtv<T>, dim3(g2), dim3(kThreads), 0, stream, (int)n_h, WTrp.p, WTrows.p, WTvals.p, vsd.p,
                       gI.p, WTv.p);
    DR_HIP(hipGetLastError());
    mark();
  }
  LoadResidual<T> residual(const T* R, int slot) {
    return LoadResidual<T>{R, Yg.p, mask.p, cfg.n_datasets == 2 ? WTv.p : nullptr,
                           cfg.n_datasets == 2 && use_cls ? WTvb.p : nullptr, (T)sigma[0], reg[0],
                           slot >= 0 ? pslot(slot) : nullptr, 0.0, gsc};
  }
  // inverse y and x passes from T1 (z < n2) to the real region field R
  void c2r_yx(T* R) {
    strided<T, true>(1, reg[2], H0, M[1], reg[1], LoadS<T>{{}, T1.p, (long)M[1] * H0, H0},
                     StoreS<T>{{}, T2.p, (long)reg[1] * H0, H0}, -1, P_YI);
    xc2r<T>(reg[1] * reg[2], LoadRowC<T>{{}, T2.p, H0}, StoreRowR<T>{{}, R, reg[0], (T)(1.0 / (double)N)}, -1,
            P_XI);
  }
  // forward x and y passes of the scattered residual into T1 (z < n2)
  void r2c_xy(LoadResidual<T> pro, int pro_slot) {
    xr2c<T>(reg[1] * reg[2], pro, StoreRowC<T>{{}, T2.p, H0}, pro_slot, P_XF);
    strided<T, false>(1, reg[2], H0, reg[1], M[1], LoadS<T>{{}, T2.p, (long)reg[1] * H0, H0},
                      StoreS<T>{{}, T1.p, (long)M[1] * H0, H0}, -1, P_YF);
  }
  // forward z pass, kick + drift, inverse z pass, in place on T1
  void zfused(MidLeapfrog<T> mid, int mid_slot) {
    const AxisFFT<T>& ad = axis_dev<T>(2);
    const int NB = M[1] * H0;
    const int tbs = tuned_tbs(P_Z, pick_tbs(ad.M, ad.M, sizeof(cx<T>), false, NB, kTileStrided),
                              pick_tbs(ad.M, ad.M, sizeof(cx<T>), false, 1 << 20, kTileStrided));
    const int TB = 1 << tbs;
    const int grid = cdiv(NB, TB);
    const size_t lds = (size_t)ad.M * TB * sizeof(cx<T>);
    // the tile holds ad.M rows, of which the first reg[2] are loaded
    DR_REQUIRE(reg[2] >= 1 && reg[2] <= ad.M && NB >= 1, DR_ELOGIC, "z pass: bad extents");
    if (mid_slot >= 0) note_blocks(mid_slot, grid);
    const long sk = (long)M[1] * H0;
    with_fft_len(ad.M, [&](auto mc) {
      constexpr int MC = decltype(mc)::value;
      // state prefetch for lengths whose 16 per-thread modes fit in registers
      if constexpr (MC > 0 && MC <= 256 && MC % (kThreads / 16) == 0) {
        if (TB == 16) {
          hipLaunchKernelGGL((k_strided2_pf<T, LoadS<T>, MidLeapfrog<T>, StoreS<T>, kEptStrided, MC>),
                             dim3(grid), dim3(kThreads), lds, stream, ad, NB, reg[2], reg[2],
                             LoadS<T>{{}, T1.p, 0, sk}, mid, StoreS<T>{{}, T1.p, 0, sk});
          return;
        }
      }
      hipLaunchKernelGGL((k_strided2<T, LoadS<T>, MidLeapfrog<T>, StoreS<T>, kEptStrided, MC>), dim3(grid),
                         dim3(kThreads), lds, stream, ad, NB, reg[2], reg[2], tbs, LoadS<T>{{}, T1.p, 0, sk}, mid,
                         StoreS<T>{{}, T1.p, 0, sk});
    });
    DR_HIP(hipGetLastError());
    mark();
  }
  int zgrid() const {
    const int NB = M[1] * H0;
    return cdiv(NB, 1 << pick_tbs(M[2], M[2], sizeof(cx<T>), false, NB, kTileStrided));
  }

  // ------------------------------------------------------------ HMC iteration
  // Leapfrog from (muA, p) to (muB, p'), MultiResParameters.h:690-700:
  //   p += eps/2 g(mu); L times { mu += eps M^-1 p; p += k eps g(mu) }.
  // Each gradient is inverse z,y,x -> W products -> forward x,y,z; the
  // forward z pass of one gradient, the kick, the drift and the inverse z
  // pass of the next gradient run as one kernel.  Reductions give the prior
  // energy at A and B, the kinetic energy at the end and both SSEs.
  void leapfrog_core(double eps, int L) {
    const long sk = (long)M[1] * H0;
    c2r_region(LoadS<T>{{}, muA.p, 0, sk}, RA.p);
    w_products(RA.p, S_SSES_A);
    r2c_xy(residual(RA.p, S_SSEH_A), S_SSEH_A);
    for (int s = 0; s < L; ++s) {
      const bool last = s == L - 1;
      if (prof && L >= 2 && s == 1) DR_HIP(hipEventRecord(ev[0], stream));
      if (prof && L >= 3 && s == 1) {
        rec = 0;
        mark();
      }
      MidLeapfrog<T> mid{s == 0 ? muA.p : muB.p, muB.p, p.p, invlam.p, sk, (T)((s == 0 ? 0.5 : 1.0) * eps),
                         (T)eps, (T)sigma[0], H0, s == 0 ? 1 : 0, pslot(S_PRIOR_A), 0.0, gsc, s == 0 ? 0 : 1};
      zfused(mid, s == 0 ? S_PRIOR_A : -1);
      c2r_yx(RB.p);
      w_products(RB.p, last ? S_SSES_B : -1);
      r2c_xy(residual(RB.p, last ? S_SSEH_B : -1), last ? S_SSEH_B : -1);
      if (rec >= 0) {
        nrec = rec;
        rec = -1;
      }
      if (prof && L >= 2 && s == L - 1) DR_HIP(hipEventRecord(ev[1], stream));
    }
    // final half kick at B with the energies of the end point
    EpiMomentum<T> em{muB.p, p.p, invlam.p, sk, (T)(0.5 * eps), (T)sigma[0], H0, nullptr, nullptr, 1, part.p,
                      pstride, S_PRIOR_B, S_TMP, S_KIN_B, S_TMP, 0.0, 0.0, 0.0, 0.0, gsc};
    const int zb = zgrid();
    note_blocks(S_PRIOR_B, zb);
    note_blocks(S_KIN_B, zb);
    note_blocks(S_TMP, zb);
    strided<T, false>(2, 1, M[1] * H0, reg[2], M[2], LoadS<T>{{}, T1.p, 0, sk}, em);
    if (compat()) {
      // imaginary chain: all L steps per mode in one pass
      const int g = (int)std::min<long>(cdiv(Nh, kThreads), 4096);
      note_blocks(S_PRIORI_A, g);
      note_blocks(S_PRIORI_B, g);
      note_blocks(S_KINI_B, g);
      hipLaunchKernelGGL(k_imag_leapfrog<T>, dim3(g), dim3(kThreads), 0, stream, Nh, H0, miA.p, miB.p, pim.p,
                         invlam.p, (T)eps, (T)(0.5 * eps), (T)eps, (T)sigma[0], L, pslot(S_PRIORI_A),
                         pslot(S_PRIORI_B), pslot(S_KINI_B), gsc);
      DR_HIP(hipGetLastError());
    }
  }

  void accept_swap() {
    drop_graphs();  // captured graphs hold the old buffer roles
    std::swap(muA.p, muB.p);
    std::swap(RA.p, RB.p);
    if (compat()) std::swap(miA.p, miB.p);
  }

  // momentum draw and leapfrog of one HMC iteration, then the energy sums
  // (MultiResParameters.h:680-705); normals injected when zr is given
  void device_iteration(double eps, int L, uint64_t it, const double* zr, const double* zi) {
    const long sk = (long)M[1] * H0;
    clear_slots();
    // injected real-space normals go through the transforms; the product
    // path draws the spectral momentum directly
    if (zr) {
      upload_real(zr, zbuf);
      noise_r2c(4 * it + 0, zbuf.p, S_E0,
                EpiNoiseMomentum<T>{p.p, invlam.p, sk, (T)sigma[0], gsc, H0, pslot(S_E0), 0.0});
    } else {
      momentum_direct(4 * it + 0, p.p, S_E0);
    }
    if (compat()) {
      if (zr) {
        upload_real(zi, zbuf);
        noise_r2c(4 * it + 1, zbuf.p, S_E0I,
                  EpiNoiseMomentum<T>{pim.p, invlam.p, sk, (T)sigma[0], gsc, H0, pslot(S_E0I), 0.0});
      } else {
        momentum_direct(4 * it + 1, pim.p, S_E0I);
      }
    }
    leapfrog_core(eps, L);
    finalize();
  }

  // energies, accept test and the kept state's statistics from the sums h
  // (:701-727).  On accept the proposal becomes state A by a pointer swap, or
  // for a captured iteration (whose graph always reads A and writes B) by
  // copies enqueued on s.
  void finish_iteration(const double* h, double u, dr_update_result* res, bool copy, hipStream_t s) override {
    const double nN = (double)N;
    const double e0 = -0.5 * (h[S_E0] + (compat() ? h[S_E0I] : 0.0)) / nN;
    const double lp0 = -0.5 * (h[S_PRIOR_A] + (compat() ? h[S_PRIORI_A] : 0.0)) / nN;
    const double ll0 = log_lik(h[S_SSEH_A], h[S_SSES_A]);
    const double lp1 = -0.5 * (h[S_PRIOR_B] + (compat() ? h[S_PRIORI_B] : 0.0)) / nN;
    const double ll1 = log_lik(h[S_SSEH_B], h[S_SSES_B]);
    const double k1 = -0.5 * (h[S_KIN_B] + (compat() ? h[S_KINI_B] : 0.0)) / nN;
    const double H0e = e0 + lp0 + ll0;
    const double H1e = lp1 + ll1 + k1;
    double R = std::exp(H1e - H0e);
    if (std::isnan(R)) R = 0.0;
    const bool acc = u < R;
    if (acc) {
      if (copy) accept_copy(s);
      else accept_swap();
      prior_sum = h[S_PRIOR_B] + (compat() ? h[S_PRIORI_B] : 0.0);
      sse_h = h[S_SSEH_B];
      sse_s = h[S_SSES_B];
    } else {
      prior_sum = h[S_PRIOR_A] + (compat() ? h[S_PRIORI_A] : 0.0);
      sse_h = h[S_SSEH_A];
      sse_s = h[S_SSES_A];
    }
    stats_valid = true;
    if (res) {
      res->R = R;
      res->mh_rate = std::min(1.0, R);
      res->H0 = H0e;
      res->H1 = H1e;
      res->log_prior0 = lp0;
      res->log_lik0 = ll0;
      res->e0 = e0;
      res->log_prior1 = lp1;
      res->log_lik1 = ll1;
      res->k1 = k1;
      res->sse_h = sse_h;
      res->sse_s = sse_s;
      res->accepted = acc ? 1 : 0;
    }
  }

  void update(double eps, int L, double u, const double* zr, const double* zi, dr_update_result* res) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "update: data not set");
    DR_REQUIRE(L >= 1, DR_EDOMAIN, "update: L must be >= 1");
    const uint64_t it = iteration++;
    device_iteration(eps, L, it, zr, zi);
    double h[NSLOT];
    read_sums(h);
    if (prof && L >= 2) accumulate_profile(L - 1);
    finish_iteration(h, u, res, false, stream);
  }

  // ------------------------------------------------------------ captured iterations
  void accept_copy(hipStream_t s) {
    DR_HIP(hipMemcpyAsync(muA.p, muB.p, Nh * sizeof(cx<T>), hipMemcpyDeviceToDevice, s));
    if (compat()) DR_HIP(hipMemcpyAsync(miA.p, miB.p, Nh * sizeof(cx<T>), hipMemcpyDeviceToDevice, s));
    DR_HIP(hipMemcpyAsync(RA.p, RB.p, (nreg + 4) * sizeof(T), hipMemcpyDeviceToDevice, s));
  }
  // forget captured graphs when buffers move; a batch compares layout_gen
  void drop_graphs() {
    for (auto& kv : gexec) (void)hipGraphExecDestroy(kv.second);
    gexec.clear();
    ++layout_gen;
  }
  void ensure_scalars() {
    if (!dsc) {
      sc_own.alloc(1);
      dsc = sc_own.p;
    }
    if (!hsc) {
      void* h = nullptr;
      DR_HIP(hipHostMalloc(&h, sizeof(StepScalars<T>), hipHostMallocDefault));
      hsc = static_cast<StepScalars<T>*>(h);
      hsc_own = true;
    }
  }
  void set_batch_slots(void* dev_scalars, void* host_scalars, double* dev_sums) override {
    DR_HIP(hipStreamSynchronize(stream));
    drop_graphs();
    if (hsc_own && hsc) (void)hipHostFree(hsc);
    hsc_own = false;
    hsc = static_cast<StepScalars<T>*>(host_scalars);
    dsc = static_cast<StepScalars<T>*>(dev_scalars);
    sums_out = dev_sums;
    if (!dsc) sc_own.release();
  }
  // the next iteration's scalars in the host staging (the eager update's
  // conversions to T and Philox streams 4 it, 4 it + 1)
  void stage_scalars(double eps) override {
    ensure_scalars();
    const uint64_t it = iteration++;
    hsc->half = (T)(0.5 * eps);
    hsc->full = (T)eps;
    hsc->sig0 = (T)sigma[0];
    hsc->sig1 = (T)sigma[1];
    hsc->stream[0] = 4 * it + 0;
    hsc->stream[1] = 4 * it + 1;
  }
  // the device work of one iteration with the leapfrog length L as a graph
  hipGraph_t capture_iteration(int L) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "update: data not set");
    DR_REQUIRE(L >= 1, DR_EDOMAIN, "update: L must be >= 1");
    ensure_scalars();
    const bool prof0 = prof;
    prof = false;
    rec = -1;
    hipGraph_t g = nullptr;
    DR_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      gsc = dsc;
      device_iteration(0.0, L, 0, nullptr, nullptr);
      gsc = nullptr;
    } catch (...) {
      gsc = nullptr;
      prof = prof0;
      hipGraph_t tmp = nullptr;
      (void)hipStreamEndCapture(stream, &tmp);
      if (tmp) (void)hipGraphDestroy(tmp);
      throw;
    }
    prof = prof0;
    DR_HIP(hipStreamEndCapture(stream, &g));
    return g;
  }
  hipGraphExec_t graph_for(int L) {
    auto f = gexec.find(L);
    if (f != gexec.end()) return f->second;
    hipGraph_t g = capture_iteration(L);
    hipGraphExec_t e = nullptr;
    const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    DR_HIP(err);
    gexec[L] = e;
    return e;
  }
  // stage the scalars, copy them, launch the graph for L and copy the sums
  // back to the host staging after it, all on s
  void enqueue_iteration(double eps, int L, hipStream_t s) override {
    hipGraphExec_t e = graph_for(L);
    stage_scalars(eps);
    DR_HIP(hipMemcpyAsync(dsc, hsc, sizeof(StepScalars<T>), hipMemcpyHostToDevice, s));
    DR_HIP(hipGraphLaunch(e, s));
  }
  // ------------------------------------------------------------ autotuning
  static bool ct_length(int n) {
    return n == 32 || n == 64 || n == 128 || n == 160 || n == 256 || n == 320 || n == 512;
  }
  void set_tuning(const int* prm) override {
    for (int i = 0; i < NPASS; ++i) tune_tbs[i] = prm[i] < 0 ? -1 : prm[i];
    for (int a = 0; a < 3; ++a) {
      const int o = prm[NPASS + a] < 0 ? 0 : prm[NPASS + a];
      DR_REQUIRE(o <= 2, DR_EDOMAIN, "radix order must be 0, 1 or 2");
      tune_order[a] = o;
      const int flen = a == 0 ? M[0] / 2 : M[a], tws = a == 0 ? 2 : 1;
      if (std::is_same<T, float>::value) adf[a] = make_axis<float>(flen, twf[a].p, tws, o);
      else add[a] = make_axis<double>(flen, twd[a].p, tws, o);
    }
    drop_graphs();
    stats_valid = false;
  }
  void get_tuning(int* prm) const override {
    for (int i = 0; i < NPASS; ++i) prm[i] = tune_tbs[i];
    for (int a = 0; a < 3; ++a) prm[NPASS + a] = tune_order[a];
  }
  // device ms of a 2-step leapfrog (3 gradients) from the current state,
  // best of 3 after one warm-up; the state A is left as it was
  double time_leapfrog() {
    auto run = [&] {
      clear_slots();
      leapfrog_core(1e-3, 2);
    };
    run();
    double best = 1e30;
    for (int r = 0; r < 3; ++r) {
      DR_HIP(hipEventRecord(ev[2], stream));
      run();
      DR_HIP(hipEventRecord(ev[3], stream));
      DR_HIP(hipEventSynchronize(ev[3]));
      float ms = 0.f;
      DR_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
      best = std::min(best, (double)ms);
    }
    return best;
  }
  // coordinate search: each pass's tile lines over every size the tile and
  // LDS allow, then each axis's radix order where the runtime stages run
  void autotune(int* prm) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "autotune: data not set");
    const bool prof0 = prof;
    prof = false;
    rec = -1;
    int cur[NPASS + 3];
    get_tuning(cur);
    const int Mh = M[0] / 2;
    const int capz = pick_tbs(M[2], M[2], sizeof(cx<T>), false, 1 << 20, kTileStrided);
    const int capy = pick_tbs(M[1], M[1], sizeof(cx<T>), false, 1 << 20, kTileStrided);
    const int capx = pick_tbs(Mh, Mh + 1, sizeof(cx<T>), true, 1 << 20, kTileX);
    const int cap[NPASS] = {capz, capy, capx, capx, capy};
    set_tuning(cur);
    double best = time_leapfrog();
    for (int pid = 0; pid < NPASS; ++pid) {
      int arg = cur[pid];
      for (int v = 0; v <= cap[pid]; ++v) {
        cur[pid] = v;
        set_tuning(cur);
        const double t = time_leapfrog();
        if (t < best) {
          best = t;
          arg = v;
        }
      }
      cur[pid] = arg;
    }
    const int len[3] = {Mh, M[1], M[2]};
    for (int a = 0; a < 3; ++a) {
      if (ct_length(len[a])) continue;
      int arg = cur[NPASS + a];
      for (int o = 0; o <= 2; ++o) {
        cur[NPASS + a] = o;
        set_tuning(cur);
        const double t = time_leapfrog();
        if (t < best) {
          best = t;
          arg = o;
        }
      }
      cur[NPASS + a] = arg;
    }
    set_tuning(cur);
    prof = prof0;
    for (int i = 0; i < NPASS + 3; ++i) prm[i] = cur[i];
  }

  // the eager launches of one iteration (the update's kernels, on the
  // chain's stream, no wait)
  void enqueue_eager(double eps, int L) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "update: data not set");
    DR_REQUIRE(L >= 1, DR_EDOMAIN, "update: L must be >= 1");
    const uint64_t it = iteration++;
    device_iteration(eps, L, it, nullptr, nullptr);
  }
  void update_graph(double eps, int L, double u, dr_update_result* res) override {
    enqueue_iteration(eps, L, stream);
    double h[NSLOT];
    read_sums(h);
    finish_iteration(h, u, res, true, stream);
    DR_HIP(hipStreamSynchronize(stream));
  }

  void leapfrog(double eps, int L, dr_update_result* res) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "leapfrog: data not set");
    DR_REQUIRE(L >= 1, DR_EDOMAIN, "leapfrog: L must be >= 1");
    clear_slots();
    quad(p.p, S_KIN_A, (T)sigma[0]);
    if (compat()) quad(pim.p, S_KINI_A, (T)sigma[0]);
    leapfrog_core(eps, L);
    finalize();
    double h[NSLOT];
    read_sums(h);
    if (prof && L >= 2) accumulate_profile(L - 1);
    const double nN = (double)N;
    accept_swap();
    prior_sum = h[S_PRIOR_B] + (compat() ? h[S_PRIORI_B] : 0.0);
    sse_h = h[S_SSEH_B];
    sse_s = h[S_SSES_B];
    stats_valid = true;
    if (res) {
      res->e0 = -0.5 * (h[S_KIN_A] + (compat() ? h[S_KINI_A] : 0.0)) / nN;
      res->log_prior0 = -0.5 * (h[S_PRIOR_A] + (compat() ? h[S_PRIORI_A] : 0.0)) / nN;
      res->log_lik0 = log_lik(h[S_SSEH_A], h[S_SSES_A]);
      res->log_prior1 = -0.5 * prior_sum / nN;
      res->log_lik1 = log_lik(sse_h, sse_s);
      res->k1 = -0.5 * (h[S_KIN_B] + (compat() ? h[S_KINI_B] : 0.0)) / nN;
      res->H0 = res->e0 + res->log_prior0 + res->log_lik0;
      res->H1 = res->k1 + res->log_prior1 + res->log_lik1;
      res->R = std::exp(res->H1 - res->H0);
      res->mh_rate = std::min(1.0, res->R);
      res->sse_h = sse_h;
      res->sse_s = sse_s;
      res->accepted = 1;
    }
  }

  void quad(const cx<T>* X, int slot, T sig) {
    const int g = (int)std::min<long>(cdiv(Nh, kThreads), 2048);
    note_blocks(slot, g);
    hipLaunchKernelGGL(k_quad<T>, dim3(g), dim3(kThreads), 0, stream, Nh, H0, X, invlam.p, sig, pslot(slot));
    DR_HIP(hipGetLastError());
  }

  void refresh_stats() override {
    if (stats_valid) return;
    clear_slots();
    const long sk = (long)M[1] * H0;
    c2r_region(LoadS<T>{{}, muA.p, 0, sk}, RA.p);
    quad(muA.p, S_PRIOR_A, (T)-1);
    if (compat()) quad(miA.p, S_PRIORI_A, (T)-1);
    if (data_set) {
      const int g = (int)std::min<long>(cdiv(nreg, kThreads), 2048);
      note_blocks(S_SSEH_A, g);
      hipLaunchKernelGGL(k_sse_h<T>, dim3(g), dim3(kThreads), 0, stream, nreg, RA.p, Yg.p, mask.p, pslot(S_SSEH_A));
      DR_HIP(hipGetLastError());
      w_products(RA.p, S_SSES_A);
    }
    finalize();
    double h[NSLOT];
    read_sums(h);
    prior_sum = h[S_PRIOR_A] + (compat() ? h[S_PRIORI_A] : 0.0);
    sse_h = h[S_SSEH_A];
    sse_s = cfg.n_datasets == 2 ? h[S_SSES_A] : 0.0;
    stats_valid = true;
  }

  // _update_tau_sq_inv (MultiResParameters.h:863-872): lambda *= t / tau^-2.
  // lambda and 1/lambda are rescaled from their values at theta[0], so
  // repeated updates do not accumulate rounding.
  void set_tau_sq_inv(double t) override {
    DR_REQUIRE(t > 0.0 && std::isfinite(t), DR_EDOMAIN, "tau^-2 must be positive and finite");
    if (!invlam0.p) {
      invlam0.alloc(Nh);
      lam0.alloc(Nh);
      DR_HIP(hipMemcpyAsync(invlam0.p, invlam.p, Nh * sizeof(T), hipMemcpyDeviceToDevice, stream));
      DR_HIP(hipMemcpyAsync(lam0.p, lam.p, Nh * sizeof(double), hipMemcpyDeviceToDevice, stream));
    }
    const double t0 = 1.0 / cfg.theta[0];
    hipLaunchKernelGGL(k_rescale_spectrum<T>, dim3(cdiv(Nh, kThreads)), dim3(kThreads), 0, stream, (long)Nh,
                       invlam0.p, lam0.p, invlam.p, lam.p, (T)(t0 / t), t / t0);
    DR_HIP(hipGetLastError());
    tau_sq_inv = t;
    stats_valid = false;
  }
  // _real_mu.squaredNorm() of the current state (:866-867): sum of Re(mu)^2
  // over the data voxels, from the region field of state A
  double mu_data_sumsq() override {
    refresh_stats();
    clear_slots();
    const int g = (int)std::min<long>(cdiv(nreg, kThreads), 2048);
    note_blocks(S_TMP, g);
    hipLaunchKernelGGL(k_sumsq_masked<T>, dim3(g), dim3(kThreads), 0, stream, (long)nreg, RA.p, mask.p,
                       pslot(S_TMP));
    DR_HIP(hipGetLastError());
    finalize();
    double h[NSLOT];
    read_sums(h);
    return h[S_TMP];
  }

  // ------------------------------------------------------------ test hooks
  void upload_real(const double* src, DevBuf<T>& dst) {
    if (dst.n != (size_t)N) dst.alloc(N);
    std::vector<T> h(N);
    for (long i = 0; i < N; ++i) h[i] = (T)src[i];
    DR_HIP(hipMemcpyAsync(dst.p, h.data(), N * sizeof(T), hipMemcpyHostToDevice, stream));
    DR_HIP(hipStreamSynchronize(stream));
  }
  void download_real(const DevBuf<T>& src, double* dst) {
    std::vector<T> h(N);
    DR_HIP(hipMemcpyAsync(h.data(), src.p, N * sizeof(T), hipMemcpyDeviceToHost, stream));
    DR_HIP(hipStreamSynchronize(stream));
    for (long i = 0; i < N; ++i) dst[i] = (double)h[i];
  }
  void set_state(const double* mr, const double* mi, const double* pr, const double* pi) override {
    if (mr) { upload_real(mr, Rf); r2c_full(Rf.p, muA.p); }
    if (pr) { upload_real(pr, Rf); r2c_full(Rf.p, p.p); }
    if (compat()) {
      if (mi) { upload_real(mi, Rf); r2c_full(Rf.p, miA.p); }
      if (pi) { upload_real(pi, Rf); r2c_full(Rf.p, pim.p); }
    }
    DR_HIP(hipStreamSynchronize(stream));
    stats_valid = false;
  }
  void get_state(double* mr, double* mi, double* pr, double* pi) override {
    if (Rf.n != (size_t)N) Rf.alloc(N);
    if (mr) { c2r_full(muA.p, Rf.p); download_real(Rf, mr); }
    if (pr) { c2r_full(p.p, Rf.p); download_real(Rf, pr); }
    if (mi) {
      if (compat()) { c2r_full(miA.p, Rf.p); download_real(Rf, mi); }
      else std::fill(mi, mi + N, 0.0);
    }
    if (pi) {
      if (compat()) { c2r_full(pim.p, Rf.p); download_real(Rf, pi); }
      else std::fill(pi, pi + N, 0.0);
    }
  }
  void gradient(double* gr, double* gi) override {
    DR_REQUIRE(data_set, DR_ELOGIC, "gradient: data not set");
    if (Rf.n != (size_t)N) Rf.alloc(N);
    const long sk = (long)M[1] * H0;
    c2r_region(LoadS<T>{{}, muA.p, 0, sk}, RA.p);
    w_products(RA.p, -1);
    r2c_region(residual(RA.p, -1), EpiGrad<T>{{}, muA.p, muB.p, invlam.p, sk}, -1, -1);
    c2r_full(muB.p, Rf.p);
    download_real(Rf, gr);
    if (gi) {
      if (compat()) {
        DR_HIP(hipMemcpyAsync(muB.p, miA.p, Nh * sizeof(cx<T>), hipMemcpyDeviceToDevice, stream));
        hipLaunchKernelGGL(k_scale_spec<T>, dim3(cdiv(Nh, kThreads)), dim3(kThreads), 0, stream, Nh, muB.p, invlam.p,
                           (T)-1);
        DR_HIP(hipGetLastError());
        c2r_full(muB.p, Rf.p);
        download_real(Rf, gi);
      } else {
        std::fill(gi, gi + N, 0.0);
      }
    }
  }

  void accumulate_profile(long n) {
    float ms = 0.f;
    DR_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    prof_ms += ms;
    prof_n += n;
    for (int i = 0; i + 1 < nrec; ++i) {
      float m = 0.f;
      DR_HIP(hipEventElapsedTime(&m, pev[i], pev[i + 1]));
      prof_k[i] += m;
    }
    if (nrec > 1) ++prof_kn;
  }
  long prof_kn = 0;
  void reset_kernel_profile() override {
    for (auto& v : prof_k) v = 0.0;
    prof_kn = 0;
    nrec = 0;
  }
  void kernel_profile(double* ms, int max_k, int* n_k, int64_t* n_steps) override {
    const int nk = std::max(0, nrec - 1);
    *n_k = nk;
    *n_steps = prof_kn;
    for (int i = 0; i < nk && i < max_k; ++i) ms[i] = prof_kn ? prof_k[i] / (double)prof_kn : 0.0;
  }
};


}  // namespace drh

// ---------------------------------------------------------------- C ABI
struct dr_chain {
  std::unique_ptr<drh::ChainBase> impl;
};

using drh::Error;
using drh::guarded;

extern "C" {

const char* dr_last_error(void) { return drh::g_last_error.c_str(); }
int dr_version(void) { return 10000; }

int dr_device_count(int* n) {
  return guarded([&] { DR_HIP(hipGetDeviceCount(n)); });
}

// torus dims: the reference rule 2^ceil(log2(2 max_ijk)) (MultiResParameters.h:191-207)
// or the caller's even embedding
static void embedding_dims(const dr_config* cfg, const int mx[3], int M[3]) {
  const bool rule = cfg->embed[0] == 0 && cfg->embed[1] == 0 && cfg->embed[2] == 0;
  for (int a = 0; a < 3; ++a) {
    if (rule) {
      const int d = 2 * mx[a];
      DR_REQUIRE(d >= 2, DR_ELOGIC, "circulant_base_3d: Cannot have dimensions <= 1");
      M[a] = (int)std::pow(2.0, std::ceil(std::log((double)d) / std::log(2.0)));
    } else {
      M[a] = cfg->embed[a];
      DR_REQUIRE(M[a] >= 2 && M[a] % 2 == 0 && M[a] >= 2 * mx[a], DR_EDOMAIN,
                 "embedding must be even and >= 2*max_ijk along every axis");
    }
    DR_REQUIRE(M[a] <= 2048, DR_EDOMAIN, "embedding axis longer than 2048");
  }
}

int dr_estimate_bytes(const dr_config* cfg, const int32_t* ijk, int64_t n_h, int64_t n_s, int64_t w_nnz,
                      int64_t* bytes) {
  return guarded([&] {
    DR_REQUIRE(cfg && ijk && bytes && n_h > 0, DR_ELOGIC, "dr_estimate_bytes: null argument or empty data");
    int mx[3] = {0, 0, 0};
    for (int64_t i = 0; i < n_h; ++i)
      for (int a = 0; a < 3; ++a) mx[a] = std::max(mx[a], ijk[3 * i + a]);
    int M[3];
    embedding_dims(cfg, mx, M);
    const int64_t s = cfg->precision == DR_FP64 ? 8 : 4, cb = 2 * s;
    const int64_t H0 = M[0] / 2 + 1, Nh = H0 * M[1] * M[2];
    const int64_t nreg = (int64_t)(mx[0] + 1) * (mx[1] + 1) * (mx[2] + 1);
    int64_t b = Nh * (5 * cb + s + 8);                                   // mu A/B, p, T1, T2, 1/lambda, lambda
    if (cfg->state_mode == DR_STATE_COMPAT) b += 3 * Nh * cb;            // imaginary chain
    int64_t ps = std::max<int64_t>({1024, Nh / 8 + 1, (int64_t)M[1] * M[2] + 1, (int64_t)M[1] * H0 + 1});
    if (n_s > 0) ps = std::max<int64_t>(ps, (n_s + 3) / 4 + n_s + 2);
    b += 16 * ps * 8;                                                    // reduction partials
    b += nreg * (4 * s + 1) + n_h * (8 + s + 24);                        // region fields, indices, summary
    if (n_s > 0) b += 2 * w_nnz * (4 + s) + n_s * (8 + 2 * s) + nreg * s; // W (CSR bound), Ys, v
    b += 2 * Nh * 16;                                                    // spectrum set-up scratch
    *bytes = b;
  });
}

int dr_create(const dr_config* cfg, const int32_t* ijk, int64_t n_h, dr_chain** out) {
  return guarded([&] {
    DR_REQUIRE(cfg && ijk && out && n_h > 0, DR_ELOGIC, "dr_create: null argument or empty data");
    DR_REQUIRE(cfg->n_datasets == 1 || cfg->n_datasets == 2, DR_ELOGIC,
               "Gradient only implemented for single or dual resolution");
    DR_REQUIRE(cfg->precision == DR_FP64 || cfg->precision == DR_FP32, DR_EDOMAIN, "unknown precision");
    DR_REQUIRE(cfg->covfun == DR_COV_RBF || cfg->covfun == DR_COV_RQ, DR_EDOMAIN, "Unknown covariance function");
    if (cfg->covfun == DR_COV_RQ) {
      DR_REQUIRE(cfg->theta[1] > 0, DR_EDOMAIN, "rational_quadratic: psi <= 0");
      DR_REQUIRE(cfg->theta[2] > 0, DR_EDOMAIN, "rational_quadratic: nu <= 0");
      DR_REQUIRE(cfg->theta[0] > 0, DR_EDOMAIN, "rational_quadratic: variance <= 0");
    }
    std::unique_ptr<drh::ChainBase> c;
    if (cfg->precision == DR_FP64) c.reset(new drh::Chain<double>());
    else c.reset(new drh::Chain<float>());
    c->cfg = *cfg;
    c->n_h = n_h;
    c->ijk.assign(ijk, ijk + 3 * n_h);
    int mx[3] = {0, 0, 0};
    for (int64_t i = 0; i < n_h; ++i)
      for (int a = 0; a < 3; ++a) {
        DR_REQUIRE(ijk[3 * i + a] >= 0, DR_EDOMAIN, "negative data index");
        mx[a] = std::max(mx[a], ijk[3 * i + a]);
      }
    embedding_dims(cfg, mx, c->M);
    c->build();
    *out = new dr_chain{std::move(c)};
  });
}

int dr_destroy(dr_chain* c) {
  return guarded([&] { delete c; });
}

int dr_spectrum_report(dr_chain* c, int dims[3], int64_t* N, int64_t* n_nonpositive, double* max_eig) {
  return guarded([&] {
    DR_REQUIRE(c, DR_ELOGIC, "null chain");
    for (int a = 0; a < 3; ++a) dims[a] = c->impl->M[a];
    *N = c->impl->N;
    *n_nonpositive = c->impl->n_nonpos;
    *max_eig = c->impl->max_eig;
  });
}

int dr_get_lambda(dr_chain* c, double* out) {
  return guarded([&] { c->impl->get_lambda(out); });
}

int dr_set_data(dr_chain* c, const double* Yh, int64_t n_h, const double* Ys, int64_t n_s, const int64_t* W_rowptr,
                const int32_t* W_cols, const double* W_vals) {
  return guarded([&] {
    auto& k = *c->impl;
    DR_REQUIRE(n_h == k.n_h, DR_ELOGIC, "dr_set_data: Yh length differs from the number of data indices");
    k.Yh.assign(Yh, Yh + n_h);
    if (k.cfg.n_datasets == 2) {
      DR_REQUIRE(Ys && n_s > 0 && W_rowptr && W_cols && W_vals, DR_ELOGIC, "Insufficient data for parameters");
      k.n_s = n_s;
      k.Ys.assign(Ys, Ys + n_s);
      k.W_rp.assign(W_rowptr, W_rowptr + n_s + 1);
      DR_REQUIRE(k.W_rp[0] == 0, DR_EDOMAIN, "W row pointer must start at 0");
      for (int64_t i = 0; i < n_s; ++i) DR_REQUIRE(k.W_rp[i + 1] >= k.W_rp[i], DR_EDOMAIN, "W row pointer decreasing");
      const int64_t nnz = k.W_rp[n_s];
      k.W_cols.assign(W_cols, W_cols + nnz);
      k.W_vals.assign(W_vals, W_vals + nnz);
    }
    k.upload_data();
  });
}

int dr_w_layout(dr_chain* c, int force_csr) {
  return guarded([&] { c->impl->force_csr = force_csr != 0; });
}

int dr_w_stats(dr_chain* c, int64_t* stats) {
  return guarded([&] {
    auto& k = *c->impl;
    stats[0] = k.use_cls ? 1 : 0;
    for (int i = 0; i < 10; ++i) stats[1 + i] = k.wstats[i];
  });
}

int dr_add_data_to_mu(dr_chain* c) {
  return guarded([&] { c->impl->add_data_to_mu(); });
}

int dr_set_output_indices(dr_chain* c, const int32_t* ijk, int64_t n_out) {
  return guarded([&] {
    auto& k = *c->impl;
    int mx[3] = {0, 0, 0};
    for (int64_t i = 0; i < n_out; ++i)
      for (int a = 0; a < 3; ++a) {
        if (ijk[3 * i + a] < 0) throw Error(DR_EDOMAIN, "MultiResParameters:set_output_indices MIN");
        if (ijk[3 * i + a] > k.M[a]) throw Error(DR_EDOMAIN, "MultiResParameters:set_output_indices MAX");
        DR_REQUIRE(ijk[3 * i + a] < k.M[a], DR_EDOMAIN, "output index on the torus boundary is not supported");
        mx[a] = std::max(mx[a], ijk[3 * i + a]);
      }
    k.out_ijk.assign(ijk, ijk + 3 * n_out);
    int r[3];
    for (int a = 0; a < 3; ++a) r[a] = std::max(k.reg[a], mx[a] + 1);
    k.set_region(r);
    k.set_outputs();
  });
}

int dr_get_mu_output(dr_chain* c, double* out, int64_t n_out) {
  return guarded([&] { c->impl->get_mu_out(out, (long)n_out); });
}

int dr_summary_reset(dr_chain* c) {
  return guarded([&] { c->impl->summary_reset(); });
}

int dr_summary_update(dr_chain* c) {
  return guarded([&] { c->impl->summary_update(); });
}

int dr_summary_read(dr_chain* c, double* sum_mu, double* sum_mu2, double* sum_abs, int64_t n_out, int64_t* n_updates) {
  return guarded([&] { c->impl->summary_read(sum_mu, sum_mu2, sum_abs, (long)n_out, n_updates); });
}

int dr_get_mu_data(dr_chain* c, double* out, int64_t n_h) {
  return guarded([&] {
    auto& k = *c->impl;
    DR_REQUIRE(n_h == k.n_h, DR_ELOGIC, "dr_get_mu_data: wrong length");
    std::vector<int32_t> g(k.n_h);
    for (int64_t i = 0; i < k.n_h; ++i) g[i] = (int32_t)k.region_index(k.ijk[3 * i], k.ijk[3 * i + 1], k.ijk[3 * i + 2]);
    k.get_mu_region_gather(g, out);
  });
}

int dr_set_sigma_sq_inv(dr_chain* c, const double* s) {
  return guarded([&] {
    auto& k = *c->impl;
    for (int i = 0; i < k.cfg.n_datasets; ++i) k.sigma[i] = s[i];
  });
}
int dr_get_sigma_sq_inv(dr_chain* c, double* s) {
  return guarded([&] {
    auto& k = *c->impl;
    for (int i = 0; i < k.cfg.n_datasets; ++i) s[i] = k.sigma[i];
  });
}
int dr_set_tau_sq_inv(dr_chain* c, double tau_sq_inv) {
  return guarded([&] { c->impl->set_tau_sq_inv(tau_sq_inv); });
}
int dr_get_tau_sq_inv(dr_chain* c, double* tau_sq_inv) {
  return guarded([&] { *tau_sq_inv = c->impl->tau_sq_inv; });
}
int dr_mu_data_sumsq(dr_chain* c, double* out) {
  return guarded([&] {
    auto& k = *c->impl;
    DR_REQUIRE(k.data_set, DR_ELOGIC, "dr_mu_data_sumsq: data not set");
    *out = k.mu_data_sumsq();
  });
}

int dr_update(dr_chain* c, double eps, int L, double u_accept, dr_update_result* res) {
  return guarded([&] { c->impl->update(eps, L, u_accept, nullptr, nullptr, res); });
}
int dr_update_with_noise(dr_chain* c, double eps, int L, double u_accept, const double* z_re, const double* z_im,
                         dr_update_result* res) {
  return guarded([&] {
    DR_REQUIRE(z_re && (z_im || !c->impl->compat()), DR_ELOGIC, "dr_update_with_noise: null noise");
    c->impl->update(eps, L, u_accept, z_re, z_im, res);
  });
}

// tuning cache: {"key": [8 ints], ...}, one entry per line
static std::map<std::string, std::vector<int>> read_tune_cache(const std::string& path) {
  std::map<std::string, std::vector<int>> m;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t q1 = line.find('"');
    if (q1 == std::string::npos) continue;
    const size_t q2 = line.find('"', q1 + 1);
    if (q2 == std::string::npos) continue;
    const size_t b1 = line.find('[', q2), b2 = line.find(']', b1 == std::string::npos ? q2 : b1);
    if (b1 == std::string::npos || b2 == std::string::npos) continue;
    std::vector<int> v;
    std::stringstream ss(line.substr(b1 + 1, b2 - b1 - 1));
    std::string tok;
    try {
      while (std::getline(ss, tok, ',')) v.push_back(std::stoi(tok));
    } catch (...) {
      continue;
    }
    if (v.size() == 8) m[line.substr(q1 + 1, q2 - q1 - 1)] = v;
  }
  return m;
}
static void write_tune_cache(const std::string& path, const std::map<std::string, std::vector<int>>& m) {
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  {
    std::ofstream f(tmp);
    DR_REQUIRE(f.is_open(), DR_ERUNTIME, "autotune: cannot write the cache " + path);
    f << "{\n";
    size_t k = 0;
    for (const auto& kv : m) {
      f << "  \"" << kv.first << "\": [";
      for (size_t i = 0; i < kv.second.size(); ++i) f << (i ? ", " : "") << kv.second[i];
      f << "]" << (++k < m.size() ? "," : "") << "\n";
    }
    f << "}\n";
  }
  DR_REQUIRE(std::rename(tmp.c_str(), path.c_str()) == 0, DR_ERUNTIME, "autotune: cannot replace the cache " + path);
}
// cache key: device architecture, precision, state mode, datasets, torus
// and image region
static std::string tune_key(const drh::ChainBase& c) {
  hipDeviceProp_t prop;
  DR_HIP(hipGetDeviceProperties(&prop, c.cfg.device));
  std::string arch = prop.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  std::ostringstream k;
  k << arch << "|" << (c.cfg.precision == DR_FP64 ? "fp64" : "fp32") << "|"
    << (c.cfg.state_mode == DR_STATE_COMPAT ? "compat" : "real") << "|" << c.cfg.n_datasets << "|" << c.M[0] << "x"
    << c.M[1] << "x" << c.M[2] << "|" << c.reg[0] << "x" << c.reg[1] << "x" << c.reg[2];
  return k.str();
}

int dr_autotune(dr_chain* c, const char* cache_path, int* params, int* from_cache) {
  return guarded([&] {
    DR_REQUIRE(c && params, DR_ELOGIC, "dr_autotune: null argument");
    auto& k = *c->impl;
    DR_HIP(hipSetDevice(k.cfg.device));
    DR_HIP(hipStreamSynchronize(k.stream));
    int prm[8];
    bool hit = false;
    std::string key;
    if (cache_path && *cache_path) {
      key = tune_key(k);
      const auto m = read_tune_cache(cache_path);
      const auto f = m.find(key);
      if (f != m.end()) {
        std::copy(f->second.begin(), f->second.end(), prm);
        k.set_tuning(prm);
        hit = true;
      }
    }
    if (!hit) {
      k.autotune(prm);
      if (cache_path && *cache_path) {
        auto m = read_tune_cache(cache_path);
        m[key] = std::vector<int>(prm, prm + 8);
        write_tune_cache(cache_path, m);
      }
    }
    std::copy(prm, prm + 8, params);
    if (from_cache) *from_cache = hit ? 1 : 0;
  });
}

int dr_set_tuning(dr_chain* c, const int* params) {
  return guarded([&] {
    DR_REQUIRE(c && params, DR_ELOGIC, "dr_set_tuning: null argument");
    DR_HIP(hipStreamSynchronize(c->impl->stream));
    c->impl->set_tuning(params);
  });
}

int dr_get_tuning(dr_chain* c, int* params) {
  return guarded([&] {
    DR_REQUIRE(c && params, DR_ELOGIC, "dr_get_tuning: null argument");
    c->impl->get_tuning(params);
  });
}

int dr_update_captured(dr_chain* c, double eps, int L, double u_accept, dr_update_result* res) {
  return guarded([&] {
    DR_REQUIRE(c, DR_ELOGIC, "null chain");
    c->impl->update_graph(eps, L, u_accept, res);
  });
}

int dr_log_prior(dr_chain* c, double* out) {
  return guarded([&] {
    auto& k = *c->impl;
    k.refresh_stats();
    *out = -0.5 * k.prior_sum / (double)k.N;
  });
}
int dr_log_likelihood(dr_chain* c, double* out) {
  return guarded([&] {
    auto& k = *c->impl;
    DR_REQUIRE(k.data_set, DR_ELOGIC, "Insufficient data for parameters");
    k.refresh_stats();
    *out = k.log_lik(k.sse_h, k.sse_s);
  });
}
int dr_log_posterior(dr_chain* c, double* out) {
  return guarded([&] {
    auto& k = *c->impl;
    DR_REQUIRE(k.data_set, DR_ELOGIC, "Insufficient data for parameters");
    k.refresh_stats();
    *out = -0.5 * k.prior_sum / (double)k.N + k.log_lik(k.sse_h, k.sse_s);
  });
}
int dr_sse(dr_chain* c, double* sse_h, double* sse_s) {
  return guarded([&] {
    auto& k = *c->impl;
    k.refresh_stats();
    *sse_h = k.sse_h;
    *sse_s = k.sse_s;
  });
}

int dr_set_state(dr_chain* c, const double* mu_re, const double* mu_im, const double* p_re, const double* p_im) {
  return guarded([&] { c->impl->set_state(mu_re, mu_im, p_re, p_im); });
}
int dr_get_state(dr_chain* c, double* mu_re, double* mu_im, double* p_re, double* p_im) {
  return guarded([&] { c->impl->get_state(mu_re, mu_im, p_re, p_im); });
}
int dr_gradient(dr_chain* c, double* g_re, double* g_im) {
  return guarded([&] { c->impl->gradient(g_re, g_im); });
}
int dr_leapfrog(dr_chain* c, double eps, int L, dr_update_result* res) {
  return guarded([&] { c->impl->leapfrog(eps, L, res); });
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Many chains advanced together, one HMC iteration per dr_batch_update.  In
// DR_BATCH_GRAPH mode every chain's iteration (momentum draw, L leapfrog
// steps, energy sums) is captured as its own graph and the graphs become
// parallel branches of one graph, launched once per iteration on the batch
// stream after one copy of all chains' scalars; one copy returns all energy
// sums.  The graphs depend on each chain's L, so a new L vector captures and
// instantiates anew.  DR_BATCH_STREAMS launches each chain's own graph on
// its own stream, DR_BATCH_EAGER the plain kernel launches.  A chain's
// kernels, scalars and summation order are the same in every mode and alone,
// so its results do not depend on the other chains.
struct dr_batch {
  std::vector<dr_chain*> chains;
  int mode = DR_BATCH_GRAPH;
  int device = 0;
  hipStream_t stream = nullptr;
  void* dsc = nullptr;
  void* hsc = nullptr;
  double* dsums = nullptr;
  double* hsums = nullptr;
  std::vector<int> cur_L;
  std::vector<uint64_t> cur_gen;
  hipGraphExec_t exec = nullptr;
  static constexpr size_t kSlot = 64;  // bytes of one chain's scalars (StepScalars<double> = 48)

  ~dr_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (dr_chain* c : chains) {
      try {
        c->impl->set_batch_slots(nullptr, nullptr, nullptr);
      } catch (...) {
      }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (dsc) (void)hipFree(dsc);
    if (dsums) (void)hipFree(dsums);
    if (hsc) (void)hipHostFree(hsc);
    if (hsums) (void)hipHostFree(hsums);
    if (stream) (void)hipStreamDestroy(stream);
  }
  double* host_sums(size_t i) { return hsums + drh::NSLOT * i; }

  void rebuild(const int* L) {
    if (exec) {
      (void)hipGraphExecDestroy(exec);
      exec = nullptr;
    }
    hipGraph_t parent = nullptr;
    DR_HIP(hipGraphCreate(&parent, 0));
    try {
      for (size_t i = 0; i < chains.size(); ++i) {
        hipGraph_t g = chains[i]->impl->capture_iteration(L[i]);
        hipGraphNode_t node;
        const hipError_t e = hipGraphAddChildGraphNode(&node, parent, nullptr, 0, g);
        (void)hipGraphDestroy(g);
        DR_HIP(e);
      }
      DR_HIP(hipGraphInstantiate(&exec, parent, nullptr, nullptr, 0));
    } catch (...) {
      (void)hipGraphDestroy(parent);
      throw;
    }
    (void)hipGraphDestroy(parent);
    cur_L.assign(L, L + chains.size());
    cur_gen.clear();
    for (dr_chain* c : chains) cur_gen.push_back(c->impl->layout_gen);
  }
  bool stale(const int* L) const {
    if (!exec || !std::equal(cur_L.begin(), cur_L.end(), L)) return true;
    for (size_t i = 0; i < chains.size(); ++i)
      if (chains[i]->impl->layout_gen != cur_gen[i]) return true;
    return false;
  }

  void update(const double* eps, const int* L, const double* u, dr_update_result* res) {
    const size_t n = chains.size();
    for (size_t i = 0; i < n; ++i) {
      DR_REQUIRE(L[i] >= 1, DR_EDOMAIN, "update: L must be >= 1");
      DR_REQUIRE(chains[i]->impl->data_set, DR_ELOGIC, "update: data not set");
      DR_HIP(hipStreamSynchronize(chains[i]->impl->stream));  // earlier work on the chain's own stream
    }
    if (mode == DR_BATCH_GRAPH) {
      if (stale(L)) rebuild(L);
      for (size_t i = 0; i < n; ++i) chains[i]->impl->stage_scalars(eps[i]);
      DR_HIP(hipMemcpyAsync(dsc, hsc, n * kSlot, hipMemcpyHostToDevice, stream));
      DR_HIP(hipGraphLaunch(exec, stream));
      DR_HIP(hipMemcpyAsync(hsums, dsums, n * drh::NSLOT * sizeof(double), hipMemcpyDeviceToHost, stream));
      DR_HIP(hipStreamSynchronize(stream));
      for (size_t i = 0; i < n; ++i) chains[i]->impl->finish_iteration(host_sums(i), u[i], &res[i], true, stream);
      DR_HIP(hipStreamSynchronize(stream));
      return;
    }
    for (size_t i = 0; i < n; ++i) {
      auto& k = *chains[i]->impl;
      if (mode == DR_BATCH_STREAMS) k.enqueue_iteration(eps[i], L[i], k.stream);
      else k.enqueue_eager(eps[i], L[i]);
      DR_HIP(hipMemcpyAsync(host_sums(i), dsums + drh::NSLOT * i, drh::NSLOT * sizeof(double),
                            hipMemcpyDeviceToHost, k.stream));
    }
    for (size_t i = 0; i < n; ++i) {
      auto& k = *chains[i]->impl;
      DR_HIP(hipStreamSynchronize(k.stream));
      k.finish_iteration(host_sums(i), u[i], &res[i], mode == DR_BATCH_STREAMS, k.stream);
    }
    for (size_t i = 0; i < n; ++i) DR_HIP(hipStreamSynchronize(chains[i]->impl->stream));
  }
};

extern "C" {

int dr_batch_create(dr_chain* const* chains, int n, int mode, dr_batch** out) {
  return guarded([&] {
    DR_REQUIRE(chains && out && n >= 1, DR_ELOGIC, "dr_batch_create: null argument or empty batch");
    DR_REQUIRE(mode == DR_BATCH_GRAPH || mode == DR_BATCH_STREAMS || mode == DR_BATCH_EAGER, DR_EDOMAIN,
               "dr_batch_create: unknown mode");
    std::unique_ptr<dr_batch> b(new dr_batch());
    b->mode = mode;
    b->device = chains[0]->impl->cfg.device;
    for (int i = 0; i < n; ++i) {
      DR_REQUIRE(chains[i], DR_ELOGIC, "dr_batch_create: null chain");
      DR_REQUIRE(chains[i]->impl->cfg.device == b->device, DR_EDOMAIN, "dr_batch_create: chains on different devices");
      for (int j = 0; j < i; ++j) DR_REQUIRE(chains[j] != chains[i], DR_ELOGIC, "dr_batch_create: chain listed twice");
    }
    DR_HIP(hipSetDevice(b->device));
    DR_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    DR_HIP(hipMalloc(&b->dsc, n * dr_batch::kSlot));
    DR_HIP(hipMalloc((void**)&b->dsums, (size_t)n * drh::NSLOT * sizeof(double)));
    DR_HIP(hipHostMalloc(&b->hsc, n * dr_batch::kSlot, hipHostMallocDefault));
    DR_HIP(hipHostMalloc((void**)&b->hsums, (size_t)n * drh::NSLOT * sizeof(double), hipHostMallocDefault));
    for (int i = 0; i < n; ++i) {
      b->chains.push_back(chains[i]);
      chains[i]->impl->set_batch_slots(static_cast<char*>(b->dsc) + i * dr_batch::kSlot,
                                       static_cast<char*>(b->hsc) + i * dr_batch::kSlot,
                                       b->dsums + (size_t)drh::NSLOT * i);
    }
    *out = b.release();
  });
}

int dr_batch_update(dr_batch* b, const double* eps, const int* L, const double* u_accept, dr_update_result* res) {
  return guarded([&] {
    DR_REQUIRE(b && eps && L && u_accept && res, DR_ELOGIC, "dr_batch_update: null argument");
    DR_HIP(hipSetDevice(b->device));
    b->update(eps, L, u_accept, res);
  });
}

int dr_batch_destroy(dr_batch* b) {
  return guarded([&] { delete b; });
}

int dr_profile_enable(dr_chain* c, int enable) {
  return guarded([&] {
    auto& k = *c->impl;
    k.prof = enable != 0;
    k.prof_ms = 0.0;
    k.prof_n = 0;
    k.reset_kernel_profile();
  });
}

int dr_profile_kernels(dr_chain* c, double* ms, int max_k, int* n_k, int64_t* n_steps) {
  return guarded([&] { c->impl->kernel_profile(ms, max_k, n_k, n_steps); });
}
int dr_profile_read(dr_chain* c, double* ms_per_grad, int64_t* n_grads, double* ms_fft_passes, double* ms_w) {
  return guarded([&] {
    auto& k = *c->impl;
    *ms_per_grad = k.prof_n ? k.prof_ms / (double)k.prof_n : 0.0;
    *n_grads = k.prof_n;
    if (ms_fft_passes) *ms_fft_passes = 0.0;
    if (ms_w) *ms_w = 0.0;
  });
}

}  // extern "C"
