// Minimum-contrast summary data for estimate_rbf on the GPU
// (estimate_kernel_parameters.h:53-101, perturbation_matrix.h:62-106).
//
// For every voxel offset p the reference sums, over voxel pairs (u, u+p)
// that are both present (not 0, not NaN) and inside the image:
//   Sxy(p) = sum x(u) x(u+p),  Sx(p) = sum x(u),  Sy(p) = sum x(u+p),  n(p).
// With x the image zeroed where missing and m its presence indicator these
// are the cross-correlations x*x, x*m, m*x, m*m.  All four come from two
// forward FFTs (x, m) and four inverse FFTs of conj(A) B on a zero-padded
// torus larger than the image by the largest offset (no wrap-around), in
// float64, using the same pass kernels as the sampler.  The per-offset
// covariance (Sxy - Sx Sy / n) / (n - 1) is formed on the host, as in the
// reference, and pair counts are the rounded m*m correlation (exact).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "dr_kernels.hpp"

namespace drh {

namespace {

template <typename T>
struct MBuf {
  T* p = nullptr;
  MBuf() = default;
  explicit MBuf(size_t n) { alloc(n); }
  MBuf(const MBuf&) = delete;
  ~MBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw Error(DR_ENOMEM, "hipMalloc failed in the minimum-contrast summary");
    }
  }
};

// masked image value (ind = 0) or presence indicator (ind = 1) on the padded torus
struct LoadMasked : NoRed {
  const double* img; int nx, ny, nz, M1, ind;
  __device__ __forceinline__ double one(int x, int y, int z) const {
    if (x >= nx || y >= ny || z >= nz) return 0.0;
    const double v = img[x + (long)nx * (y + (long)ny * z)];
    if (isnan(v) || v == 0.0) return 0.0;
    return ind ? 1.0 : v;
  }
  __device__ __forceinline__ cx<double> load2(int row, int n) const {
    const int y = row % M1, z = row / M1;
    return mk<double>(one(2 * n, y, z), one(2 * n + 1, y, z));
  }
};

// P_ab = conj(A) B for (x,x), (x,m), (m,x), (m,m)
__global__ void k_corr_products(long Nh, const cx<double>* __restrict__ X, const cx<double>* __restrict__ Mk,
                                cx<double>* __restrict__ P0, cx<double>* __restrict__ P1, cx<double>* __restrict__ P2,
                                cx<double>* __restrict__ P3) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nh) return;
  const cx<double> x = X[i], m = Mk[i];
  P0[i] = cmulc(x, x);
  P1[i] = cmulc(m, x);
  P2[i] = cmulc(x, m);
  P3[i] = cmulc(m, m);
}

// read the four correlation grids at the requested offsets (mod M)
__global__ void k_gather_offsets(int n, const int* __restrict__ off, int M0, int M1, int M2,
                                 const double* __restrict__ R0, const double* __restrict__ R1,
                                 const double* __restrict__ R2, const double* __restrict__ R3,
                                 double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = ((off[3 * i] % M0) + M0) % M0;
  const int y = ((off[3 * i + 1] % M1) + M1) % M1;
  const int z = ((off[3 * i + 2] % M2) + M2) % M2;
  const long g = x + (long)M0 * (y + (long)M1 * z);
  out[4 * i] = R0[g];
  out[4 * i + 1] = R1[g];
  out[4 * i + 2] = R2[g];
  out[4 * i + 3] = R3[g];
}

int smooth_even(int n) {
  for (int m = std::max(2, n);; ++m) {
    if (m % 2) continue;
    int r = m;
    for (int f : {2, 3, 5, 7})
      while (r % f == 0) r /= f;
    if (r == 1) return m;
  }
}

// float64 3D transforms on one torus (full, unpruned).  Buffers (twiddle
// tables, pass scratch T1/T2) are borrowed: plan() only records them.
struct Fft3d {
  int M[3] = {0, 0, 0}, H0 = 0;
  long N = 0, Nh = 0;
  hipStream_t s = nullptr;
  AxisFFT<double> ax[3]{};
  cx<double>* T1 = nullptr;
  cx<double>* T2 = nullptr;

  static void sizes(const int m[3], long* N, long* Nh) {
    *N = (long)m[0] * m[1] * m[2];
    *Nh = (long)(m[0] / 2 + 1) * m[1] * m[2];
  }
  void plan(const int m[3], hipStream_t st, cx<double>* const tw[3], cx<double>* t1, cx<double>* t2) {
    s = st;
    for (int a = 0; a < 3; ++a) M[a] = m[a];
    H0 = M[0] / 2 + 1;
    sizes(M, &N, &Nh);
    for (int a = 0; a < 3; ++a) ax[a] = make_axis<double>(a == 0 ? M[0] / 2 : M[a], tw[a], a == 0 ? 2 : 1);
    T1 = t1;
    T2 = t2;
  }
  template <bool INV, class Pro, class Epi>
  void strided(int a, int n_outer, int NB, Pro pro, Epi epi) {
    const AxisFFT<double>& ad = ax[a];
    const int tbs = pick_tbs(ad.M, ad.M, sizeof(cx<double>), false, NB, kTileStrided);
    const int TB = 1 << tbs;
    const size_t lds = (size_t)ad.M * TB * sizeof(cx<double>);
    hipLaunchKernelGGL((k_strided<double, INV, Pro, Epi>), dim3(cdiv(NB, TB), n_outer), dim3(kThreads), lds, s, ad,
                       NB, ad.M, ad.M, tbs, pro, epi);
    DR_HIP(hipGetLastError());
  }
  int xtbs(int nrows) const {
    const int Mh = ax[0].M;
    const int lines = std::max(std::max(1, nrows / 2048), cdiv(1024, std::max(Mh, 1)));
    return pick_tbs(Mh, Mh + 1, sizeof(cx<double>), true, lines, kTileX);
  }
  template <class Pro>
  void r2c(Pro pro, cx<double>* X) {
    const int rows = M[1] * M[2], tbs = xtbs(rows), TB = 1 << tbs, Mh = ax[0].M;
    const size_t lds = (size_t)(Mh + 1) * (TB + 1) * sizeof(cx<double>);
    hipLaunchKernelGGL((k_r2c_x<double, Pro, StoreRowC<double>>), dim3(cdiv(rows, TB)), dim3(kThreads), lds, s, ax[0],
                       rows, tbs, pro, StoreRowC<double>{{}, T2, H0});
    DR_HIP(hipGetLastError());
    strided<false>(1, M[2], H0, LoadS<double>{{}, T2, (long)M[1] * H0, H0}, StoreS<double>{{}, T1, (long)M[1] * H0, H0});
    strided<false>(2, 1, M[1] * H0, LoadS<double>{{}, T1, 0, (long)M[1] * H0}, StoreS<double>{{}, X, 0, (long)M[1] * H0});
  }
  void c2r(const cx<double>* X, double* dst) {
    strided<true>(2, 1, M[1] * H0, LoadS<double>{{}, X, 0, (long)M[1] * H0}, StoreS<double>{{}, T1, 0, (long)M[1] * H0});
    strided<true>(1, M[2], H0, LoadS<double>{{}, T1, (long)M[1] * H0, H0}, StoreS<double>{{}, T2, (long)M[1] * H0, H0});
    const int rows = M[1] * M[2], tbs = xtbs(rows), TB = 1 << tbs, Mh = ax[0].M;
    const size_t lds = (size_t)(Mh + 1) * (TB + 1) * sizeof(cx<double>);
    hipLaunchKernelGGL((k_c2r_x<double, LoadRowC<double>, StoreRowR<double>>), dim3(cdiv(rows, TB)), dim3(kThreads),
                       lds, s, ax[0], rows, tbs, LoadRowC<double>{{}, T2, H0},
                       StoreRowR<double>{{}, dst, M[0], 1.0 / (double)N});
    DR_HIP(hipGetLastError());
  }
};

}  // namespace

// unique MCE offsets in first-occurrence order (perturbation_matrix.h:20-106)
std::vector<int> mce_offsets() {
  static const int U[14][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},
                               {1, 0, 1},  {-1, 1, 0}, {1, 0, -1}, {0, 1, 1},  {0, 1, -1},
                               {1, 1, 1},  {1, 1, -1}, {-1, 1, 1}, {-1, 1, -1}};
  const int na = 18, nr = 25;
  std::vector<int> P;
  std::unordered_set<long> seen;
  auto push = [&](int x, int y, int z) {
    const long key = ((long)(x + 64) << 16) | ((long)(y + 64) << 8) | (long)(z + 64);
    if (seen.insert(key).second) P.insert(P.end(), {x, y, z});
  };
  for (int c = 1; c <= na; ++c)
    for (int b = 1; b <= na; ++b)
      for (int a = 1; a <= na; ++a)
        for (int j = 0; j < 14; ++j) push(U[j][0] * a, U[j][1] * b, U[j][2] * c);
  for (int i = 0; i < nr - na; ++i)
    for (int j = 0; j < 3; ++j) {
      int v[3] = {0, 0, 0};
      v[j] = na + i + 1;
      push(v[0], v[1], v[2]);
    }
  return P;
}

// correlation sums (Sxy, Sx, Sy, n) at each offset, float64
static std::vector<double> corr_sums(const double* img, const int dims[3], const std::vector<int>& off, int device) {
  DR_REQUIRE(img && dims, DR_ELOGIC, "minimum contrast: null argument");
  for (int a = 0; a < 3; ++a) DR_REQUIRE(dims[a] >= 1, DR_EDOMAIN, "minimum contrast: empty image");
  int reach[3] = {0, 0, 0};
  const long n = (long)off.size() / 3;
  for (long i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) reach[a] = std::max(reach[a], std::abs(off[3 * i + a]));
  int M[3];
  for (int a = 0; a < 3; ++a) {
    M[a] = smooth_even(dims[a] + reach[a] + 1);
    DR_REQUIRE(M[a] <= 2048, DR_EDOMAIN, "minimum contrast: image axis too long");
  }
  DR_HIP(hipSetDevice(device));
  hipStream_t s;
  DR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct SG { hipStream_t s; ~SG() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } sg{s};
  long N, Nh;
  Fft3d::sizes(M, &N, &Nh);
  MBuf<cx<double>> tw[3], T1(Nh), T2(Nh);
  cx<double>* twp[3];
  for (int a = 0; a < 3; ++a) {
    const auto t = twiddles<double>(M[a]);
    tw[a].alloc(M[a]);
    DR_HIP(hipMemcpyAsync(tw[a].p, t.data(), M[a] * sizeof(cx<double>), hipMemcpyHostToDevice, s));
    DR_HIP(hipStreamSynchronize(s));
    twp[a] = tw[a].p;
  }
  Fft3d f;
  f.plan(M, s, twp, T1.p, T2.p);
  const long nvox = (long)dims[0] * dims[1] * dims[2];
  MBuf<double> d_img(nvox);
  DR_HIP(hipMemcpyAsync(d_img.p, img, nvox * sizeof(double), hipMemcpyHostToDevice, s));
  MBuf<cx<double>> X(f.Nh), Mk(f.Nh), P[4];
  f.r2c(LoadMasked{{}, d_img.p, dims[0], dims[1], dims[2], M[1], 0}, X.p);
  f.r2c(LoadMasked{{}, d_img.p, dims[0], dims[1], dims[2], M[1], 1}, Mk.p);
  for (auto& q : P) q.alloc(f.Nh);
  hipLaunchKernelGGL(k_corr_products, dim3(cdiv(f.Nh, 256)), dim3(256), 0, s, f.Nh, X.p, Mk.p, P[0].p, P[1].p,
                     P[2].p, P[3].p);
  DR_HIP(hipGetLastError());
  MBuf<double> R[4];
  for (int k = 0; k < 4; ++k) {
    R[k].alloc(f.N);
    f.c2r(P[k].p, R[k].p);
  }
  MBuf<int> d_off(off.size());
  MBuf<double> d_out(4 * n);
  DR_HIP(hipMemcpyAsync(d_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_gather_offsets, dim3(cdiv(n, 256)), dim3(256), 0, s, (int)n, d_off.p, M[0], M[1], M[2], R[0].p,
                     R[1].p, R[2].p, R[3].p, d_out.p);
  DR_HIP(hipGetLastError());
  std::vector<double> out(4 * n);
  DR_HIP(hipMemcpyAsync(out.data(), d_out.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  DR_HIP(hipStreamSynchronize(s));
  for (long i = 0; i < n; ++i) out[4 * i + 3] = std::nearbyint(out[4 * i + 3]);
  return out;
}

// ||[p, 0] Q^T|| in float, rows accumulated left to right (estimate_kernel_parameters.h:70)
static float offset_distance(const float q[16], const int* p) {
  float ss = 0.0f;
  for (int r = 0; r < 4; ++r) {
    float v = 0.0f;
    for (int c = 0; c < 3; ++c) {
      const volatile float prod = q[4 * r + c] * (float)p[c];
      v = v + prod;
    }
    const volatile float sq = v * v;
    ss = ss + sq;
  }
  return sqrt_rn(ss);
}

// ---------------------------------------------------------------- allocation-free summary
// dr_mce_summary_noalloc keeps every intermediate in one fixed device array
// per GPU, `scratch`, and the offset table in `mce_offset_table`.  Host code
// never passes these names to a kernel: on each device the first call looks
// up their device addresses with hipGetSymbolAddress (that device current)
// and every later use goes through the recorded pointers.
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
constexpr int kNumMceOffsets = 25348;
constexpr int kMaxNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static __device__ __align__(256) int mce_offset_table[3 * kNumMceOffsets];

struct NoAllocDevice {
  bool ready = false;
  unsigned char* scr = nullptr;  // device address of `scratch` on this device
  int* off = nullptr;            // device address of `mce_offset_table`
};
static NoAllocDevice g_noalloc[kMaxNoAllocDevices];
static std::mutex g_noalloc_mu;

// one-time set-up on `device` (current): symbol addresses and the offset table
static const NoAllocDevice& noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocDevice& d = g_noalloc[device];
  if (d.ready) return d;
  void* p = nullptr;
  DR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  d.scr = static_cast<unsigned char*>(p);
  p = nullptr;
  DR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(mce_offset_table)));
  d.off = static_cast<int*>(p);
  const std::vector<int> P = mce_offsets();
  DR_REQUIRE((long)P.size() == 3L * kNumMceOffsets, DR_ELOGIC, "minimum contrast: offset table size changed");
  DR_HIP(hipMemcpy(d.off, P.data(), P.size() * sizeof(int), hipMemcpyHostToDevice));
  d.ready = true;
  return d;
}

namespace {

// exp(-2 pi i e / M): cos and sin of pi * (2e/M), the argument reduced exactly
__global__ void k_twiddles(int M, cx<double>* __restrict__ t) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  double sn, cs;
  sincospi(2.0 * (double)e / (double)M, &sn, &cs);
  t[e] = mk<double>(cs, -sn);
}

// one of the four products conj(A) B of k_corr_products
__global__ void k_corr_product_one(long Nh, const cx<double>* __restrict__ X, const cx<double>* __restrict__ Mk,
                                   int k, cx<double>* __restrict__ P) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nh) return;
  const cx<double> x = X[i], m = Mk[i];
  const cx<double> a = (k == 0 || k == 2) ? x : m;
  const cx<double> b = (k == 0 || k == 1) ? x : m;
  P[i] = cmulc(a, b);
}

// column k of k_gather_offsets
__global__ void k_gather_one(int n, const int* __restrict__ off, int M0, int M1, int M2,
                             const double* __restrict__ R, int k, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = ((off[3 * i] % M0) + M0) % M0;
  const int y = ((off[3 * i + 1] % M1) + M1) % M1;
  const int z = ((off[3 * i + 2] % M2) + M2) % M2;
  out[4 * i + k] = R[x + (long)M0 * (y + (long)M1 * z)];
}

struct QForm {
  float q[16];
};

// dr_mce_summary's host loop on the device, with the same rounding: pair
// counts rounded, (Sxy - Sx Sy / n) / (n - 1) in float64 then float, and the
// float distance ||[p, 0] Q^T|| with rows accumulated left to right
__global__ void k_mce_finalize(int n, const double* __restrict__ sums, const int* __restrict__ off, QForm qf,
                               float* __restrict__ cov, float* __restrict__ dist, int32_t* __restrict__ npairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sxy = sums[4 * i], sx = sums[4 * i + 1], sy = sums[4 * i + 2];
  const long np = (long)rint(sums[4 * i + 3]);
  npairs[i] = (int32_t)np;
  cov[i] = np > 1 ? (float)__ddiv_rn(__dsub_rn(sxy, __ddiv_rn(__dmul_rn(sx, sy), (double)np)), (double)(np - 1))
                  : 0.0f;
  float ss = 0.0f;
  for (int r = 0; r < 4; ++r) {
    float v = 0.0f;
    for (int c = 0; c < 3; ++c) v = add_rn(v, mul_rn(qf.q[4 * r + c], (float)off[3 * i + c]));
    ss = add_rn(ss, mul_rn(v, v));
  }
  dist[i] = sqrt_rn(ss);
}

}  // namespace

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Scratch bytes the allocation-free summary needs for an image of `dims`
// (twiddles, T1, T2, the two forward transforms, one product, one
// correlation grid, the per-offset sums); fills the padded torus M.
static size_t noalloc_bytes(const int dims[3], int M[3]) {
  for (int a = 0; a < 3; ++a) M[a] = smooth_even(dims[a] + 25 + 1);  // largest offset component: 25
  long N, Nh;
  Fft3d::sizes(M, &N, &Nh);
  size_t b = 0;
  for (int a = 0; a < 3; ++a) b += align256((size_t)M[a] * sizeof(cx<double>));
  b += 5 * align256((size_t)Nh * sizeof(cx<double>));
  b += align256((size_t)N * sizeof(double));
  b += align256((size_t)4 * kNumMceOffsets * sizeof(double));
  return b;
}

}  // namespace drh

extern "C" {

int dr_mce_offsets(int64_t* n, int32_t* offsets_xyz) {
  return drh::guarded([&] {
    DR_REQUIRE(n, DR_ELOGIC, "dr_mce_offsets: null count");
    const std::vector<int> P = drh::mce_offsets();
    *n = (int64_t)(P.size() / 3);
    if (offsets_xyz) std::copy(P.begin(), P.end(), offsets_xyz);
  });
}

int dr_mce_summary(const double* img, const int dims[3], const float qform[16], int device, float* covariance,
                   float* distance, int32_t* npairs) {
  return drh::guarded([&] {
    DR_REQUIRE(covariance && distance && npairs && qform, DR_ELOGIC, "dr_mce_summary: null output");
    const std::vector<int> P = drh::mce_offsets();
    const std::vector<double> s = drh::corr_sums(img, dims, P, device);
    const long n = (long)P.size() / 3;
    for (long i = 0; i < n; ++i) {
      const double sxy = s[4 * i], sx = s[4 * i + 1], sy = s[4 * i + 2];
      const long np = (long)s[4 * i + 3];
      npairs[i] = (int32_t)np;
      covariance[i] = np > 1 ? (float)((sxy - sx * sy / (double)np) / (double)(np - 1)) : 0.0f;
      distance[i] = drh::offset_distance(qform, &P[3 * i]);
    }
  });
}

int dr_mce_summary_noalloc(const double* d_img, const int dims[3], const float qform[16], int device, void* stream,
                           float* d_covariance, float* d_distance, int32_t* d_npairs) {
  return drh::guarded([&] {
    using namespace drh;
    DR_REQUIRE(d_img && dims && qform && d_covariance && d_distance && d_npairs, DR_ELOGIC,
               "dr_mce_summary_noalloc: null argument");
    for (int a = 0; a < 3; ++a) DR_REQUIRE(dims[a] >= 1, DR_EDOMAIN, "minimum contrast: empty image");
    DR_REQUIRE(device >= 0 && device < kMaxNoAllocDevices, DR_EDOMAIN, "dr_mce_summary_noalloc: bad device");
    int M[3];
    const size_t need = noalloc_bytes(dims, M);
    if (need > kNoAllocScratchBytes)
      throw Error(DR_EDOMAIN, "dr_mce_summary_noalloc: image " + std::to_string(dims[0]) + "x" +
                                  std::to_string(dims[1]) + "x" + std::to_string(dims[2]) + " needs " +
                                  std::to_string(need) + " bytes of scratch; the limit is " +
                                  std::to_string(kNoAllocScratchBytes));
    DR_HIP(hipSetDevice(device));
    const NoAllocDevice& d = noalloc_setup(device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    long N, Nh;
    Fft3d::sizes(M, &N, &Nh);
    unsigned char* cur = d.scr;
    auto carve = [&](size_t bytes) {
      unsigned char* p = cur;
      cur += align256(bytes);
      return p;
    };
    cx<double>* tw[3];
    for (int a = 0; a < 3; ++a) tw[a] = reinterpret_cast<cx<double>*>(carve((size_t)M[a] * sizeof(cx<double>)));
    auto cbuf = [&]() { return reinterpret_cast<cx<double>*>(carve((size_t)Nh * sizeof(cx<double>))); };
    cx<double>* T1 = cbuf();
    cx<double>* T2 = cbuf();
    cx<double>* X = cbuf();
    cx<double>* Mk = cbuf();
    cx<double>* P = cbuf();
    double* R = reinterpret_cast<double*>(carve((size_t)N * sizeof(double)));
    double* sums = reinterpret_cast<double*>(carve((size_t)4 * kNumMceOffsets * sizeof(double)));
    DR_REQUIRE((size_t)(cur - d.scr) == need, DR_ELOGIC, "dr_mce_summary_noalloc: scratch layout mismatch");
    for (int a = 0; a < 3; ++a) {
      hipLaunchKernelGGL(k_twiddles, dim3(cdiv(M[a], 256)), dim3(256), 0, s, M[a], tw[a]);
      DR_HIP(hipGetLastError());
    }
    Fft3d f;
    f.plan(M, s, tw, T1, T2);
    f.r2c(LoadMasked{{}, d_img, dims[0], dims[1], dims[2], M[1], 0}, X);
    f.r2c(LoadMasked{{}, d_img, dims[0], dims[1], dims[2], M[1], 1}, Mk);
    for (int k = 0; k < 4; ++k) {
      hipLaunchKernelGGL(k_corr_product_one, dim3(cdiv(Nh, 256)), dim3(256), 0, s, Nh, X, Mk, k, P);
      DR_HIP(hipGetLastError());
      f.c2r(P, R);
      hipLaunchKernelGGL(k_gather_one, dim3(cdiv(kNumMceOffsets, 256)), dim3(256), 0, s, kNumMceOffsets, d.off, M[0],
                         M[1], M[2], R, k, sums);
      DR_HIP(hipGetLastError());
    }
    QForm qf;
    for (int i = 0; i < 16; ++i) qf.q[i] = qform[i];
    hipLaunchKernelGGL(k_mce_finalize, dim3(cdiv(kNumMceOffsets, 256)), dim3(256), 0, s, kNumMceOffsets, sums, d.off,
                       qf, d_covariance, d_distance, d_npairs);
    DR_HIP(hipGetLastError());
  });
}

int dr_mce_binned(const double* img, const int dims[3], const float qform[16], int device, int nbins, double maxdist,
                  double* bin_cov, double* bin_dist, int64_t* bin_npairs) {
  return drh::guarded([&] {
    DR_REQUIRE(nbins >= 1 && maxdist > 0 && bin_cov && bin_dist && bin_npairs && qform, DR_ELOGIC,
               "dr_mce_binned: bad arguments");
    // offsets in the half space (z > 0, or z = 0 and y > 0, or z = y = 0 and x >= 0) within maxdist
    int reach[3];
    for (int c = 0; c < 3; ++c) {
      const float col = std::sqrt(qform[c] * qform[c] + qform[4 + c] * qform[4 + c] + qform[8 + c] * qform[8 + c]);
      reach[c] = (int)std::ceil(maxdist / std::max(col, 1e-6f));
    }
    std::vector<int> P;
    std::vector<float> dist;
    for (int z = 0; z <= reach[2]; ++z)
      for (int y = -reach[1]; y <= reach[1]; ++y)
        for (int x = -reach[0]; x <= reach[0]; ++x) {
          if (z == 0 && (y < 0 || (y == 0 && x < 0))) continue;
          const int p[3] = {x, y, z};
          const float d = drh::offset_distance(qform, p);
          if (d <= (float)maxdist) {
            P.insert(P.end(), {x, y, z});
            dist.push_back(d);
          }
        }
    const std::vector<double> s = drh::corr_sums(img, dims, P, device);
    std::vector<double> sxy(nbins, 0.0), sx(nbins, 0.0), sy(nbins, 0.0), sd(nbins, 0.0);
    std::vector<int64_t> nn(nbins, 0);
    const double w = maxdist / nbins;
    for (size_t i = 0; i < dist.size(); ++i) {
      const int b = std::min(nbins - 1, (int)(dist[i] / w));
      const int64_t np = (int64_t)s[4 * i + 3];
      sxy[b] += s[4 * i];
      sx[b] += s[4 * i + 1];
      sy[b] += s[4 * i + 2];
      nn[b] += np;
      sd[b] += (double)dist[i] * (double)np;
    }
    for (int b = 0; b < nbins; ++b) {
      bin_npairs[b] = nn[b];
      bin_cov[b] = nn[b] > 1 ? (sxy[b] - sx[b] * sy[b] / (double)nn[b]) / (double)(nn[b] - 1) : 0.0;
      bin_dist[b] = nn[b] > 0 ? sd[b] / (double)nn[b] : (b + 0.5) * w;
    }
  });
}

}  // extern "C"
