// Load/store functors fused into the FFT passes, and the non-FFT kernels
// of the gradient: CSR W.mu, W^T v, reductions.
//
// Spectral arrays hold the half spectrum S[kz][ky][kx], kx in [0, M0/2],
// of real torus fields.  invlam[k] = 1/lambda_k for lambda_k > 0 and 0 on
// the low-rank set Z (Re lambda <= 0), which implements every "zero Z" of
// the reference (MultiResParameters.h:309-319) with a single array.
#pragma once

#include "dr_fft.hpp"

namespace drh {

struct NoRed {
  __device__ __forceinline__ void finish(double*, int) {}
};

// Per-iteration scalars of one chain in device memory.  The functors and
// kernels of a captured HMC iteration read them (bind()) in place of their
// by-value fields, so one graph serves every step size and noise precision.
// The values are the host's conversions to T, the same ones the by-value
// fields of an eager launch hold.
template <typename T> struct StepScalars {
  T half, full;        // (T)(eps / 2), (T)eps
  T sig0, sig1;        // sigma_h^-2, sigma_s^-2
  uint64_t stream[2];  // Philox streams of the momentum draw (real, imaginary chain)
};

// weight of a half-spectrum entry in a full-torus sum: 1 on the kx = 0 and
// kx = M0/2 planes, 2 elsewhere (Hermitian partner not stored)
__device__ __forceinline__ double herm_w(int kx, int H0) { return (kx == 0 || kx == H0 - 1) ? 1.0 : 2.0; }

__device__ __forceinline__ void red_write(double v, double* red, double* part, int blk) {
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blk] = s;
}

// ---------------------------------------------------------------- strided functors
template <typename T> struct LoadS : NoRed {
  const cx<T>* src; long so, sk;
  __device__ __forceinline__ cx<T> load(int o, int k, int b) const { return src[o * so + (long)k * sk + b]; }
};
template <typename T> struct StoreS : NoRed {
  cx<T>* dst; long so, sk;
  __device__ __forceinline__ void store(int o, int k, int b, cx<T> v) const { dst[o * so + (long)k * sk + b] = v; }
};

// position update mu += eps * M^{-1} p fused into the first inverse pass
// (MultiResParameters.h:693-698): M^{-1} = 1 / (1/lambda + sigma_h^-2), 0 on Z.
template <typename T> struct LoadPosUpdate : NoRed {
  const cx<T>* src; cx<T>* dst; const cx<T>* p; const T* invlam; long sk; T eps; T sig;
  const cx<T>* srci; cx<T>* dsti; const cx<T>* pim;  // imaginary chain (may be null)
  __device__ __forceinline__ cx<T> load(int, int k, int b) const {
    const long i = (long)k * sk + b;
    const T il = invlam[i];
    const T c = il > (T)0 ? eps / (il + sig) : (T)0;
    const cx<T> m = src[i] + p[i] * c;
    dst[i] = m;
    if (dsti) dsti[i] = srci[i] + pim[i] * c;
    return m;
  }
};

// momentum update p += coef * (F(S r) - mu/lambda) fused into the last
// forward pass (MultiResParameters.h:690,699 with _compute_gradient :349-374);
// optional reductions of the prior and kinetic energies (Parseval).
template <typename T> struct EpiMomentum {
  const cx<T>* X; cx<T>* p; const T* invlam; long sk; T coef; T sig; int H0;
  const cx<T>* Xi; cx<T>* pim;
  int do_red; double* part; long pstride; int s_prior, s_priori, s_kin, s_kini;
  double a0, a1, a2, a3;
  const StepScalars<T>* sc;  // half kick
  __device__ __forceinline__ void bind() {
    if (sc) {
      coef = sc->half;
      sig = sc->sig0;
    }
  }
  // the state one mode's update reads (k_strided fetches several modes'
  // state before updating them)
  struct State { T il; cx<T> x, p, xi, pi; };
  __device__ __forceinline__ State fetch(int, int k, int b) const {
    const long i = (long)k * sk + b;
    State s{invlam[i], X[i], p[i], mk<T>(0, 0), mk<T>(0, 0)};
    if (pim) {
      s.xi = Xi[i];
      s.pi = pim[i];
    }
    return s;
  }
  __device__ __forceinline__ void store(int o, int k, int b, cx<T> v) { store(o, k, b, v, fetch(o, k, b)); }
  __device__ __forceinline__ void store(int, int k, int b, cx<T> v, const State& s) {
    const long i = (long)k * sk + b;
    const T il = s.il;
    const cx<T> x = s.x;
    const cx<T> pn = s.p + (v - x * il) * coef;
    p[i] = pn;
    cx<T> xi = mk<T>(0, 0), pin = mk<T>(0, 0);
    if (pim) {
      xi = s.xi;
      pin = s.pi - xi * (il * coef);
      pim[i] = pin;
    }
    if (do_red) {
      const double w = herm_w(b % H0, H0);
      const double ild = (double)il;
      const double ilm = il > (T)0 ? 1.0 / (double)(il + sig) : 0.0;
      a0 += w * ild * ((double)x.x * x.x + (double)x.y * x.y);
      a2 += w * ilm * ((double)pn.x * pn.x + (double)pn.y * pn.y);
      a1 += w * ild * ((double)xi.x * xi.x + (double)xi.y * xi.y);
      a3 += w * ilm * ((double)pin.x * pin.x + (double)pin.y * pin.y);
    }
  }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (!do_red) return;
    red_write(a0, red, part + s_prior * pstride, blk);
    red_write(a1, red, part + s_priori * pstride, blk);
    red_write(a2, red, part + s_kin * pstride, blk);
    red_write(a3, red, part + s_kini * pstride, blk);
  }
};

// kick then drift between two gradients (MultiResParameters.h:690-699):
//   p  += coef (F(S r) - X/lambda)      with the gradient taken at X
//   X' =  X + eps p / lambda_m          written to Xd and returned for the
//                                       inverse pass of the next gradient.
// Optionally reduces the prior energy of X.
// Non-temporal (streaming) access for data that the next kernels do not
// reuse: the leapfrog state in the z pass, the boundary W entries.  Leaving
// them out of the caches keeps the transforms' intermediate fields cached
// (measured at 1 mm: W mu 101 -> 97 us, W^T v 95 -> 91 us per step).
template <typename T>
__device__ __forceinline__ cx<T> nt_ld(const cx<T>* q) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(q));
  return mk(x[0], x[1]);
}
template <typename T>
__device__ __forceinline__ void nt_st(cx<T>* q, cx<T> v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2*>(q));
}
template <typename T> struct MidLeapfrog {
  const cx<T>* X; cx<T>* Xd; cx<T>* p; const T* invlam; long sk; T coef; T eps; T sig; int H0;
  int do_red; double* part; double a0;
  const StepScalars<T>* sc; int coef_full;  // coef = full (1) or half (0) step
  __device__ __forceinline__ void bind() {
    if (sc) {
      coef = coef_full ? sc->full : sc->half;
      eps = sc->full;
      sig = sc->sig0;
    }
  }
  // the state one mode's update reads; each mode is read and written by one
  // thread only, so a thread may fetch its modes' state early
  struct State { cx<T> x, p; T il; };
  __device__ __forceinline__ State fetch(int k, int b) const {
    const long i = (long)k * sk + b;
    // 1/lambda keeps cached loads: a 16-column tile reads half of each of its
    // 128-byte lines, and the neighbouring tile the other half
    return State{nt_ld(X + i), nt_ld(p + i), invlam[i]};
  }
  __device__ __forceinline__ cx<T> apply(int k, int b, cx<T> v, const State& s) {
    const long i = (long)k * sk + b;
    const cx<T> pn = s.p + (v - s.x * s.il) * coef;
    nt_st(p + i, pn);
    const T c = s.il > (T)0 ? eps / (s.il + sig) : (T)0;
    const cx<T> m = s.x + pn * c;
    nt_st(Xd + i, m);
    if (do_red) a0 += herm_w(b % H0, H0) * (double)s.il * ((double)s.x.x * s.x.x + (double)s.x.y * s.x.y);
    return m;
  }
  __device__ __forceinline__ cx<T> apply(int k, int b, cx<T> v) { return apply(k, b, v, fetch(k, b)); }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (do_red) red_write(a0, red, part, blk);
  }
};

// The data-free imaginary chain of the reference's complex state evolves under
// the prior alone, so every Fourier mode is an independent linear oscillator:
// all L leapfrog steps run in registers in one pass per iteration, with the
// same per-step arithmetic as the real chain.  Reduces its prior energy before
// and after and its final kinetic energy.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_imag_leapfrog(long Nh, int H0, const cx<T>* __restrict__ muA, cx<T>* __restrict__ muB, cx<T>* __restrict__ p,
                const T* __restrict__ invlam, T eps, T half_coef, T full_coef, T sig, int L, double* partA,
                double* partB, double* partK, const StepScalars<T>* sc) {
  __shared__ double red[kThreads];
  if (sc) {
    eps = sc->full;
    half_coef = sc->half;
    full_coef = sc->full;
    sig = sc->sig0;
  }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < Nh; i += (long)gridDim.x * kThreads) {
    const T il = invlam[i];
    const T c = il > (T)0 ? eps / (il + sig) : (T)0;
    cx<T> m = muA[i];
    cx<T> q = p[i];
    const double w = herm_w((int)(i % H0), H0);
    a0 += w * (double)il * ((double)m.x * m.x + (double)m.y * m.y);
    q = q - m * (il * half_coef);
    for (int s = 0; s < L; ++s) {
      m = m + q * c;
      q = q - m * (il * (s == L - 1 ? half_coef : full_coef));
    }
    muB[i] = m;
    p[i] = q;
    const double ilm = il > (T)0 ? 1.0 / (double)(il + sig) : 0.0;
    a1 += w * (double)il * ((double)m.x * m.x + (double)m.y * m.y);
    a2 += w * ilm * ((double)q.x * q.x + (double)q.y * q.y);
  }
  red_write(a0, red, partA, blockIdx.x);
  red_write(a1, red, partB, blockIdx.x);
  red_write(a2, red, partK, blockIdx.x);
}

// gradient in the spectral domain: G = F(S r) - X/lambda
template <typename T> struct EpiGrad : NoRed {
  const cx<T>* X; cx<T>* G; const T* invlam; long sk;
  __device__ __forceinline__ void store(int, int k, int b, cx<T> v) const {
    const long i = (long)k * sk + b;
    G[i] = v - X[i] * invlam[i];
  }
};

// momentum draw p = sqrt(lambda_m) F(z), 0 on Z (MultiResParameters.h:811-814)
// p = sqrt(lambda_m) F(z), 0 on Z (MultiResParameters.h:800-815), and the
// sum that gives E0 = -1/2 sum |z|^2 (:787-799) by Parseval: w |p|^2 /
// lambda_m off Z, term for term as the final half kick sums the end point's
// kinetic energy (EpiMomentum), plus w |F(z)|^2 on Z, whose energy the
// reference's E0 holds although p does not.  Summing the spectrum rather
// than z matters in float32: a transform rounds the field's energy by a few
// ulp, and that difference between E0 and the end point's kinetic energy
// would shift every accept ratio (by about exp(2.7) at the 1 mm workload).
template <typename T> struct EpiNoiseMomentum {
  cx<T>* p; const T* invlam; long sk; T sig;
  const StepScalars<T>* sc;
  int H0; double* part; double acc;
  __device__ __forceinline__ void bind() {
    if (sc) sig = sc->sig0;
  }
  struct State { T il; };
  __device__ __forceinline__ State fetch(int, int k, int b) const { return State{invlam[(long)k * sk + b]}; }
  __device__ __forceinline__ void store(int o, int k, int b, cx<T> v) { store(o, k, b, v, fetch(o, k, b)); }
  // sqrt in T: the correctly rounded float square root equals the float
  // rounding of the double one
  __device__ __forceinline__ void store(int, int k, int b, cx<T> v, const State& s) {
    const long i = (long)k * sk + b;
    const T il = s.il;
    const cx<T> pn = il > (T)0 ? v * sqrt(il + sig) : mk<T>(0, 0);
    p[i] = pn;
    if (part) {
      const double e = il > (T)0 ? 1.0 / (double)(il + sig) * ((double)pn.x * pn.x + (double)pn.y * pn.y)
                                 : (double)v.x * v.x + (double)v.y * v.y;
      acc += herm_w(b % H0, H0) * e;
    }
  }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (part) red_write(acc, red, part, blk);
  }
};

// The momentum draw in one pass over the half spectrum.  For a real field z
// of iid N(0, 1) on the N-point torus, F(z) is Hermitian white noise: every
// mode k with a distinct partner -k has Re and Im iid N(0, N / 2) and F(z)(-k)
// = conj F(z)(k); a self-conjugate mode (k = -k) is real N(0, N).  So the
// spectral momentum sqrt(lambda_m) F(z) of EpiNoiseMomentum is drawn here
// directly, without the three transforms: one normal pair per mode, keyed by
// the mode's index, and in the planes kx = 0 and kx = M0 / 2 (where the half
// spectrum holds both k and -k) by the smaller index of the pair, the other
// taking the conjugate.  Float64: the pair of key q comes from Philox block q
// (53-bit uniforms from its four words).  Float32: from block q / 2, words
// (a, c) for even q and (b, d) for odd q, so a thread's two consecutive modes
// share one block.  E0 is summed as EpiNoiseMomentum sums it.  Even M0, M1,
// M2 (Nh is even, below 2^31: the index arithmetic is 32-bit).
struct ModeKey {
  uint32_t key, kx;
  bool self, cj;
};
__device__ __forceinline__ ModeKey mode_key(uint32_t i, uint32_t H0, uint32_t M1, uint32_t M2) {
  const uint32_t row = i / H0, kx = i - row * H0;
  ModeKey m{i, kx, false, false};
  if (kx == 0 || kx == H0 - 1) {
    const uint32_t kz = row / M1, ky = row - kz * M1;
    const uint32_t ip = (((M2 - kz) % M2) * M1 + (M1 - ky) % M1) * H0 + kx;
    m.self = ip == i;
    m.cj = ip < i;
    if (m.cj) m.key = ip;
  }
  return m;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_momentum_direct(long Nh, int H0, int M1, int M2, uint64_t seed, uint64_t stream, int which, T sig,
                  const T* __restrict__ invlam, cx<T>* __restrict__ p, double* part, const StepScalars<T>* sc,
                  T sd_pair, T sd_self) {
  __shared__ double red[kThreads];
  if (sc) {
    stream = sc->stream[which];
    sig = sc->sig0;
  }
  double acc = 0.0;
  for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; 2 * (long)t < Nh; t += gridDim.x * kThreads) {
    ModeKey m[2];
    T a[2], b[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) m[u] = mode_key(2 * t + u, (uint32_t)H0, (uint32_t)M1, (uint32_t)M2);
    if constexpr (sizeof(T) == 4) {
      const u4 r0 = philox_block(seed, stream, (uint64_t)(m[0].key >> 1));
      const u4 r1 = (m[1].key >> 1) == (m[0].key >> 1) ? r0 : philox_block(seed, stream, (uint64_t)(m[1].key >> 1));
      box_muller_f((m[0].key & 1) ? r0.b : r0.a, (m[0].key & 1) ? r0.d : r0.c, a[0], b[0]);
      box_muller_f((m[1].key & 1) ? r1.b : r1.a, (m[1].key & 1) ? r1.d : r1.c, a[1], b[1]);
    } else {
#pragma unroll
      for (int u = 0; u < 2; ++u) normal_pair(seed, stream, (uint64_t)m[u].key, a[u], b[u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t i = 2 * t + u;
      const cx<T> xi = m[u].self ? mk<T>(a[u] * sd_self, (T)0)
                                 : mk<T>(a[u] * sd_pair, (m[u].cj ? -b[u] : b[u]) * sd_pair);
      const T il = invlam[i];
      const cx<T> pn = il > (T)0 ? xi * sqrt(il + sig) : mk<T>(0, 0);
      p[i] = pn;
      const double e = il > (T)0 ? 1.0 / (double)(il + sig) * ((double)pn.x * pn.x + (double)pn.y * pn.y)
                                 : (double)xi.x * xi.x + (double)xi.y * xi.y;
      acc += herm_w((int)m[u].kx, H0) * e;
    }
  }
  red_write(acc, red, part, blockIdx.x);
}

// initial state mu = sqrt(lambda) F(z) / 10, 0 on Z (MultiResParameters.h:752-755)
template <typename T> struct EpiInitMu : NoRed {
  cx<T>* X; const double* lam; long sk;
  __device__ __forceinline__ void store(int, int k, int b, cx<T> v) const {
    const long i = (long)k * sk + b;
    const double l = lam[i];
    X[i] = l > 0.0 ? v * (T)(sqrt(l) / 10.0) : mk<T>(0, 0);
  }
};

template <typename T> struct EpiAddTo : NoRed {
  cx<T>* X; long sk;
  __device__ __forceinline__ void store(int, int k, int b, cx<T> v) const {
    const long i = (long)k * sk + b;
    X[i] = X[i] + v;
  }
};

// eigenvalues: lambda = Re F(c); invlam for the working precision; weighted
// count of Re lambda <= 0 (MultiResParameters.h:253-262)
template <typename T> struct EpiSpectrum {
  double* lam; T* invlam; long sk; int H0; double* part; double acc;
  __device__ __forceinline__ void store(int, int k, int b, cx<double> v) {
    const long i = (long)k * sk + b;
    lam[i] = v.x;
    invlam[i] = v.x > 0.0 ? (T)(1.0 / v.x) : (T)0;
    if (!(v.x > 0.0)) acc += herm_w(b % H0, H0);
  }
  __device__ __forceinline__ void finish(double* red, int blk) { red_write(acc, red, part, blk); }
};

// ---------------------------------------------------------------- x-pass functors
template <typename T> struct LoadRowC : NoRed {
  const cx<T>* src; int H0;
  __device__ __forceinline__ cx<T> load(int row, int k) const { return src[(long)row * H0 + k]; }
};
template <typename T> struct StoreRowC : NoRed {
  cx<T>* dst; int H0;
  __device__ __forceinline__ void store(int row, int k, cx<T> v) const { dst[(long)row * H0 + k] = v; }
};
template <typename T> struct LoadRowR : NoRed {
  const T* src; int n_in;
  __device__ __forceinline__ cx<T> load2(int row, int n) const {
    const int x = 2 * n;
    const long base = (long)row * n_in;
    return mk<T>(x < n_in ? src[base + x] : (T)0, x + 1 < n_in ? src[base + x + 1] : (T)0);
  }
};
template <typename T> struct StoreRowR : NoRed {
  T* dst; int n_out; T scale;
  __device__ __forceinline__ void store2(int row, int n, cx<T> z) const {
    const int x = 2 * n;
    const long base = (long)row * n_out;
    if (x < n_out) dst[base + x] = z.x * scale;
    if (x + 1 < n_out) dst[base + x + 1] = z.y * scale;
  }
};

// standard normal field on the full torus, pair index = flat / 2; the float64
// sum of squares of the working-precision values gives E0 (:785-809)
template <typename T> struct LoadNoise {
  uint64_t seed, stream; int M0h; double* part; double acc;
  const StepScalars<T>* sc; int which;  // stream = sc->stream[which]
  __device__ __forceinline__ void bind() {
    if (sc) stream = sc->stream[which];
  }
  // normals in the chain's precision
  __device__ __forceinline__ cx<T> load2(int row, int n) {
    T a, b;
    normal_pair(seed, stream, (uint64_t)row * (uint64_t)M0h + (uint64_t)n, a, b);
    acc += (double)a * (double)a + (double)b * (double)b;
    return mk<T>(a, b);
  }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (part) red_write(acc, red, part, blk);
  }
};
// injected normal field (test hook), same E0 reduction
template <typename T> struct LoadNoiseBuf {
  const T* z; int M0; double* part; double acc;
  __device__ __forceinline__ cx<T> load2(int row, int n) {
    const long i = (long)row * M0 + 2 * n;
    const T a = z[i], b = z[i + 1];
    acc += (double)a * (double)a + (double)b * (double)b;
    return mk<T>(a, b);
  }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (part) red_write(acc, red, part, blk);
  }
};

// likelihood residual on the image region, scattered onto the torus rows
// (MultiResParameters.h:357-374): r = sigma_h^-2 (Yh - mu) + W^T v at data
// voxels, 0 elsewhere; accumulates ||Yh - mu[I]||^2 (:502-503).
template <typename T> struct LoadResidual {
  const T* R; const T* Yg; const uint8_t* mask; const T* WTv; const T* WTvb; T sig; int n0; double* part; double acc;
  const StepScalars<T>* sc;
  __device__ __forceinline__ void bind() {
    if (sc) sig = sc->sig0;
  }
  // split load: fetch2 reads the inputs of samples 2n, 2n+1 of a row without
  // side effects (any row < nrows, any n: x is clamped into the row), apply2
  // forms the residual and accumulates the SSE for the samples inside the row
  struct Raw { T y[2], r[2], w[2], wb[2]; uint8_t m[2]; };
  static constexpr bool kSplitLoad = true;
  __device__ __forceinline__ Raw fetch2(int row, int n) const {
    Raw q;
    const long base = (long)row * n0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const long g = base + min(2 * n + e, n0 - 1);
      q.m[e] = mask[g];
      q.y[e] = Yg[g];
      q.r[e] = R[g];
      q.w[e] = WTv ? WTv[g] : (T)0;
      q.wb[e] = WTvb ? WTvb[g] : (T)0;
    }
    return q;
  }
  __device__ __forceinline__ T one(const Raw& q, int e) {
    if (!q.m[e]) return (T)0;
    const T d = q.y[e] - q.r[e];
    acc += (double)d * (double)d;
    T r = sig * d;
    if (WTv) r += q.w[e];
    if (WTvb) r += q.wb[e];
    return r;
  }
  __device__ __forceinline__ cx<T> apply2(const Raw& q, int n) {
    const int x = 2 * n;
    const T a = x < n0 ? one(q, 0) : (T)0;
    const T b = x + 1 < n0 ? one(q, 1) : (T)0;
    return mk<T>(a, b);
  }
  __device__ __forceinline__ cx<T> load2(int row, int n) { return apply2(fetch2(row, n), n); }
  __device__ __forceinline__ void finish(double* red, int blk) {
    if (part) red_write(acc, red, part, blk);
  }
};

// circulant base c = cov(||Q (d(i), d(j), d(k))||) with d(t) = min(t, M-t);
// distance in float32 (impl/circulant_base.inl:60-85), kernel in float64
// (impl/kernels.inl:7-15, 58-74).
struct LoadBase : NoRed {
  int M0, M1, M2; float q[9]; double tau2, psi, nu; int covfun;
  __device__ __forceinline__ double cov(int x, int y, int z) const {
    const float dx = (float)(x <= M0 / 2 ? x : M0 - x);
    const float dy = (float)(y <= M1 / 2 ? y : M1 - y);
    const float dz = (float)(z <= M2 / 2 ? z : M2 - z);
    float c[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float v = mul_rn(q[3 * r + 0], dx);
      v = add_rn(v, mul_rn(q[3 * r + 1], dy));
      v = add_rn(v, mul_rn(q[3 * r + 2], dz));
      c[r] = v;
    }
    const float ss = add_rn(add_rn(mul_rn(c[0], c[0]), mul_rn(c[1], c[1])), mul_rn(c[2], c[2]));
    const double d = (double)sqrt_rn(ss);
    if (covfun == 0) return tau2 * exp(-psi * pow(fabs(d), nu));
    return tau2 * pow(1.0 + psi / nu * d * d, -nu);
  }
  __device__ __forceinline__ cx<double> load2(int row, int n) const {
    const int y = row % M1, z = row / M1;
    return mk<double>(cov(2 * n, y, z), cov(2 * n + 1, y, z));
  }
};

// ---------------------------------------------------------------- other kernels
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// v_i = sigma_s^-2 (Ys_i - sum_j W_ij mu[col_ij]); one wavefront per row;
// accumulates ||Ys - W mu||^2 (MultiResParameters.h:361-363, 505-507).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_wmu(int n_s, const int* __restrict__ rp, const int* __restrict__ cols, const T* __restrict__ vals,
      const T* __restrict__ R, const T* __restrict__ Ys, T sig_s, T* __restrict__ v, double* part,
      const StepScalars<T>* sc) {
  __shared__ double red[kThreads];
  if (sc) sig_s = sc->sig1;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  double e2 = 0.0;
  if (row < n_s) {
    T acc = (T)0;
    const int j1 = rp[row + 1];
    for (int j = rp[row] + lane; j < j1; j += 64) acc += vals[j] * R[cols[j]];
    acc = wave_sum(acc);
    if (lane == 0) {
      const T e = Ys[row] - acc;
      v[row] = sig_s * e;
      e2 = (double)e * (double)e;
    }
  }
  if (part) red_write(e2, red, part, blockIdx.x);
}

// (W^T v) at each data voxel, written to its image-region grid position;
// transposed CSR, 16 lanes per data voxel (deterministic order).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_wtv(int n_h, const int* __restrict__ rp, const int* __restrict__ rows, const T* __restrict__ vals,
      const T* __restrict__ v, const int* __restrict__ gidx, T* __restrict__ out) {
  const int lane = threadIdx.x & 15;
  const int d = blockIdx.x * (kThreads / 16) + (threadIdx.x >> 4);
  T acc = (T)0;
  if (d < n_h) {
    const int j1 = rp[d + 1];
    for (int j = rp[d] + lane; j < j1; j += 16) acc += vals[j] * v[rows[j]];
  }
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 16);
  if (d < n_h && lane == 0) out[gidx[d]] = acc;
}

// Stencil + boundary form of the same two products.  The rows of W that are
// equal up to a translation of the image-region grid and make up the largest
// such class ("interior" rows: every neighbour present) share one stencil of
// nD (region-grid offset, weight) pairs.  Their bases (first entry) lie on a
// lattice c + sigma*Z^3, so W^T v restricted to them is, at a voxel x, a sum
// over only the stencil entries whose 3D displacement is congruent to x - c
// modulo sigma (the voxel's residue class), read from a grid Vl holding
// v at lattice coordinates.  All other rows ("boundary" rows) keep explicit
// entries with 3D offsets packed in 15 bits (5 per axis).

// packed 3D offset (dx | dy << 5 | dz << 10) -> region-grid flat offset
// (R0 and R01 below 2^24: 24-bit multiplies run at full rate, 32-bit ones at
// a quarter)
__device__ __forceinline__ int unpack3(unsigned q, int R0, int R01) {
  return (int)((q & 31u) + __umul24((unsigned)R0, (q >> 5) & 31u) + __umul24((unsigned)R01, q >> 10));
}

// four packed entries: offsets (4 x 16 bits) and weights, read as one vector each
template <typename T>
struct alignas(4 * sizeof(T)) Quad {
  T v[4];
};
// streaming reads of the boundary entries (non-temporal).  Every load
// instruction of a wavefront reads 64 consecutive quads: with 16 short runs
// per instruction instead, whose lines the next instruction needs again,
// non-temporal reads took W^T v from 97 to 148 us.
template <typename T>
__device__ __forceinline__ Quad<T> nt_quad(const Quad<T>* p) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  const v4 x = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return Quad<T>{{x[0], x[1], x[2], x[3]}};
}
__device__ __forceinline__ uint2 nt_u2(const uint2* p) {
  typedef unsigned v2 __attribute__((ext_vector_type(2)));
  const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return make_uint2(x[0], x[1]);
}
__device__ __forceinline__ unsigned lo16(unsigned x) { return x & 0xffffu; }
__device__ __forceinline__ unsigned hi16(unsigned x) { return x >> 16; }

// lanes per voxel of the boundary W^T v part (the transposed entry lists are
// interleaved in blocks of this many quads).  Measured at 1 mm with the
// wavefront-contiguous lists: W^T v 79 / 76.5 / 80 / 92 us with 1 / 2 / 4 / 8
// lanes per voxel.
constexpr int kWtvLanes = 2;

// Argument bundles of the four parts of the two products.
// bit patterns of T: weights read through them are wave-uniform loads the
// compiler may serve from the scalar cache (a T pointer could alias the output)
template <typename T> struct Bits;
template <> struct Bits<float> { using type = uint32_t; };
template <> struct Bits<double> { using type = uint64_t; };
__device__ __forceinline__ float from_bits(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ double from_bits(uint64_t u) { return __longlong_as_double((long long)u); }

template <typename T> struct WmuInt {
  int n; const int* row; const int* base; const int* lidx; const int* bpar; const int* wstart; const int* woff;
  const T* ww; int ncw; T* Vl;
};
// far rows: explicit region indices, v to the compact array Vf
template <typename T> struct WmuFar {
  int n; const int* row; const int* rp; const int* col; const T* w; T* Vf;
};
// boundary rows: packed offsets from each row's base voxel, v of row k to Vc[k]
template <typename T> struct WmuBnd {
  int n; const int* row; const int* rp; const int* base; const uint16_t* d; const T* w;
  int R0, R01; T* Vc;
};
template <typename T> struct WtvInt {
  const int* tiles; const int* cls_start; const int* ldo; const typename Bits<T>::type* w; const T* Vl;
  const uint8_t* mask; T* out;
  int ncls, TD[3], LT[3], c[3], sg[3], omin[3], lo[3], lv[3], R[3];
};
// touched voxels: Vc index = base[k] + 16-bit delta per entry
template <typename T> struct WtvBnd {
  // gs: entry offsets of the wavefronts' shared quad ranges
  int n; const int* g; const int* gs; const int* base; const uint16_t* d; const T* w; const T* Vc; T* out;
  const int* fp; const int* fidx; const T* fw; const T* Vf;  // far-row entries per voxel
};

// four consecutive values from a 16-byte (float) / 32-byte (double) aligned address
__device__ __forceinline__ void ld4(const float* p, float* v) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void ld4(const double* p, double* v) {
  const double2 q0 = reinterpret_cast<const double2*>(p)[0], q1 = reinterpret_cast<const double2*>(p)[1];
  v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
}

// lanes per boundary W mu row (the row's quads are interleaved in blocks of
// this many): a whole wavefront per row.  Measured at 1 mm: W mu 92.7 us
// with 64 lanes per row, 96.1 us with 32.
constexpr int kWmuBndLanes = 64;

// lanes per interior W mu row.  Fewer lanes per row put more rows, at
// neighbouring lattice points, in one wavefront: a load instruction then
// reads fewer distinct cache lines (the interior rows are bound by L1 tag
// lookups).  Measured at 1 mm, W mu 98 / 96 / 94 / 103 us with 8 / 4 / 2 / 1
// lanes per row.
constexpr int kWmuLanes = 2;

// Interior rows: kWmuLanes lanes per row; the workgroup's rows share base
// parity p (base mod 4), whose table of aligned 4-voxel windows (offset, 4
// weights) is staged in LDS.  Lane q takes windows q, q + kWmuLanes, ... with
// one vector load each; the partial sums meet in a butterfly.
template <typename T>
__device__ __forceinline__ double wmu_int_block(const WmuInt<T>& a, int blk, const T* __restrict__ R,
                                                const T* __restrict__ Ys, T sig_s) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* sw = reinterpret_cast<T*>(lds_raw);
  int* so = reinterpret_cast<int*>(lds_raw + (size_t)a.ncw * 4 * sizeof(T));
  const int p = a.bpar[blk];
  const int cs = a.wstart[p], nc = a.wstart[p + 1] - cs;
  for (int i = threadIdx.x; i < nc; i += kThreads) {
    so[i] = a.woff[cs + i];
#pragma unroll
    for (int j = 0; j < 4; ++j) sw[4 * i + j] = a.ww[4 * (cs + i) + j];
  }
  __syncthreads();
  constexpr int LN = kWmuLanes;
  const int q = threadIdx.x & (LN - 1);
  const int k = (blk * kThreads + (int)threadIdx.x) / LN;
  const int row = k < a.n ? a.row[k] : -1;
  T acc = (T)0;
  if (row >= 0) {
    const T* Rb = R + a.base[k];
    constexpr int B = 4;
    int c = q;
    for (; c + LN * (B - 1) < nc; c += LN * B) {
      T v[B][4];
#pragma unroll
      for (int b = 0; b < B; ++b) ld4(Rb + so[c + LN * b], v[b]);
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc += sw[4 * (c + LN * b) + j] * v[b][j];
    }
    for (; c < nc; c += LN) {
      T v[4];
      ld4(Rb + so[c], v);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc += sw[4 * c + j] * v[j];
    }
  }
#pragma unroll
  for (int o = 1; o < LN; o <<= 1) acc += __shfl_xor(acc, o, LN);
  double e2 = 0.0;
  if (row >= 0 && q == 0) {
    const T e = Ys[row] - acc;
    a.Vl[a.lidx[k]] = sig_s * e;
    e2 = (double)e * (double)e;
  }
  return e2;
}

// boundary rows, kWmuBndLanes lanes per row; entry lists start on multiples of 4
// and are padded to whole quads (offset 0, weight 0)
template <typename T>
__device__ __forceinline__ double wmu_bnd_block(const WmuBnd<T>& a, int blk, const T* __restrict__ R,
                                                const T* __restrict__ Ys, T sig_s) {
  constexpr int LB = kWmuBndLanes;
  const int lane = threadIdx.x & (LB - 1);
  const int k = blk * (kThreads / LB) + (int)threadIdx.x / LB;
  double e2 = 0.0;
  if (k < a.n) {
    const int j1 = a.rp[k + 1] >> 2;
    const T* Rb = R + a.base[k];
    const uint2* d4 = reinterpret_cast<const uint2*>(a.d);
    const Quad<T>* w4 = reinterpret_cast<const Quad<T>*>(a.w);
    T acc = (T)0;
    for (int j = (a.rp[k] >> 2) + lane; j < j1; j += LB) {
      const uint2 dq = nt_u2(d4 + j);
      const Quad<T> wq = nt_quad(w4 + j);
      acc += wq.v[0] * Rb[unpack3(lo16(dq.x), a.R0, a.R01)];
      acc += wq.v[1] * Rb[unpack3(hi16(dq.x), a.R0, a.R01)];
      acc += wq.v[2] * Rb[unpack3(lo16(dq.y), a.R0, a.R01)];
      acc += wq.v[3] * Rb[unpack3(hi16(dq.y), a.R0, a.R01)];
    }
#pragma unroll
    for (int o = LB / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LB);
    if (lane == 0) {
      const T e = Ys[a.row[k]] - acc;
      a.Vc[k] = sig_s * e;
      e2 = (double)e * (double)e;
    }
  }
  return e2;
}

// far rows (neighbourhoods too wide to pack), one wavefront per row
template <typename T>
__device__ __forceinline__ double wmu_far_block(const WmuFar<T>& a, int blk, const T* __restrict__ R,
                                                const T* __restrict__ Ys, T sig_s) {
  const int lane = threadIdx.x & 63;
  const int k = blk * (kThreads / 64) + (threadIdx.x >> 6);
  double e2 = 0.0;
  if (k < a.n) {
    T acc = (T)0;
    for (int j = a.rp[k] + lane; j < a.rp[k + 1]; j += 64) acc += a.w[j] * R[a.col[j]];
    acc = wave_sum(acc);
    if (lane == 0) {
      const T e = Ys[a.row[k]] - acc;
      a.Vf[k] = sig_s * e;
      e2 = (double)e * (double)e;
    }
  }
  return e2;
}

// v = sigma_s (Ys - W R) for every row and ||Ys - W R||^2: blocks [0, gi)
// take the interior rows, [gi, gi + gb) the packed boundary rows and the rest
// the far rows, in one launch so the latency-bound gathers and the streaming
// boundary entries overlap
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_wmu_sb(WmuInt<T> ai, WmuBnd<T> ab, WmuFar<T> af, int gi, int gb, const T* __restrict__ R,
         const T* __restrict__ Ys, T sig_s, double* part, const StepScalars<T>* sc) {
  __shared__ double red[kThreads];
  if (sc) sig_s = sc->sig1;
  const int blk = blockIdx.x;
  const double e2 = blk < gi        ? wmu_int_block(ai, blk, R, Ys, sig_s)
                    : blk < gi + gb ? wmu_bnd_block(ab, blk - gi, R, Ys, sig_s)
                                    : wmu_far_block(af, blk - gi - gb, R, Ys, sig_s);
  if (part) red_write(e2, red, part, blockIdx.x);
}

// W^T v of the interior rows (written, not added), one voxel tile per
// workgroup.  The tile T0 x T1 x T2 stages the box of lattice values v (Vl,
// zero outside the lattice grid) that its voxels reach in LDS, then walks the
// residue classes: a voxel x of class rho (x = c + sg u + rho) takes
// sum_e w_e v[u + o_e] over its class's sub-stencil, in sub-stencil order, with
// the entry offsets o_e pre-flattened into the box (ldo).  Lanes run along
// lattice x, so the LDS reads of a wavefront are consecutive.  Data voxels get
// the sum, other region voxels of the tile 0.  Weights are passed as bit
// patterns (Bits<T>).
__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

template <typename T>
__device__ __forceinline__ void wtv_int_block(const WtvInt<T>& a, int blk) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* sv = reinterpret_cast<T*>(lds_raw);
  int X0[3], u0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    X0[d] = a.tiles[3 * blk + d];
    u0[d] = floordiv(X0[d] - a.c[d], a.sg[d]) + a.omin[d];
  }
  // stage the box: 8 lattice rows (i1, i2) per wavefront per trip, 8 loads in
  // flight per lane
  {
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    const int nrow = a.LT[1] * a.LT[2];
    for (int i0 = ln; i0 < a.LT[0]; i0 += 64) {
      const int v0 = u0[0] + i0 - a.lo[0];
      const bool xin = v0 >= 0 && v0 < a.lv[0];
      for (int r0 = wv * 8; r0 < nrow; r0 += 8 * (kThreads / 64)) {
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int row = r0 + q, v1 = u0[1] + row % a.LT[1] - a.lo[1], v2 = u0[2] + row / a.LT[1] - a.lo[2];
          v[q] = (T)0;
          if (row < nrow && xin && v1 >= 0 && v1 < a.lv[1] && v2 >= 0 && v2 < a.lv[2])
            v[q] = a.Vl[v0 + (long)a.lv[0] * (v1 + (long)a.lv[1] * v2)];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (r0 + q < nrow) sv[i0 + a.LT[0] * (r0 + q)] = v[q];
      }
    }
  }
  __syncthreads();
  // one residue class per wavefront, four voxels per lane: each batch of 8
  // (offset, weight) pairs feeds 32 independent LDS reads and multiply-adds
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int k = wave; k < a.ncls; k += kThreads / 64) {
    const int rho[3] = {k % a.sg[0], (k / a.sg[0]) % a.sg[1], k / (a.sg[0] * a.sg[1])};
    int r[3], n[3], ub[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r[d] = ((rho[d] - (X0[d] - a.c[d])) % a.sg[d] + a.sg[d]) % a.sg[d];
      n[d] = r[d] < a.TD[d] ? (a.TD[d] - r[d] + a.sg[d] - 1) / a.sg[d] : 0;
      ub[d] = floordiv(X0[d] + r[d] - a.c[d], a.sg[d]) - u0[d];
    }
    const int nv = n[0] * n[1] * n[2];
    const int e0 = a.cls_start[k], e1 = a.cls_start[k + 1];
    for (int f0 = 0; f0 < nv; f0 += 4 * 64) {  // uniform bounds: entry loads stay scalar
      int bs[4], jj[4][3];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int f = f0 + lane + 64 * p < nv ? f0 + lane + 64 * p : 0;
        jj[p][0] = f % n[0];
        jj[p][1] = (f / n[0]) % n[1];
        jj[p][2] = f / (n[0] * n[1]);
        bs[p] = (ub[0] + jj[p][0]) + a.LT[0] * ((ub[1] + jj[p][1]) + a.LT[1] * (ub[2] + jj[p][2]));
      }
      T acc[4] = {(T)0, (T)0, (T)0, (T)0};
      int e = e0;
      for (; e + 8 <= e1; e += 8) {
        int o[8];
        T w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          o[q] = a.ldo[e + q];
          w[q] = from_bits(a.w[e + q]);
        }
        T v[4][8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int p = 0; p < 4; ++p) v[p][q] = sv[bs[p] + o[q]];
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int p = 0; p < 4; ++p) acc[p] += w[q] * v[p][q];
      }
      for (; e < e1; ++e) {
        const int o = a.ldo[e];
        const T w = from_bits(a.w[e]);
#pragma unroll
        for (int p = 0; p < 4; ++p) acc[p] += w * sv[bs[p] + o];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        if (f0 + lane + 64 * p >= nv) continue;
        const int x0 = X0[0] + r[0] + a.sg[0] * jj[p][0], x1 = X0[1] + r[1] + a.sg[1] * jj[p][1],
                  x2 = X0[2] + r[2] + a.sg[2] * jj[p][2];
        if (x0 < a.R[0] && x1 < a.R[1] && x2 < a.R[2]) {
          const long g = x0 + (long)a.R[0] * (x1 + (long)a.R[1] * x2);
          a.out[g] = a.mask[g] ? acc[p] : (T)0;
        }
      }
    }
  }
}

// W^T v of the boundary and far rows at the voxels they touch (written to its
// own field, which the residual adds); kWtvLanes lanes per voxel, quads of
// entries (boundary row - the voxel's first boundary row, as 16-bit deltas
// into the compact boundary v; pads are delta 0, weight 0) interleaved over
// the wavefront's voxels so that its loads read consecutive quads, then the
// voxel's far-row entries (far row index, weight).
template <typename T>
__device__ __forceinline__ void wtv_bnd_block(const WtvBnd<T>& a, int blk) {
  const int lane = threadIdx.x & (kWtvLanes - 1);
  const int k = blk * (kThreads / kWtvLanes) + (threadIdx.x / kWtvLanes);
  T acc = (T)0;
  int gj = 0;
  if (k < a.n) {
    gj = a.g[k];
    const int grp = k / (64 / kWtvLanes);  // the wavefront's voxels share a quad range
    const int j1 = a.gs[grp + 1] >> 2;
    const T* Vg = a.Vc + a.base[k];
    const uint2* d4 = reinterpret_cast<const uint2*>(a.d);
    const Quad<T>* w4 = reinterpret_cast<const Quad<T>*>(a.w);
    for (int j = (a.gs[grp] >> 2) + (int)(threadIdx.x & 63); j < j1; j += 64) {
      const uint2 dq = nt_u2(d4 + j);
      const Quad<T> wq = nt_quad(w4 + j);
      acc += wq.v[0] * Vg[lo16(dq.x)];
      acc += wq.v[1] * Vg[hi16(dq.x)];
      acc += wq.v[2] * Vg[lo16(dq.y)];
      acc += wq.v[3] * Vg[hi16(dq.y)];
    }
    for (int j = a.fp[k] + lane; j < a.fp[k + 1]; j += kWtvLanes) acc += a.fw[j] * a.Vf[a.fidx[j]];
  }
#pragma unroll
  for (int o = kWtvLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWtvLanes);
  if (k < a.n && lane == 0) a.out[gj] = acc;
}

// W^T v: blocks [0, gi) interior voxels, the rest boundary voxels
template <typename T>
__global__ void __launch_bounds__(kThreads) k_wtv_sb(WtvInt<T> ai, WtvBnd<T> ab, int gi) {
  if ((int)blockIdx.x < gi) wtv_int_block(ai, blockIdx.x);
  else wtv_bnd_block(ab, blockIdx.x - gi);
}

// mcmc_summary<T>::update (gaussian_process_model.h:205-224) at the output
// voxels with float64 accumulators.  Pass 1 gathers mu and the largest |mu|
// (atomicMax on the bit pattern of a non-negative double, which orders like
// its value); pass 2 adds mu, mu^2 and |mu| / max(max |mu|, 1e-6).
static __global__ void __launch_bounds__(64) k_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0) p[threadIdx.x] = 0ull;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_sum_gather(long n, const int* __restrict__ g, const T* __restrict__ R, T* __restrict__ out,
             unsigned long long* amax) {
  __shared__ double red[kThreads];
  double m = 0.0;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
    const T v = R[g[i]];
    out[i] = v;
    m = fmax(m, fabs((double)v));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(amax, (unsigned long long)__double_as_longlong(red[0]));
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_sum_acc(long n, const T* __restrict__ v, const unsigned long long* __restrict__ amax, double* __restrict__ s1,
          double* __restrict__ s2, double* __restrict__ sa) {
  const double M = fmax(__longlong_as_double((long long)amax[0]), 1e-6);
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
    const double x = (double)v[i];
    s1[i] += x;
    s2[i] += x * x;
    sa[i] += fabs(x) / M;
  }
}

// sum over the half spectrum of w |X|^2 a(k) with a = invlam (prior) or
// 1/(invlam + sig) (kinetic); sig < 0 selects the prior form.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_quad(long Nh, int H0, const cx<T>* __restrict__ X, const T* __restrict__ invlam, T sig, double* part) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < Nh; i += (long)gridDim.x * kThreads) {
    const T il = invlam[i];
    double a = (double)il;
    if (sig >= (T)0) a = il > (T)0 ? 1.0 / (double)(il + sig) : 0.0;
    const cx<T> x = X[i];
    acc += herm_w((int)(i % H0), H0) * a * ((double)x.x * x.x + (double)x.y * x.y);
  }
  red_write(acc, red, part, blockIdx.x);
}

// ||Yh - mu[I]||^2 over the region
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_sse_h(long n, const T* __restrict__ R, const T* __restrict__ Yg, const uint8_t* __restrict__ mask, double* part) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (long g = (long)blockIdx.x * kThreads + threadIdx.x; g < n; g += (long)gridDim.x * kThreads) {
    if (mask[g]) {
      const T d = Yg[g] - R[g];
      acc += (double)d * (double)d;
    }
  }
  red_write(acc, red, part, blockIdx.x);
}

// sum of R^2 over the data voxels (mask) of a region field
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_sumsq_masked(long n, const T* __restrict__ R, const uint8_t* __restrict__ mask, double* part) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (long g = (long)blockIdx.x * kThreads + threadIdx.x; g < n; g += (long)gridDim.x * kThreads)
    if (mask[g]) acc += (double)R[g] * (double)R[g];
  red_write(acc, red, part, blockIdx.x);
}

// lambda and 1/lambda rescaled from their base values (the tau^-2 update)
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_rescale_spectrum(long n, const T* __restrict__ il0, const double* __restrict__ l0, T* __restrict__ il,
                   double* __restrict__ l, T s_il, double s_l) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    il[i] = il0[i] * s_il;
    l[i] = l0[i] * s_l;
  }
}

struct FinalizeArgs { int nslot; int nblk[32]; long pstride; };
__global__ void __launch_bounds__(kThreads) k_finalize(FinalizeArgs a, const double* part, double* sums);

template <typename T>
__global__ void k_gather(long n, const int* __restrict__ idx, const T* __restrict__ src, T* __restrict__ dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

template <typename T>
__global__ void k_scale_spec(long n, cx<T>* X, const T* __restrict__ invlam, T s) {
  // X = s * X * invlam  (imaginary-chain prior gradient)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i] = X[i] * (s * invlam[i]);
}

}  // namespace drh
