// Mixed-radix 3D FFT passes for gfx950, staged through LDS.
//
// A 3D transform on the x-fastest grid is three 1D passes.  Each pass kernel
// loads a tile of TB lines (TB a power of two) into LDS, runs a Stockham
// autosort FFT over the tile (radices 2,3,4,5,7,8, butterflies in registers,
// one LDS round trip per stage), and writes the tile back.  Loads and stores
// go through functor objects so the spectral updates of the leapfrog step,
// the residual of the likelihood and the energy reductions are fused into
// the passes instead of costing extra sweeps over HBM.
//
//  * k_strided : transform along y or z.  The batch index b (contiguous in
//                memory) maps to consecutive lanes, so global loads/stores
//                are coalesced; LDS layout [k][b].
//  * k_r2c_x   : real rows along x -> half spectrum (M0/2+1), via a complex
//                FFT of length M0/2 on packed even/odd samples.
//  * k_c2r_x   : the inverse of k_r2c_x.
//
// Transforms are unnormalised (FFTW conventions): forward uses
// exp(-2 pi i nk/M), inverse exp(+2 pi i nk/M).  Input lengths may be
// shorter than M (zero padding) and outputs may be truncated, which prunes
// the passes to the image region of the embedding torus.
#pragma once

#include <cmath>
#include <type_traits>
#include <utility>
#include <vector>

#include "dr_common.hpp"

namespace drh {

constexpr int kMaxStages = 16;
// complex elements per lane per stage, and the tile (TB * M per workgroup):
// strided (y/z) passes batch 16, the x passes 8 (more workgroups in flight
// for their short rows)
constexpr int kEptStrided = 16;
constexpr int kEptX = 8;
constexpr int kTileStrided = kEptStrided * kThreads;
constexpr int kTileX = kEptX * kThreads;

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin
// by linear id.  xcd_tile maps the linear id to a tile index so that each
// XCD receives a contiguous range of tiles: neighbouring tiles, which share
// cache lines at their edges, then sit in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int lin, int total) {
  const int x = lin % kXcds, j = lin / kXcds;
  const int q = total / kXcds, r = total % kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

template <typename T>
struct AxisFFT {
  int M;                  // transform length
  int nst;                // number of stages
  int radix[kMaxStages];  // radix of each stage
  const cx<T>* tw;        // table exp(-2 pi i e / (M * tws)), length M * tws
  int tws;                // stride into the table
};

// ---------------------------------------------------------------- butterflies
template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 2, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    const cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
    const cx<T> t2 = v[1] + v[3], t3 = rot90<T, INV>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
  }
};

template <typename T, bool INV> struct Dft<T, 8, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, INV>::run(e);
    Dft<T, 4, INV>::run(o);
    const T h = (T)0.70710678118654752440;
    // w8^k for k = 1,2,3 (conjugated for the inverse)
    const cx<T> o1 = INV ? mk<T>((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
                         : mk<T>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    const cx<T> o2 = rot90<T, INV>(o[2]);
    const cx<T> o3 = INV ? mk<T>(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h)
                         : mk<T>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
};

// Radix 16 as 4 x 4: with n = 4 n1 + n2 and k = k1 + 4 k2,
//   X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[4 n1 + n2] W4^(n1 k1).
struct Tw16 {  // cos and sin of 2 pi j / 16, j = 0..9
  static constexpr double c[10] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                   0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                   -1.0, -0.92387953251128675613};
  static constexpr double s[10] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                                   1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                   0.0, -0.38268343236508977173};
};
template <typename T, bool INV> struct Dft<T, 16, INV> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> a[4][4];  // a[n2][k1]
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cx<T> t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<T, 4, INV>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        const int j = n2 * k1;  // W16^j = exp(-+ 2 pi i j / 16)
        if (j == 0) {
          a[n2][k1] = t[k1];
        } else {
          const cx<T> w = mk<T>((T)Tw16::c[j], INV ? (T)Tw16::s[j] : -(T)Tw16::s[j]);
          a[n2][k1] = cmul(t[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<T> u[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<T, 4, INV>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = u[k2];
    }
  }
};

// Odd radix R = 2m+1 via the symmetric/antisymmetric split:
//   X_q = x0 + sum_k cos(2 pi kq/R) t_k  -/+ i sum_k sin(2 pi kq/R) u_k
// with t_k = x_k + x_{R-k}, u_k = x_k - x_{R-k}.
template <int R> struct OddTab;
template <> struct OddTab<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <> struct OddTab<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct OddTab<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};

template <typename T, int R, bool INV> struct DftOdd {
  static __device__ __forceinline__ void run(cx<T>* v) {
    constexpr int m = R / 2;
    cx<T> t[m + 1], u[m + 1];
#pragma unroll
    for (int k = 1; k <= m; ++k) {
      t[k] = v[k] + v[R - k];
      u[k] = v[k] - v[R - k];
    }
    const cx<T> x0 = v[0];
    cx<T> s0 = x0;
#pragma unroll
    for (int k = 1; k <= m; ++k) s0 = s0 + t[k];
#pragma unroll
    for (int q = 1; q <= m; ++q) {
      cx<T> a = x0, b = mk<T>(0, 0);
#pragma unroll
      for (int k = 1; k <= m; ++k) {
        const int j = (k * q) % R;
        a = a + t[k] * (T)OddTab<R>::c[j];
        b = b + u[k] * (T)OddTab<R>::s[j];
      }
      // forward: X_q = a - i b, X_{R-q} = a + i b
      const cx<T> ib = mk<T>(-b.y, b.x);
      v[q] = INV ? a + ib : a - ib;
      v[R - q] = INV ? a - ib : a + ib;
    }
    v[0] = s0;
  }
};
template <typename T, bool INV> struct Dft<T, 3, INV> : DftOdd<T, 3, INV> {};
template <typename T, bool INV> struct Dft<T, 5, INV> : DftOdd<T, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 7, INV> : DftOdd<T, 7, INV> {};

// ---------------------------------------------------------------- Stockham stage
// LDS holds TB = 2^tbs lines; element k of line b sits at lds[k*P + b].
// Stage with span Ns: butterfly j reads j + r*M/R, twiddles by
// w^(r * (j mod Ns) * M/(Ns R)), writes (j - j mod Ns) R + j mod Ns + r Ns.
template <typename T, int R, bool INV, int EPT>
__device__ __forceinline__ void fft_stage(cx<T>* lds, int P, int tbs, const AxisFFT<T>& ax, int Ns) {
  const int TB = 1 << tbs;
  const int BF = ax.M / R;
  const int total = BF << tbs;
  const int estep = (ax.M / (Ns * R)) * ax.tws;
  constexpr int IT = (EPT + R - 1) / R;
  cx<T> v[IT][R];
  int dst[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int q = threadIdx.x + it * kThreads;
    dst[it] = -1;
    if (q < total) {
      const int b = q & (TB - 1);
      const int j = q >> tbs;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = lds[(j + r * BF) * P + b];
      const int jm = j % Ns;
      if (jm != 0) {
        const int e = jm * estep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = ax.tw[r * e];
          v[it][r] = INV ? cmulc(v[it][r], w) : cmul(v[it][r], w);
        }
      }
      Dft<T, R, INV>::run(v[it]);
      dst[it] = ((j - jm) * R + jm) * P + b;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    if (dst[it] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dst[it] + r * Ns * P] = v[it][r];
    }
  }
  __syncthreads();
}

// The same stages with the length, span and twiddle stride known at compile
// time (the hot passes at common lengths): all index arithmetic folds.
struct Factors {
  int n;
  int r[kMaxStages];
};
constexpr Factors factor_ct(int M) {  // the radix order of factorize()
  Factors f{};
  int m = M;
  while (m % 16 == 0) { f.r[f.n++] = 16; m /= 16; }
  while (m % 8 == 0 && m > 8) { f.r[f.n++] = 8; m /= 8; }
  if (m % 8 == 0) { f.r[f.n++] = 8; m /= 8; }
  while (m % 4 == 0) { f.r[f.n++] = 4; m /= 4; }
  while (m % 2 == 0) { f.r[f.n++] = 2; m /= 2; }
  while (m % 3 == 0) { f.r[f.n++] = 3; m /= 3; }
  while (m % 5 == 0) { f.r[f.n++] = 5; m /= 5; }
  while (m % 7 == 0) { f.r[f.n++] = 7; m /= 7; }
  return f;
}

template <typename T, int R, bool INV, int EPT, int M, int NS, int TWS>
__device__ __forceinline__ void fft_stage_ct(cx<T>* lds, int P, int tbs, const cx<T>* __restrict__ tw) {
  const int TB = 1 << tbs;
  constexpr int BF = M / R;
  const int total = BF << tbs;
  constexpr int estep = (M / (NS * R)) * TWS;
  constexpr int IT = (EPT + R - 1) / R;
  cx<T> v[IT][R];
  int dst[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int q = threadIdx.x + it * kThreads;
    dst[it] = -1;
    if (q < total) {
      const int b = q & (TB - 1);
      const int j = q >> tbs;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = lds[(j + r * BF) * P + b];
      const int jm = j % NS;
      if (NS > 1 && jm != 0) {
        const int e = jm * estep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = tw[r * e];
          v[it][r] = INV ? cmulc(v[it][r], w) : cmul(v[it][r], w);
        }
      }
      Dft<T, R, INV>::run(v[it]);
      dst[it] = ((j - jm) * R + jm) * P + b;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    if (dst[it] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[dst[it] + r * NS * P] = v[it][r];
    }
  }
  __syncthreads();
}

template <typename T, bool INV, int EPT, int M, int TWS, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_ct(cx<T>* lds, int P, int tbs, const cx<T>* tw) {
  constexpr Factors F = factor_ct(M);
  if constexpr (S < F.n) {
    constexpr int R = F.r[S];
    fft_stage_ct<T, R, INV, EPT, M, NS, TWS>(lds, P, tbs, tw);
    fft_ct<T, INV, EPT, M, TWS, S + 1, NS * R>(lds, P, tbs, tw);
  }
}

// runtime-length stages, or the compile-time pipeline when MC > 0 (MC == ax.M)
template <typename T, bool INV, int EPT, int MC, int TWS>
__device__ __forceinline__ void fft_run(cx<T>* lds, int P, int tbs, const AxisFFT<T>& ax);

template <typename T, bool INV, int EPT>
__device__ __forceinline__ void fft_lds(cx<T>* lds, int P, int tbs, const AxisFFT<T>& ax) {
  int Ns = 1;
  for (int s = 0; s < ax.nst; ++s) {
    const int R = ax.radix[s];
    switch (R) {
      case 16: fft_stage<T, 16, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 8: fft_stage<T, 8, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 4: fft_stage<T, 4, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 2: fft_stage<T, 2, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 3: fft_stage<T, 3, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 5: fft_stage<T, 5, INV, EPT>(lds, P, tbs, ax, Ns); break;
      case 7: fft_stage<T, 7, INV, EPT>(lds, P, tbs, ax, Ns); break;
      default: break;
    }
    Ns *= R;
  }
}

template <typename T, bool INV, int EPT, int MC, int TWS>
__device__ __forceinline__ void fft_run(cx<T>* lds, int P, int tbs, const AxisFFT<T>& ax) {
  if constexpr (MC > 0) fft_ct<T, INV, EPT, MC, TWS>(lds, P, tbs, ax.tw);
  else fft_lds<T, INV, EPT>(lds, P, tbs, ax);
}

// ---------------------------------------------------------------- pass kernels
// Functor interfaces:
//   strided:  cx<T> Pro::load(int o, int k, int b);   void Epi::store(int o, int k, int b, cx<T>);
//   r2c_x:    cx<T> Pro::load2(int row, int n);  (samples 2n, 2n+1 packed)   void Epi::store(row, k, cx<T>)
//   c2r_x:    cx<T> Pro::load(int row, int k);   void Epi::store2(int row, int n, cx<T> z)  (samples 2n, 2n+1)
// and void finish(double* red, int block) on both for optional reductions.

// Prologues that split into a side-effect-free fetch and an apply
// (kSplitLoad) fill a tile in groups of kFillGroup elements per thread: all
// loads of a group are issued before its first apply and LDS write, so a
// thread waits once per group instead of once per element (or more: the
// residual's mask decided which loads ran).  Plain single loads gain nothing
// from it (the y passes were 1-2 us slower).
template <class P, class = void> struct is_split_load : std::false_type {};
template <class P> struct is_split_load<P, std::void_t<decltype(P::kSplitLoad)>> : std::bool_constant<P::kSplitLoad> {};
constexpr int kFillGroup = 4;
// fetch(idx) must be safe for every idx in [0, total); put(idx, v) stores
// element idx.  Elements are put in increasing idx per thread, as the plain
// strided loop does, so reductions inside put keep their order.
template <int G, class Fetch, class Put>
__device__ __forceinline__ void fill_grouped(int total, Fetch&& fetch, Put&& put) {
  using V = decltype(fetch(0));
  for (int i0 = threadIdx.x; i0 < total; i0 += G * kThreads) {
    V v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = fetch(min(i0 + u * kThreads, total - 1));
#pragma unroll
    for (int u = 0; u < G; ++u)
      if (i0 + u * kThreads < total) put(i0 + u * kThreads, v[u]);
  }
}

// Per-iteration scalars (step sizes, noise precisions, noise stream ids) of
// functors built for a captured HIP graph are read from device memory when a
// kernel starts: a functor with a bind() member refreshes its by-value
// fields there; functors without one are left alone.
template <class F, class = void>
struct HasBind : std::false_type {};
template <class F>
struct HasBind<F, std::void_t<decltype(std::declval<F&>().bind())>> : std::true_type {};
template <class F>
__device__ __forceinline__ void bind_scalars(F& f) {
  if constexpr (HasBind<F>::value) f.bind();
}

// An epilogue with a State type and fetch(o, k, b) / store(o, k, b, v, state)
// gets the state of 4 modes per thread fetched before they are updated.
template <class F, class = void>
struct HasFetch : std::false_type {};
template <class F>
struct HasFetch<F, std::void_t<typename F::State>> : std::true_type {};

template <typename T, bool INV, class Pro, class Epi, int EPT = kEptStrided, int MC = 0>
__global__ void __launch_bounds__(kThreads)
k_strided(AxisFFT<T> ax, int NB, int k_in, int k_out, int tbs, Pro pro, Epi epi) {
  bind_scalars(pro);
  bind_scalars(epi);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[kThreads];
  const int TB = 1 << tbs;
  const int tile = xcd_tile(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  const int o = tile / gridDim.x;
  const int b0 = (tile % gridDim.x) * TB;
  const int M = MC > 0 ? MC : ax.M;
  for (int idx = threadIdx.x; idx < M * TB; idx += kThreads) {
    const int b = idx & (TB - 1), k = idx >> tbs;
    cx<T> v = mk<T>(0, 0);
    if (k < k_in && b0 + b < NB) v = pro.load(o, k, b0 + b);
    lds[k * TB + b] = v;
  }
  __syncthreads();
  fft_run<T, INV, EPT, MC, 1>(lds, TB, tbs, ax);
  if constexpr (HasFetch<Epi>::value) {  // same modes, same order per thread
    for (int idx0 = threadIdx.x; idx0 < k_out * TB; idx0 += 4 * kThreads) {
      typename Epi::State st[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = idx0 + u * kThreads, b = idx & (TB - 1), k = idx >> tbs;
        if (idx < k_out * TB && b0 + b < NB) st[u] = epi.fetch(o, k, b0 + b);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = idx0 + u * kThreads, b = idx & (TB - 1), k = idx >> tbs;
        if (idx < k_out * TB && b0 + b < NB) epi.store(o, k, b0 + b, lds[k * TB + b], st[u]);
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < k_out * TB; idx += kThreads) {
      const int b = idx & (TB - 1), k = idx >> tbs;
      if (b0 + b < NB) epi.store(o, k, b0 + b, lds[k * TB + b]);
    }
  }
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  pro.finish(red, blk);
  epi.finish(red, blk);
}

// Forward transform, pointwise update, inverse transform of the same tile:
// the leapfrog kick and drift in the spectral domain sit between the forward
// z pass of one gradient and the inverse z pass of the next.
//   cx<T> Mid::apply(int k, int b, cx<T> v)  for every k < M.
template <typename T, class Pro, class Mid, class Epi, int EPT = kEptStrided, int MC = 0>
__global__ void __launch_bounds__(kThreads)
k_strided2(AxisFFT<T> ax, int NB, int k_in, int k_out, int tbs, Pro pro, Mid mid, Epi epi) {
  bind_scalars(pro);
  bind_scalars(mid);
  bind_scalars(epi);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[kThreads];
  const int TB = 1 << tbs;
  // XCD-contiguous tiles: neighbouring tiles share the cache lines of the
  // 4-byte 1/lambda rows, so both halves are fetched into one L2
  const int b0 = xcd_tile(blockIdx.x, gridDim.x) * TB;
  const int M = MC > 0 ? MC : ax.M;
  for (int idx = threadIdx.x; idx < M * TB; idx += kThreads) {
    const int b = idx & (TB - 1), k = idx >> tbs;
    cx<T> v = mk<T>(0, 0);
    if (k < k_in && b0 + b < NB) v = pro.load(0, k, b0 + b);
    lds[k * TB + b] = v;
  }
  __syncthreads();
  fft_run<T, false, EPT, MC, 1>(lds, TB, tbs, ax);
  for (int idx = threadIdx.x; idx < M * TB; idx += kThreads) {
    const int b = idx & (TB - 1), k = idx >> tbs;
    if (b0 + b < NB) lds[k * TB + b] = mid.apply(k, b0 + b, lds[k * TB + b]);
  }
  __syncthreads();
  fft_run<T, true, EPT, MC, 1>(lds, TB, tbs, ax);
  for (int idx = threadIdx.x; idx < k_out * TB; idx += kThreads) {
    const int b = idx & (TB - 1), k = idx >> tbs;
    if (b0 + b < NB) epi.store(0, k, b0 + b, lds[k * TB + b]);
  }
  const int blk = blockIdx.x;
  pro.finish(red, blk);
  mid.finish(red, blk);
  epi.finish(red, blk);
}

// k_strided2 for 16-column tiles of a compile-time length MC: each thread
// fetches the state of the first NP of its MC * 16 / kThreads modes
// (Mid::fetch) before the tile's load and forward transform, so that those
// loads are in flight during the transform (NP bounds the registers held
// across the transform); the rest are fetched in groups of 4 after it.  The modes
// are the ones k_strided2's Mid loop gives the thread: column b = thread
// mod 16, rows thread / 16 + e * kThreads / 16.
template <typename T, class Pro, class Mid, class Epi, int EPT, int MC>
__global__ void __launch_bounds__(kThreads)
k_strided2_pf(AxisFFT<T> ax, int NB, int k_in, int k_out, Pro pro, Mid mid, Epi epi) {
  bind_scalars(pro);
  bind_scalars(mid);
  bind_scalars(epi);
  constexpr int TB = 16, tbs = 4, KS = kThreads / TB, NE = MC / KS;
  constexpr int NP = NE < 32 / (int)sizeof(T) ? NE : 32 / (int)sizeof(T);
  static_assert(NE * KS == MC, "MC must be a multiple of kThreads / 16");
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[kThreads];
  const int b0 = xcd_tile(blockIdx.x, gridDim.x) * TB;
  const int bt = threadIdx.x & (TB - 1), kt = threadIdx.x >> tbs;
  const bool in = b0 + bt < NB;
  const int bf = in ? b0 + bt : b0;  // b0 < NB: the clamped fetch stays in bounds
  typename Mid::State st[NP];
#pragma unroll
  for (int e = 0; e < NP; ++e) st[e] = mid.fetch(kt + e * KS, bf);
  // the tile's k_in input rows in groups of 4 loads per thread (lds index =
  // k * TB + b = idx), then the zero rows; Pro::load has no side effects
  fill_grouped<4>(
      k_in * TB, [&](int idx) { return pro.load(0, idx >> tbs, min(b0 + (idx & (TB - 1)), NB - 1)); },
      [&](int idx, cx<T> v) { lds[idx] = b0 + (idx & (TB - 1)) < NB ? v : mk<T>(0, 0); });
  for (int idx = k_in * TB + threadIdx.x; idx < MC * TB; idx += kThreads) lds[idx] = mk<T>(0, 0);
  __syncthreads();
  fft_run<T, false, EPT, MC, 1>(lds, TB, tbs, ax);
#pragma unroll
  for (int e = 0; e < NP; ++e) {
    const int k = kt + e * KS;
    if (in) lds[k * TB + bt] = mid.apply(k, b0 + bt, lds[k * TB + bt], st[e]);
  }
#pragma unroll
  for (int e0 = NP; e0 < NE; e0 += 4) {
    typename Mid::State sr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e0 + u < NE) sr[u] = mid.fetch(kt + (e0 + u) * KS, bf);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = kt + (e0 + u) * KS;
      if (e0 + u < NE && in) lds[k * TB + bt] = mid.apply(k, b0 + bt, lds[k * TB + bt], sr[u]);
    }
  }
  __syncthreads();
  fft_run<T, true, EPT, MC, 1>(lds, TB, tbs, ax);
  for (int idx = threadIdx.x; idx < k_out * TB; idx += kThreads) {
    const int b = idx & (TB - 1), k = idx >> tbs;
    if (b0 + b < NB) epi.store(0, k, b0 + b, lds[k * TB + b]);
  }
  const int blk = blockIdx.x;
  pro.finish(red, blk);
  mid.finish(red, blk);
  epi.finish(red, blk);
}

// Real rows of length M0 (first n_in samples nonzero) -> M0/2+1 complex.
// ax describes the complex FFT of length Mh = M0/2 with tws = 2 into the
// length-M0 table, so ax.tw[k] = exp(-2 pi i k / M0).
template <typename T, class Pro, class Epi, int EPT = kEptX, int MC = 0>
__global__ void __launch_bounds__(kThreads)
k_r2c_x(AxisFFT<T> ax, int nrows, int tbs, Pro pro, Epi epi) {
  bind_scalars(pro);
  bind_scalars(epi);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[kThreads];
  const int TB = 1 << tbs;
  const int P = TB + 1;
  const int Mh = MC > 0 ? MC : ax.M;
  const int row0 = blockIdx.x * TB;
  if constexpr (MC > 0 && is_split_load<Pro>::value) {
    fill_grouped<kFillGroup>(
        TB * Mh,
        [&](int idx) {
          const int l = idx / Mh, n = idx - l * Mh;
          return pro.fetch2(min(row0 + l, nrows - 1), n);
        },
        [&](int idx, const typename Pro::Raw& q) {
          const int l = idx / Mh, n = idx - l * Mh;
          lds[n * P + l] = row0 + l < nrows ? pro.apply2(q, n) : mk<T>(0, 0);
        });
  } else {
    for (int idx = threadIdx.x; idx < TB * Mh; idx += kThreads) {
      const int l = idx / Mh, n = idx - l * Mh;
      const int row = row0 + l;
      cx<T> v = mk<T>(0, 0);
      if (row < nrows) v = pro.load2(row, n);
      lds[n * P + l] = v;
    }
  }
  __syncthreads();
  fft_run<T, false, EPT, MC, 2>(lds, P, tbs, ax);
  const int H = Mh + 1;
  for (int idx = threadIdx.x; idx < TB * H; idx += kThreads) {
    const int l = idx / H, k = idx - l * H;
    const int row = row0 + l;
    if (row < nrows) {
      const cx<T> zk = lds[(k == Mh ? 0 : k) * P + l];
      const cx<T> zc = conj(lds[(k == 0 ? 0 : Mh - k) * P + l]);
      const cx<T> e = (zk + zc) * (T)0.5;
      const cx<T> d = zk - zc;
      const cx<T> od = mk<T>(d.y * (T)0.5, -d.x * (T)0.5);  // (zk - zc) / (2i)
      const cx<T> w = ax.tw[k];                              // exp(-2 pi i k / M0)
      epi.store(row, k, e + cmul(w, od));
    }
  }
  const int blk = blockIdx.x;
  pro.finish(red, blk);
  epi.finish(red, blk);
}

// M0/2+1 complex per row -> real rows; the epilogue receives packed pairs
// (x[2n], x[2n+1]) = (Re z[n], Im z[n]) unscaled.
template <typename T, class Pro, class Epi, int EPT = kEptX, int MC = 0>
__global__ void __launch_bounds__(kThreads)
k_c2r_x(AxisFFT<T> ax, int nrows, int tbs, Pro pro, Epi epi) {
  bind_scalars(pro);
  bind_scalars(epi);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[kThreads];
  const int TB = 1 << tbs;
  const int P = TB + 1;
  const int Mh = MC > 0 ? MC : ax.M;
  const int H = Mh + 1;
  const int row0 = blockIdx.x * TB;
  for (int idx = threadIdx.x; idx < TB * H; idx += kThreads) {
    const int l = idx / H, k = idx - l * H;
    const int row = row0 + l;
    cx<T> v = mk<T>(0, 0);
    if (row < nrows) v = pro.load(row, k);
    lds[k * P + l] = v;
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[Mh-k]) + i (X[k] - conj X[Mh-k]) exp(+2 pi i k / M0)
  const int hp = Mh / 2 + 1;
  for (int idx = threadIdx.x; idx < TB * hp; idx += kThreads) {
    const int l = idx / hp, k = idx - l * hp;
    const cx<T> a = lds[k * P + l];
    const cx<T> c = lds[(Mh - k) * P + l];
    const cx<T> s1 = a + conj(c), d1 = a - conj(c);
    const cx<T> t1 = cmulc(d1, ax.tw[k]);
    const cx<T> zk = mk<T>(s1.x - t1.y, s1.y + t1.x);
    if (k > 0 && Mh - k != k) {
      const cx<T> s2 = c + conj(a), d2 = c - conj(a);
      const cx<T> t2 = cmulc(d2, ax.tw[Mh - k]);
      lds[(Mh - k) * P + l] = mk<T>(s2.x - t2.y, s2.y + t2.x);
    }
    lds[k * P + l] = zk;
  }
  __syncthreads();
  fft_run<T, true, EPT, MC, 2>(lds, P, tbs, ax);
  for (int idx = threadIdx.x; idx < TB * Mh; idx += kThreads) {
    const int l = idx / Mh, n = idx - l * Mh;
    const int row = row0 + l;
    if (row < nrows) epi.store2(row, n, lds[n * P + l]);
  }
  const int blk = blockIdx.x;
  pro.finish(red, blk);
  epi.finish(red, blk);
}

// ---------------------------------------------------------------- host plan helpers
std::vector<int> factorize(int M);

// twiddle table exp(-2 pi i e / M), e in [0, M), rounded from long double
template <typename T>
std::vector<cx<T>> twiddles(int M) {
  std::vector<cx<T>> t(M);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int e = 0; e < M; ++e) {
    const long double a = -2.0L * pi * (long double)e / (long double)M;
    t[e] = mk<T>((T)std::cos(a), (T)std::sin(a));
  }
  return t;
}

// Stage radices in factorize()'s order, into the descriptor's fixed array
// (no heap use, so the allocation-free entry points can build plans).
// order 1 runs the odd radices (3, 5, 7) first, order 2 reverses the default
// sequence; the compile-time pipelines ignore it.
template <typename T>
AxisFFT<T> make_axis(int fft_len, const cx<T>* tw, int tws, int order = 0) {
  AxisFFT<T> a{};
  a.M = fft_len;
  if (fft_len < 1) throw Error(DR_EDOMAIN, "FFT length must be positive");
  int m = fft_len;
  auto take = [&](int r) {
    if (a.nst >= kMaxStages) throw Error(DR_EDOMAIN, "FFT length needs too many stages");
    a.radix[a.nst++] = r;
    m /= r;
  };
  while (m % 16 == 0) take(16);
  while (m % 8 == 0 && m > 8) take(8);
  if (m % 8 == 0) take(8);
  while (m % 4 == 0) take(4);
  while (m % 2 == 0) take(2);
  for (int r : {3, 5, 7})
    while (m % r == 0) take(r);
  if (m != 1) throw Error(DR_EDOMAIN, "FFT length " + std::to_string(fft_len) + " has a prime factor > 7");
  if (order == 1) {
    int r[kMaxStages], n = 0;
    for (int i = 0; i < a.nst; ++i)
      if (a.radix[i] % 2) r[n++] = a.radix[i];
    for (int i = 0; i < a.nst; ++i)
      if (a.radix[i] % 2 == 0) r[n++] = a.radix[i];
    for (int i = 0; i < a.nst; ++i) a.radix[i] = r[i];
  } else if (order == 2) {
    for (int i = 0, j = a.nst - 1; i < j; ++i, --j) {
      const int t = a.radix[i];
      a.radix[i] = a.radix[j];
      a.radix[j] = t;
    }
  }
  a.tw = tw;
  a.tws = tws;
  return a;
}

// Largest power-of-two tile TB with TB*len <= tile_elems and the LDS bytes
// within 64 KiB; x-passes use a padded pitch TB+1 and len+1 rows.
int pick_tbs(int len, int rows_lds, int elem_bytes, bool padded, int max_lines, int tile_elems);

}  // namespace drh
