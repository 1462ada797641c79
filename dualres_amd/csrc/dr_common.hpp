// Shared definitions for the dualres HIP engine (gfx950 / CDNA4).
//
// Complex arithmetic on float2/double2-like structs, status/error plumbing
// for the C ABI, a counter-based Philox4x32-10 generator (Salmon, Moraes,
// Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) with
// Box-Muller normals, and deterministic block reductions in float64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/dualres_hip.h"

namespace drh {

// ---------------------------------------------------------------- errors
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DR_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e__ = (expr);                                                  \
    if (e__ != hipSuccess)                                                    \
      throw ::drh::Error(DR_EHIP, std::string(#expr ": ") + hipGetErrorString(e__)); \
  } while (0)

#define DR_REQUIRE(cond, code, msg)                                           \
  do {                                                                        \
    if (!(cond)) throw ::drh::Error((code), (msg));                           \
  } while (0)

void set_last_error(const std::string& m);

template <class F>
int guarded(F&& f) {
  try {
    f();
    return DR_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc& e) {
    set_last_error(std::string("allocation failed: ") + e.what());
    return DR_ENOMEM;
  } catch (const std::domain_error& e) {
    set_last_error(e.what());
    return DR_EDOMAIN;
  } catch (const std::logic_error& e) {
    set_last_error(e.what());
    return DR_ELOGIC;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return DR_ERUNTIME;
  } catch (...) {
    set_last_error("unknown error");
    return DR_ERUNTIME;
  }
}

// ---------------------------------------------------------------- complex
template <typename T> struct cx;
template <> struct alignas(8) cx<float> { float x, y; };
template <> struct alignas(16) cx<double> { double x, y; };

template <typename T> __host__ __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> __host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __host__ __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {  // a * conj(b)
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __host__ __device__ __forceinline__ cx<T> rot90(cx<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// ---------------------------------------------------------------- Philox
struct u4 { uint32_t a, b, c, d; };

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// Philox4x32 with 10 rounds; constants from the SC'11 paper.
__host__ __device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(M0, c.a), lo0 = M0 * c.a;
    const uint32_t hi1 = mulhi32(M1, c.c), lo1 = M1 * c.c;
    u4 n;
    n.a = hi1 ^ c.b ^ k0;
    n.b = lo1;
    n.c = hi0 ^ c.d ^ k1;
    n.d = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Two standard normals (float64) for counter (pair, stream): the pair index
// addresses elements 2*pair and 2*pair+1 of a flat field.
__host__ __device__ __forceinline__ void normal_pair(uint64_t seed, uint64_t stream, uint64_t pair,
                                                     double& z0, double& z1) {
  u4 c;
  c.a = (uint32_t)pair;
  c.b = (uint32_t)(pair >> 32);
  c.c = (uint32_t)stream;
  c.d = (uint32_t)(stream >> 32);
  const u4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  // 53-bit uniforms strictly inside (0, 1)
  const double u1 = ((double)(((uint64_t)(r.a >> 5) << 26) | (r.b >> 6)) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(((uint64_t)(r.c >> 5) << 26) | (r.d >> 6)) + 0.5) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  double s, co;
  sincospi(2.0 * u2, &s, &co);
  z0 = rad * co;
  z1 = rad * s;
}

// The Philox block of counter (pair, stream).
__host__ __device__ __forceinline__ u4 philox_block(uint64_t seed, uint64_t stream, uint64_t pair) {
  u4 c;
  c.a = (uint32_t)pair;
  c.b = (uint32_t)(pair >> 32);
  c.c = (uint32_t)stream;
  c.d = (uint32_t)(stream >> 32);
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Box-Muller in float32 arithmetic from two 32-bit words (24-bit uniforms
// strictly inside (0, 1), so |z| < 5.9): float32 logarithm and sine/cosine
// cost a fraction of the float64 ones.
__host__ __device__ __forceinline__ void box_muller_f(uint32_t w1, uint32_t w2, float& z0, float& z1) {
  const float u1 = ((float)(w1 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((float)(w2 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float rad = sqrtf(-2.0f * logf(u1));
  float s, co;
  sincospif(2.0f * u2, &s, &co);
  z0 = rad * co;
  z1 = rad * s;
}

// The pair in float32 for the float32 chain, from words a and c of the block.
__host__ __device__ __forceinline__ void normal_pair(uint64_t seed, uint64_t stream, uint64_t pair, float& z0,
                                                     float& z1) {
  const u4 r = philox_block(seed, stream, pair);
  box_muller_f(r.a, r.c, z0, z1);
}

// ---------------------------------------------------------------- separately rounded float ops
// HIP's __fmul_rn / __fadd_rn are plain '*' and '+', which hipcc's default
// -ffp-contract=fast fuses into v_fma_f32 once inlined.  The reference's
// distances are separately rounded float products and sums (x86-64 without
// FMA), so these carry the contraction switch off in their own bodies: the
// IR operations lose the contract flag and no later fusion can merge them.
__host__ __device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__host__ __device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__host__ __device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}

// ---------------------------------------------------------------- exact float sqrt
// Correctly rounded float square root.  The hardware estimate is within a
// couple of ulp; it is corrected with the exact midpoint tests
// (r - u/2)^2 <= s <= (r + u/2)^2, evaluated in double where the 25-bit
// midpoints square exactly.  Matches x86 sqrtss, which the reference uses.
__host__ __device__ __forceinline__ float sqrt_rn(float s) {
  if (!(s > 0.0f)) return sqrtf(s);
  float r = sqrtf(s);
  const double sd = (double)s;
#pragma unroll
  for (int it = 0; it < 3; ++it) {
    const double lo = 0.5 * ((double)r + (double)nextafterf(r, 0.0f));
    const double hi = 0.5 * ((double)r + (double)nextafterf(r, 3.0e38f));
    if (lo * lo > sd) r = nextafterf(r, 0.0f);
    else if (hi * hi < sd) r = nextafterf(r, 3.0e38f);
    else break;
  }
  return r;
}

// ---------------------------------------------------------------- reductions
// Deterministic block sum of one double per thread.  Uses shared scratch of
// blockDim.x doubles; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x >> 1; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

constexpr int kThreads = 256;

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace drh
