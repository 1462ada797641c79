// Sparse local-kriging matrix W on the GPU (kriging_matrix.h:41-343).
//
//  1. k_pattern (one wavefront per stdres voxel): the reference's neighbour
//     search over the offset ball P, in P order, with its flat-offset range
//     check; emits (bounded column, offset index, k') and a 128-bit
//     position-sensitive hash per row.
//  2. Host: rows with equal (count, hash) share one kriging system; equality
//     is then verified element by element on the device (k_verify).
//  3. Unique systems, sorted by size and padded with an identity block to a
//     multiple of 64, are factorised in batches by a blocked left-looking
//     float64 Cholesky over 64x64 tiles, two launches per tile
//     column: k_chol_diag subtracts the finished columns' products from the
//     diagonal tile, factors it and inverts the factor; k_chol_panel does
//     the same subtraction for the tiles below and multiplies them by that
//     inverse (the products on the FP64 matrix cores).  Then solved by tiles
//     with those inverses (k_solve_chol), then
//     normalised by their sum (kriging_matrix.h:318-321).  Systems whose
//     Cholesky meets a nonpositive pivot (K with its unit diagonal is
//     indefinite when tau^2 > 1) are factorised again by a blocked LU with
//     partial pivoting (k_getrf_panel / k_trsm_lu / k_gemm_lu_mfma) and
//     solved by k_solve_lu.
//  4. Rows gather their system's weights; rows whose weights sum to NaN stay
//     empty, as in the reference.
// K and k' entries are float32 (the reference evaluates rbf<float>), taken as
// the float32 rounding of the float64 kernel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "dr_common.hpp"

namespace drh {

namespace {

constexpr int NB = 64;

struct Geo {
  int nx, ny, nz;
  long nvox;
  int lo[3], bd[3];
  long nbidx;
  float map[16], qs[16], qh[16];
  double tau2, psi, nu;  // float-rounded parameters, evaluated in double
  float eps0;
  int nP;
};

__device__ __forceinline__ float dot4(const float* q, int r, const float* v) {
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) acc = add_rn(acc, mul_rn(q[4 * r + c], v[c]));
  return acc;
}
// host twin (x86-64 without FMA: separately rounded products and sums)
static float dot4_h(const float* q, int r, const float* v) {
  float acc = 0.0f;
  for (int c = 0; c < 4; ++c) {
    const volatile float prod = q[4 * r + c] * v[c];
    acc = acc + prod;
  }
  return acc;
}

__host__ __device__ __forceinline__ float kern_f(const Geo& g, float d) {
  return (float)(g.tau2 * exp(-g.psi * pow(fabs((double)d), g.nu)));
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// pass = 0: counts and hashes; pass = 1: write entries at rp[row]
__global__ void __launch_bounds__(256)
k_pattern(Geo g, int n_rows, const int* __restrict__ sijk, const int* __restrict__ P, const uint8_t* __restrict__ present,
          const int* __restrict__ bidx, int pass, int* __restrict__ count, uint64_t* __restrict__ h1,
          uint64_t* __restrict__ h2, int* __restrict__ oob, const long* __restrict__ rp, int* __restrict__ cols,
          int* __restrict__ pidx, float* __restrict__ kp) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;  // whole wavefront exits together
  float s4[4] = {(float)sijk[3 * row], (float)sijk[3 * row + 1], (float)sijk[3 * row + 2], 1.0f};
  int base[3];
  float xs[4];
#pragma unroll
  for (int r = 0; r < 3; ++r) base[r] = (int)truncf(dot4(g.map, r, s4));
#pragma unroll
  for (int r = 0; r < 4; ++r) xs[r] = dot4(g.qs, r, s4);
  int n = 0, noob = 0;
  uint64_t a1 = 0, a2 = 0;
  const long wbase = pass ? rp[row] : 0;
  for (int j0 = 0; j0 < g.nP; j0 += 64) {
    const int j = j0 + lane;
    bool valid = j < g.nP;
    int hx = 0, hy = 0, hz = 0;
    long b = -1;
    int col = -1;
    if (valid) {
      hx = base[0] + P[3 * j];
      hy = base[1] + P[3 * j + 1];
      hz = base[2] + P[3 * j + 2];
      const long off = (long)hz * g.ny * g.nx + (long)hy * g.nx + hx;
      valid = off >= 0 && off < g.nvox && present[off];
      if (valid) {
        b = (long)(hz - g.lo[2]) * g.bd[0] * g.bd[1] + (long)(hy - g.lo[1]) * g.bd[0] + (hx - g.lo[0]);
        if (b < 0 || b >= g.nbidx) {
          valid = false;
          ++noob;
        } else {
          col = bidx[b];
          valid = col >= 0;
        }
      }
    }
    float k = 0.0f;
    if (valid) {
      const float h4[4] = {(float)hx, (float)hy, (float)hz, 1.0f};
      float ss = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float d = sub_rn(xs[r], dot4(g.qh, r, h4));
        ss = add_rn(ss, mul_rn(d, d));
      }
      k = kern_f(g, add_rn(sqrt_rn(ss), g.eps0));
    }
    const uint64_t bal = __ballot(valid);
    const int before = __popcll(bal & ((1ULL << lane) - 1ULL));
    if (valid) {
      const int pos = n + before;
      if (pass) {
        cols[wbase + pos] = col;
        pidx[wbase + pos] = j;
        kp[wbase + pos] = k;
      } else {
        uint32_t kb;
        memcpy(&kb, &k, 4);
        const uint64_t key = ((uint64_t)(uint32_t)j << 32) ^ kb;
        a1 += mix64(key + 0x9E3779B97F4A7C15ULL * (uint64_t)(pos + 1));
        a2 += mix64((key * 0xD6E8FEB86659FD93ULL) ^ (uint64_t)(pos + 7) * 0xA0761D6478BD642FULL);
      }
    }
    n += __popcll(bal);
  }
  if (!pass) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a1 += __shfl_xor(a1, o, 64);
      a2 += __shfl_xor(a2, o, 64);
      noob += __shfl_xor(noob, o, 64);
    }
    if (lane == 0) {
      count[row] = n;
      h1[row] = a1;
      h2[row] = a2;
      oob[row] = noob;
    }
  }
}

// rows sharing a system must have identical (offset index, k') lists
__global__ void k_verify(int n_rows, const long* __restrict__ rp, const int* __restrict__ rep,
                         const int* __restrict__ pidx, const float* __restrict__ kp, int* __restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  const int r = rep[row];
  if (r == row) return;
  const long a = rp[row], b = rp[r], n = rp[row + 1] - rp[row];
  int mism = 0;
  for (long t = lane; t < n; t += 64) {
    if (pidx[a + t] != pidx[b + t]) mism = 1;
    if (__float_as_uint(kp[a + t]) != __float_as_uint(kp[b + t])) mism = 1;
  }
  const uint64_t bal = __ballot(mism != 0);
  if (lane == 0 && bal) bad[row] = 1;
}

struct SysDesc {
  long off;   // entry offset of the representative row
  int n;      // system size
  long sol;   // offset into the solution buffer
};

// A_s = [K[pidx, pidx] 0; 0 I] padded to np; b_s = [k'; 0].  lower != 0
// skips the 64x64 tiles above the diagonal, which the Cholesky never reads.
__global__ void k_assemble(const SysDesc* __restrict__ sd, int np, const float* __restrict__ K, int nP,
                           const int* __restrict__ pidx, const float* __restrict__ kp, double* __restrict__ A,
                           double* __restrict__ bv, int lower) {
  const int s = blockIdx.y;
  const SysDesc d = sd[s];
  const long npp = (long)np * np;
  double* As = A + (long)s * npp;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < npp; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t / np), j = (int)(t - (long)i * np);
    if (lower && j / NB > i / NB) continue;
    double v;
    if (i < d.n && j < d.n) v = (double)K[(long)pidx[d.off + i] * nP + pidx[d.off + j]];
    else v = (i == j) ? 1.0 : 0.0;
    As[t] = v;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < np; i += blockDim.x) bv[(long)s * np + i] = i < d.n ? (double)kp[d.off + i] : 0.0;
}

// FP64 matrix-core tile products.  Each wavefront owns a 32x32 quadrant of a
// 64x64 tile as 2x2 blocks of 16x16, accumulated by v_mfma_f64_16x16x4f64
// over 64 columns, 4 at a time.  Operand layout of the instruction (checked
// on the device with distinct per-lane values): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register t of lane l
// holds D[4 t + l / 16][l % 16].
typedef double dr_v4f64 __attribute__((ext_vector_type(4)));
typedef double Tile[NB][NB + 1];

struct TileLane {  // this thread's place in the tile products
  int r0, c0, li, lk;
  __device__ TileLane() {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    r0 = (wave >> 1) * 32;
    c0 = (wave & 1) * 32;
    li = lane & 15;
    lk = lane >> 4;
  }
  // tile row / column of accumulator register t of block (a, b)
  __device__ int row(int a, int t) const { return r0 + 16 * a + 4 * t + lk; }
  __device__ int col(int b) const { return c0 + 16 * b + li; }
};

// acc += X Y^T over 64 columns, X and Y row-major 64x64 tiles in LDS
__device__ __forceinline__ void mfma_xyt(const Tile& X, const Tile& Y, const TileLane& q, dr_v4f64 acc[2][2]) {
#pragma unroll 4
  for (int kk = 0; kk < NB; kk += 4) {
    double av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) av[a] = X[q.r0 + 16 * a + q.li][kk + q.lk];
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = Y[q.c0 + 16 * b + q.li][kk + q.lk];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
}

__device__ __forceinline__ void zero_acc(dr_v4f64 acc[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = dr_v4f64{0.0, 0.0, 0.0, 0.0};
}

// Left-looking Cholesky, tile (BI, BJ) of tile column j:
// cv = A[BI][BJ] - sum_{l<j} L[BI][l] L[BJ][l]^T, accumulated over the whole
// depth j * 64 in registers (a right-looking update reads and writes every
// trailing tile once per panel).  The tile pair of panel l + 1 is fetched
// into registers while the matrix cores work on panel l from LDS.  Returns
// with every thread past its last read of Pi and Pj.
__device__ __forceinline__ void col_update(const double* As, int np, int BI, int BJ, int j, Tile& Pi, Tile& Pj,
                                           const TileLane& q, double cv[2][2][4]) {
  const int lc = threadIdx.x & 63, lr = threadIdx.x >> 6;  // loads: column lc of rows lr + 4 r
  double gi[16], gj[16];
  auto fetch = [&](int l) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      gi[r] = As[(long)(BI + lr + 4 * r) * np + l * NB + lc];
      gj[r] = As[(long)(BJ + lr + 4 * r) * np + l * NB + lc];
    }
  };
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) cv[a][b][t] = As[(long)(BI + q.row(a, t)) * np + BJ + q.col(b)];
  dr_v4f64 acc[2][2];
  zero_acc(acc);
  if (j > 0) fetch(0);
  for (int l = 0; l < j; ++l) {
    __syncthreads();  // the previous panel's products are done with LDS
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      Pi[lr + 4 * r][lc] = gi[r];
      Pj[lr + 4 * r][lc] = gj[r];
    }
    __syncthreads();
    if (l + 1 < j) fetch(l + 1);
    mfma_xyt(Pi, Pj, q, acc);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) cv[a][b][t] -= acc[a][b][t];
  __syncthreads();
}

__device__ __forceinline__ void cv_to_lds(const double cv[2][2][4], const TileLane& q, Tile& T) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) T[q.row(a, t)][q.col(b)] = cv[a][b][t];
}

// value of lane l (uniform) of a double held across the wavefront
__device__ __forceinline__ double lane_value(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One wavefront factors the 64x64 tile T in place, lane c holding row c in
// registers: for j = 0..63, L[j][j] = sqrt(A[j][j]), L[c][j] = A[c][j] /
// L[j][j] for c > j, then A[c][m] -= L[c][j] L[m][j] for m > j (rows c < m
// update entries above the diagonal, which are never read).  Returns 1 on a
// nonpositive pivot.
__device__ __forceinline__ int potrf_wave(Tile& T) {
  const int c = threadIdx.x & 63;
  double a[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) a[m] = T[c][m];
  int bad = 0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const double d = lane_value(a[j], j);
    bad |= !(d > 0.0);
    const double sd = sqrt(d);
    a[j] = c == j ? sd : a[j] / sd;
#pragma unroll
    for (int m = j + 1; m < NB; ++m) a[m] -= a[j] * lane_value(a[j], m);
  }
#pragma unroll
  for (int m = 0; m < NB; ++m) T[c][m] = a[m];
  return bad;
}

// The inverse of the lower factor in T, by the whole workgroup: diagonal
// blocks of size s = 1, 2, 4, .. 32 are inverted from those of size s / 2 by
// [A 0; B C]^-1 = [A^-1 0; -C^-1 B A^-1  C^-1], the product B A^-1 kept for
// the second product in the block's unused upper triangle of W.  Leaves
// L^-1[i][c] at T[c][i] for i > c (the inverse's diagonal is 1 / L[i][i]);
// W is scratch.
__device__ __forceinline__ void trtri_block(Tile& T, Tile& W) {
  for (int t = threadIdx.x; t < NB; t += 256) W[t][t] = 1.0 / T[t][t];
  __syncthreads();
  for (int s = 1; s < NB; s *= 2) {
    const int n = (NB / (2 * s)) * s * s;  // elements of all blocks at this level
    for (int t = threadIdx.x; t < n; t += 256) {  // U[cc][r] = (B A^-1)[r][cc]
      const int b = (t / (s * s)) * 2 * s, r = (t / s) % s, cc = t % s;
      double v = 0.0;
      for (int k = cc; k < s; ++k) v += T[b + s + r][b + k] * W[b + k][b + cc];
      W[b + cc][b + s + r] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += 256) {  // X[r][cc] = -(C^-1 U^T)[r][cc]
      const int b = (t / (s * s)) * 2 * s, r = (t / s) % s, cc = t % s;
      double v = 0.0;
      for (int k = 0; k <= r; ++k) v += W[b + s + r][b + s + k] * W[b + cc][b + s + k];
      W[b + s + r][b + cc] = -v;
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NB * NB; t += 256) {
    const int i = t / NB, c = t % NB;
    if (i > c) T[c][i] = W[i][c];
  }
}

__device__ __forceinline__ void lds_to_tile(const Tile& T, double* As, int np, int BI, int BJ) {
  for (int t = threadIdx.x; t < NB * NB; t += 256) {
    const int i = t / NB, jj = t % NB;
    As[(long)(BI + i) * np + BJ + jj] = T[i][jj];
  }
}

// Cholesky, tile column j, diagonal tile: update, factor, invert
__global__ void __launch_bounds__(256) k_chol_diag(double* A, int np, int j, int* fail) {
  __shared__ Tile Pi, Pj;
  double* As = A + (long)blockIdx.x * np * np;
  const TileLane q;
  double cv[2][2][4];
  col_update(As, np, j * NB, j * NB, j, Pi, Pj, q, cv);
  cv_to_lds(cv, q, Pi);
  __syncthreads();
  if (threadIdx.x < 64) {
    const int bad = potrf_wave(Pi);
    if (bad && threadIdx.x == 0) fail[blockIdx.x] = 1;
  }
  __syncthreads();
  trtri_block(Pi, Pj);
  __syncthreads();
  lds_to_tile(Pi, As, np, j * NB, j * NB);
}

// Cholesky, tile column j, the tile rows i = j + 1 + blockIdx.y below the
// diagonal: update, then L[i][j] = C L[j][j]^-T as a product with the
// inverse k_chol_diag left above the diagonal (in place of a
// column-by-column substitution, 64 dependent steps per tile)
__global__ void __launch_bounds__(256) k_chol_panel(double* A, int np, int j) {
  __shared__ Tile Pi, Pj;
  double* As = A + (long)blockIdx.x * np * np;
  const int BI = (j + 1 + blockIdx.y) * NB, o = j * NB;
  const TileLane q;
  double cv[2][2][4];
  col_update(As, np, BI, o, j, Pi, Pj, q, cv);
  cv_to_lds(cv, q, Pi);
  const double* D = As + (long)o * np + o;
  for (int t = threadIdx.x; t < NB * NB; t += 256) {  // Pj[i][l] = L^-1[i][l]
    const int i = t / NB, l = t % NB;
    Pj[i][l] = l < i ? D[(long)l * np + i] : l == i ? 1.0 / D[(long)i * np + i] : 0.0;
  }
  __syncthreads();
  dr_v4f64 acc[2][2];
  zero_acc(acc);
  mfma_xyt(Pi, Pj, q, acc);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) As[(long)(BI + q.row(a, t)) * np + o + q.col(b)] = acc[a][b][t];
}

// ---------------------------------------------------------------- LU with partial pivoting
// Systems whose Cholesky meets a nonpositive pivot (the unit-diagonal K with
// tau^2 > 1 is indefinite) are solved by a blocked right-looking LU with
// partial pivoting on the device, in the role of the reference's
// colPivHouseholderQr (kriging_matrix.h:309-313): any factorisation of a
// nonsingular system gives its solution up to rounding.  Row swaps span the
// whole row (both factored and trailing columns) and the right-hand side, so
// the factors satisfy P A = L U and b is permuted in place.  The identity
// padding stays decoupled: its rows are zero in the data columns and never
// win a pivot search there.

// one workgroup per system: unblocked LU of the panel of columns [o, o + 64)
// over rows [o, np); fail = 1 when a pivot column is all zero (singular)
__global__ void __launch_bounds__(256) k_getrf_panel(double* A, double* bv, int np, int k, int* fail) {
  __shared__ double rv[256];
  __shared__ int ri[256];
  double* As = A + (long)blockIdx.x * np * np;
  double* bs = bv + (long)blockIdx.x * np;
  const int o = k * NB;
  for (int j = 0; j < NB; ++j) {
    const int c = o + j;
    // pivot: largest |A[i][c]| over i >= c, the first such row on ties
    double best = -1.0;
    int bi = c;
    for (int i = c + (int)threadIdx.x; i < np; i += 256) {
      const double v = fabs(As[(long)i * np + c]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    rv[threadIdx.x] = best;
    ri[threadIdx.x] = bi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) {
        const double v2 = rv[threadIdx.x + st];
        const int i2 = ri[threadIdx.x + st];
        if (v2 > rv[threadIdx.x] || (v2 == rv[threadIdx.x] && i2 < ri[threadIdx.x])) {
          rv[threadIdx.x] = v2;
          ri[threadIdx.x] = i2;
        }
      }
      __syncthreads();
    }
    const double pmax = rv[0];
    const int p = ri[0];
    __syncthreads();  // rv / ri are rewritten by the next column's search
    if (!(pmax > 0.0)) {  // same value in every thread: a uniform branch
      if (threadIdx.x == 0) fail[blockIdx.x] = 1;
      continue;
    }
    if (p != c) {
      for (int t = threadIdx.x; t < np; t += 256) {
        const double u = As[(long)c * np + t];
        As[(long)c * np + t] = As[(long)p * np + t];
        As[(long)p * np + t] = u;
      }
      if (threadIdx.x == 0) {
        const double u = bs[c];
        bs[c] = bs[p];
        bs[p] = u;
      }
      __syncthreads();
    }
    const double d = As[(long)c * np + c];
    for (int i = c + 1 + (int)threadIdx.x; i < np; i += 256) As[(long)i * np + c] /= d;
    __syncthreads();
    const int w = NB - j - 1;
    if (w > 0) {
      const long cnt = (long)(np - c - 1) * w;
      for (long t = threadIdx.x; t < cnt; t += 256) {
        const int i = c + 1 + (int)(t / w), cc = c + 1 + (int)(t % w);
        As[(long)i * np + cc] -= As[(long)i * np + c] * As[(long)c * np + cc];
      }
    }
    __syncthreads();
  }
}

// U12 = L11^-1 A12 for the tile column (k + 1 + blockIdx.y) of the panel's
// row block (L11 unit lower)
__global__ void __launch_bounds__(256) k_trsm_lu(double* A, int np, int k) {
  __shared__ double L[NB][NB + 1];
  __shared__ double X[NB][NB + 1];
  double* As = A + (long)blockIdx.x * np * np;
  const int o = k * NB, c0 = (k + 1 + blockIdx.y) * NB;
  for (int t = threadIdx.x; t < NB * NB; t += 256) {
    const int i = t / NB, j = t % NB;
    L[i][j] = As[(long)(o + i) * np + o + j];
    X[i][j] = As[(long)(o + i) * np + c0 + j];
  }
  __syncthreads();
  for (int j = 0; j < NB - 1; ++j) {
    const int m = NB - j - 1;
    for (int t = threadIdx.x; t < m * NB; t += 256) {
      const int i = j + 1 + t / NB, c = t % NB;
      X[i][c] -= L[i][j] * X[j][c];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NB * NB; t += 256) {
    const int i = t / NB, j = t % NB;
    As[(long)(o + i) * np + c0 + j] = X[i][j];
  }
}

// A22 -= L21 U12 over every trailing tile, blockIdx.y = bi * m + bj
// on the FP64 matrix cores
__global__ void __launch_bounds__(256) k_gemm_lu_mfma(double* A, int np, int k, int m) {
  __shared__ Tile Pi, Pj;
  double* As = A + (long)blockIdx.x * np * np;
  const int BI = (k + 1 + blockIdx.y / m) * NB, BJ = (k + 1 + blockIdx.y % m) * NB, o = k * NB;
  for (int t = threadIdx.x; t < NB * NB; t += 256) {
    const int i = t / NB, l = t % NB;
    Pi[i][l] = As[(long)(BI + i) * np + o + l];  // L21
    Pj[l][i] = As[(long)(o + i) * np + BJ + l];  // U12^T
  }
  __syncthreads();
  const TileLane q;
  dr_v4f64 acc[2][2];
  zero_acc(acc);
  mfma_xyt(Pi, Pj, q, acc);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) As[(long)(BI + q.row(a, t)) * np + BJ + q.col(b)] -= acc[a][b][t];
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Solves the factored system and normalises: w /= sum(w) (a zero sum counts
// as 1, kriging_matrix.h:318-321); writes the n weights to sol.  One
// workgroup per system; y holds the right-hand side, then the solution.
__device__ __forceinline__ double block_dot(double v, double* part) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  v = wsum(v);
  if (lane == 0) part[w] = v;
  __syncthreads();
  const double t = (part[0] + part[1]) + (part[2] + part[3]);
  __syncthreads();
  return t;
}

__device__ __forceinline__ void normalise_out(const double* y, int n, double* part, double* out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += y[i];
  double tot = block_dot(acc, part);
  if (tot == 0.0) tot = 1.0;
  for (int i = threadIdx.x; i < n; i += 256) out[i] = y[i] / tot;
}

// Cholesky: L L^T w = b by 64-row tiles, with the inverses of the diagonal
// tiles that k_chol_diag left above their diagonals.  Forward, tile k:
// r = b_k - sum_{l<k} L[k][l] y_l, y_k = L[k][k]^-1 r; backward:
// r = y_k - sum_{l>k} L[l][k]^T x_l, x_k = L[k][k]^-T r.  Tiles past the
// system's size are identity padding with a zero right-hand side and are
// skipped.
__global__ void __launch_bounds__(256) k_solve_chol(const double* A, int np, const SysDesc* __restrict__ sd,
                                                    const double* __restrict__ bv, double* __restrict__ sol) {
  extern __shared__ double y[];
  __shared__ double part[4][NB];
  __shared__ double rk[NB];
  const int s = blockIdx.x;
  const SysDesc d = sd[s];
  const double* As = A + (long)s * np * np;
  const int nt = (d.n + NB - 1) / NB, nn = nt * NB;
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < nn; i += 256) y[i] = bv[(long)s * np + i];
  __syncthreads();
  for (int k = 0; k < nt; ++k) {
    const int o = k * NB;
    const double* D = As + (long)o * np + o;
    double acc[16];  // rows o + 16 q + rr, the lanes over the columns
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) acc[rr] = 0.0;
    for (int c = lane; c < o; c += 64) {
      const double yc = y[c];
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) acc[rr] += As[(long)(o + 16 * q + rr) * np + c] * yc;
    }
    double mine = 0.0;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      const double v = wsum(acc[rr]);
      if (lane == rr) mine = v;
    }
    if (lane < 16) rk[16 * q + lane] = y[o + 16 * q + lane] - mine;
    __syncthreads();
    double v = 0.0;  // (L^-1 r)[lane], columns 16 q ..
#pragma unroll 4
    for (int c = 16 * q; c < 16 * q + 16; ++c) {
      const double lic = c < lane ? D[(long)c * np + lane] : c == lane ? 1.0 / D[(long)lane * np + lane] : 0.0;
      v += lic * rk[c];
    }
    part[q][lane] = v;
    __syncthreads();
    if (q == 0) y[o + lane] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    __syncthreads();
  }
  for (int k = nt - 1; k >= 0; --k) {
    const int o = k * NB;
    const double* D = As + (long)o * np + o;
    double v = 0.0;  // column o + lane, rows o + 64 + q, step 4
    for (int i = o + NB + q; i < nn; i += 4) v += As[(long)i * np + o + lane] * y[i];
    part[q][lane] = v;
    __syncthreads();
    if (q == 0) rk[lane] = y[o + lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
    __syncthreads();
    v = 0.0;  // (L^-T r)[lane], rows 16 q ..
#pragma unroll 4
    for (int i = 16 * q; i < 16 * q + 16; ++i) {
      const double lic = i > lane ? D[(long)lane * np + i] : i == lane ? 1.0 / D[(long)lane * np + lane] : 0.0;
      v += lic * rk[i];
    }
    part[q][lane] = v;
    __syncthreads();
    if (q == 0) y[o + lane] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    __syncthreads();
  }
  normalise_out(y, d.n, &part[0][0], sol + d.sol);
}

// LU: L U w = b with b already permuted, L unit lower and U upper in A; a
// system flagged singular gets NaN weights, which drop its rows as a NaN sum
// does.
__global__ void __launch_bounds__(256) k_solve_lu(const double* A, int np, const SysDesc* __restrict__ sd,
                                                  const double* __restrict__ bv, const int* __restrict__ fail,
                                                  double* __restrict__ sol) {
  extern __shared__ double y[];
  __shared__ double part[4];
  const int s = blockIdx.x;
  const SysDesc d = sd[s];
  const double* As = A + (long)s * np * np;
  const int n = d.n;
  if (fail[s]) {
    for (int i = threadIdx.x; i < n; i += 256) sol[d.sol + i] = __builtin_nan("");
    return;
  }
  for (int i = threadIdx.x; i < n; i += 256) y[i] = bv[(long)s * np + i];
  __syncthreads();
  for (int j = 0; j < n; ++j) {  // forward: row j of L against the solved y[0..j)
    const double* Lj = As + (long)j * np;
    double acc = 0.0;
    for (int i = threadIdx.x; i < j; i += 256) acc += Lj[i] * y[i];
    acc = block_dot(acc, part);
    if (threadIdx.x == 0) y[j] -= acc;
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; --j) {  // back: row j of U against the solved y(j..n)
    const double* Uj = As + (long)j * np;
    double acc = 0.0;
    for (int i = j + 1 + (int)threadIdx.x; i < n; i += 256) acc += Uj[i] * y[i];
    acc = block_dot(acc, part);
    if (threadIdx.x == 0) y[j] = (y[j] - acc) / Uj[j];
    __syncthreads();
  }
  normalise_out(y, n, part, sol + d.sol);
}

// vals[e] = sol[sys_sol[uid(row)] + (e - rp[row])]; nanrow[row] = 1 when a
// weight of the row is NaN
__global__ void k_gather_rows(int n_rows, const long* __restrict__ rp, const int* __restrict__ uid,
                              const long* __restrict__ usol, const double* __restrict__ sol, double* __restrict__ vals,
                              int* __restrict__ nanrow) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  const long a = rp[row], n = rp[row + 1] - a;
  int nan = 0;
  if (n > 0) {
    const long so = usol[uid[row]];
    for (long t = lane; t < n; t += 64) {
      const double v = sol[so + t];
      vals[a + t] = v;
      nan |= v != v;
    }
  }
  const uint64_t bal = __ballot(nan);
  if (lane == 0) nanrow[row] = bal != 0;
}

// kept rows moved to their compacted offsets rpn (dropped rows have none)
__global__ void k_compact_rows(int n_rows, const long* __restrict__ rp, const long* __restrict__ rpn,
                               const int* __restrict__ cols, const double* __restrict__ vals, int* __restrict__ cols_out,
                               double* __restrict__ vals_out) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_rows) return;
  const long a = rp[row], b = rpn[row], n = rpn[row + 1] - b;
  for (long t = lane; t < n; t += 64) {
    cols_out[b + t] = cols[a + t];
    vals_out[b + t] = vals[a + t];
  }
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  DBuf() = default;
  explicit DBuf(size_t n) { alloc(n); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void**)&p, n * sizeof(T)) != hipSuccess) {
      p = nullptr;
      throw Error(DR_ENOMEM, "hipMalloc failed in kriging build");
    }
  }
  void up(const T* h, size_t n, hipStream_t s) { DR_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s)); }
  void down(T* h, size_t n, hipStream_t s) { DR_HIP(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, s)); }
};

float fsq(float a) {
  const volatile float r = a * a;
  return r;
}

}  // namespace

// ---------------------------------------------------------------- host build
// The CSR stays on the device where it was built until it is copied out.
struct KrigingCSR {
  long nrow = 0, ncol = 0, n_unique = 0, nnz = 0;
  long n_lu = 0;  // unique systems solved by the pivoted LU (Cholesky failed)
  std::vector<int64_t> rp;
  DBuf<int32_t> cols;
  DBuf<double> vals;
};

static float norm3f(const float* v) {
  float s = 0.0f;
  for (int c = 0; c < 3; ++c) {
    const volatile float sq = v[c] * v[c];
    s = s + sq;
  }
  return sqrt_rn(s);
}

KrigingCSR kriging_build(const double* hr, const int hd[3], const float qh[16], const double* st, const int sdims[3],
                         const float qs[16], const double theta[3], double radius, int device) {
  DR_REQUIRE(hr && st && hd && sdims && qh && qs && theta, DR_ELOGIC, "kriging_build: null argument");
  DR_REQUIRE(radius >= 0, DR_EDOMAIN, "neighborhood_perturbation: radius cannot be negative");
  DR_HIP(hipSetDevice(device));
  const int nx = hd[0], ny = hd[1], nz = hd[2];
  const long nvox = (long)nx * ny * nz;
  // neighbourhood offsets (kriging_matrix.h:63-106)
  float vd[3];
  for (int c = 0; c < 3; ++c) {
    float s = 0.0f;
    for (int r = 0; r < 4; ++r) s += qh[4 * r + c] * qh[4 * r + c];
    vd[c] = sqrt_rn(s);
  }
  const float rad = (float)radius;
  int rng[3];
  for (int c = 0; c < 3; ++c) rng[c] = (int)(rad / vd[c]);
  std::vector<int> P;
  std::vector<float> X;
  for (int z = -rng[2]; z <= rng[2]; ++z)
    for (int y = -rng[1]; y <= rng[1]; ++y)
      for (int x = -rng[0]; x <= rng[0]; ++x) {
        const float p4[4] = {(float)x, (float)y, (float)z, 0.0f};
        float xx[3];
        for (int r = 0; r < 3; ++r) xx[r] = dot4_h(qh, r, p4);
        if (norm3f(xx) <= rad) {
          P.insert(P.end(), {x, y, z});
          X.insert(X.end(), {xx[0], xx[1], xx[2]});
        }
      }
  const int nP = (int)(P.size() / 3);
  DR_REQUIRE(nP > 0, DR_ELOGIC, "neighborhood_perturbation: no perturbations within radius");
  DR_REQUIRE(nP <= 5000, DR_ELOGIC, "get_sparse_kriging_array_data: neighborhood >5e3 voxels");
  Geo g{};
  g.tau2 = (double)(float)theta[0];
  g.psi = (double)(float)theta[1];
  g.nu = (double)(float)theta[2];
  g.eps0 = 1e-5f;
  g.nP = nP;
  // K (kriging_matrix.h:111-138): unit diagonal
  std::vector<float> K((size_t)nP * nP);
  for (int i = 0; i < nP; ++i) {
    K[(size_t)i * nP + i] = 1.0f;
    for (int j = 0; j < i; ++j) {
      const float d[3] = {X[3 * i] - X[3 * j], X[3 * i + 1] - X[3 * j + 1], X[3 * i + 2] - X[3 * j + 2]};
      const float ss = (fsq(d[0]) + fsq(d[1])) + fsq(d[2]);
      const float v = kern_f(g, sqrt_rn(ss));
      K[(size_t)i * nP + j] = v;
      K[(size_t)j * nP + i] = v;
    }
  }
  // highres bounding box and bounded nonzero index (nifti_manipulation.inl:210-275)
  auto pres = [](double v) { return !(std::isnan(v) || v == 0.0); };
  std::vector<uint8_t> present(nvox);
  int lo[3] = {nx, ny, nz}, hi[3] = {-1, -1, -1};
  long nnz_hr = 0;
  for (long f = 0; f < nvox; ++f) {
    present[f] = pres(hr[f]);
    if (present[f]) {
      const int c[3] = {(int)(f % nx), (int)((f / nx) % ny), (int)(f / ((long)nx * ny))};
      for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], c[a]); hi[a] = std::max(hi[a], c[a]); }
      ++nnz_hr;
    }
  }
  DR_REQUIRE(nnz_hr > 0, DR_EDOMAIN, "kriging_build: high-resolution image has no nonzero voxels");
  int bd[3];
  for (int a = 0; a < 3; ++a) bd[a] = hi[a] - lo[a] + 1;
  const long nb = (long)bd[0] * bd[1] * bd[2];
  std::vector<int> bidx(nb);
  {
    int cnt = 0;
    for (int k = 0; k < bd[2]; ++k)
      for (int j = 0; j < bd[1]; ++j)
        for (int i = 0; i < bd[0]; ++i) {
          const long f = (long)(k + lo[2]) * nx * ny + (long)(j + lo[1]) * nx + (i + lo[0]);
          bidx[(long)k * bd[0] * bd[1] + (long)j * bd[0] + i] = present[f] ? cnt++ : -1;
        }
  }
  // stdres nonzero ijk (nifti_manipulation.inl:334-360)
  const long snvox = (long)sdims[0] * sdims[1] * sdims[2];
  std::vector<int> sijk;
  for (long f = 0; f < snvox; ++f)
    if (pres(st[f])) {
      sijk.push_back((int)(f % sdims[0]));
      sijk.push_back((int)((f / sdims[0]) % sdims[1]));
      sijk.push_back((int)(f / ((long)sdims[0] * sdims[1])));
    }
  const int n_rows = (int)(sijk.size() / 3);
  DR_REQUIRE(n_rows > 0, DR_EDOMAIN, "kriging_build: standard-resolution image has no nonzero voxels");
  // map = Qh^-1 Qs (float32 rounding of the float64 product)
  {
    double a[16], inv[16];
    for (int i = 0; i < 16; ++i) a[i] = qh[i];
    // Gauss-Jordan inverse of the 4x4
    double m[4][8];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? a[4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
      int p = c;
      for (int r = c + 1; r < 4; ++r)
        if (std::fabs(m[r][c]) > std::fabs(m[p][c])) p = r;
      DR_REQUIRE(m[p][c] != 0.0, DR_EDOMAIN, "kriging_build: singular high-resolution qform");
      for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[p][j]);
      const double d = m[c][c];
      for (int j = 0; j < 8; ++j) m[c][j] /= d;
      for (int r = 0; r < 4; ++r)
        if (r != c) {
          const double f = m[r][c];
          for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) inv[4 * i + j] = m[i][4 + j];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += inv[4 * i + k] * (double)qs[4 * k + j];
        g.map[4 * i + j] = (float)s;
      }
  }
  for (int i = 0; i < 16; ++i) { g.qs[i] = qs[i]; g.qh[i] = qh[i]; }
  g.nx = nx; g.ny = ny; g.nz = nz; g.nvox = nvox;
  for (int a = 0; a < 3; ++a) { g.lo[a] = lo[a]; g.bd[a] = bd[a]; }
  g.nbidx = nb;

  hipStream_t s;
  DR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct SG { hipStream_t s; ~SG() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } sg{s};
  DBuf<int> d_sijk(sijk.size()), d_P(P.size()), d_bidx(nb), d_count(n_rows), d_oob(n_rows);
  DBuf<uint8_t> d_pres(nvox);
  DBuf<uint64_t> d_h1(n_rows), d_h2(n_rows);
  d_sijk.up(sijk.data(), sijk.size(), s);
  d_P.up(P.data(), P.size(), s);
  d_bidx.up(bidx.data(), nb, s);
  d_pres.up(present.data(), nvox, s);
  const int gblk = cdiv(n_rows, 4);
  hipLaunchKernelGGL(k_pattern, dim3(gblk), dim3(256), 0, s, g, n_rows, d_sijk.p, d_P.p, d_pres.p, d_bidx.p, 0,
                     d_count.p, d_h1.p, d_h2.p, d_oob.p, (const long*)nullptr, (int*)nullptr, (int*)nullptr,
                     (float*)nullptr);
  DR_HIP(hipGetLastError());
  std::vector<int> cnt(n_rows), oobv(n_rows);
  std::vector<uint64_t> h1(n_rows), h2(n_rows);
  d_count.down(cnt.data(), n_rows, s);
  d_h1.down(h1.data(), n_rows, s);
  d_h2.down(h2.data(), n_rows, s);
  d_oob.down(oobv.data(), n_rows, s);
  DR_HIP(hipStreamSynchronize(s));
  std::vector<long> rp(n_rows + 1, 0);
  for (int i = 0; i < n_rows; ++i) rp[i + 1] = rp[i] + cnt[i];
  const long nnz = rp[n_rows];
  DR_REQUIRE(nnz < (1L << 31), DR_EDOMAIN, "kriging_build: too many nonzeros");
  // dedup
  struct Key { uint64_t a, b; int n; bool operator==(const Key& o) const { return a == o.a && b == o.b && n == o.n; } };
  struct KH { size_t operator()(const Key& k) const { return (size_t)(k.a ^ (k.b * 31) ^ (uint64_t)k.n); } };
  std::unordered_map<Key, int, KH> seen;
  std::vector<int> rep(n_rows), uid(n_rows), urow;
  for (int i = 0; i < n_rows; ++i) {
    if (cnt[i] == 0) { rep[i] = i; uid[i] = -1; continue; }
    Key k{h1[i], h2[i], cnt[i]};
    auto it = seen.find(k);
    if (it == seen.end()) {
      const int u = (int)urow.size();
      seen.emplace(k, u);
      urow.push_back(i);
      rep[i] = i;
      uid[i] = u;
    } else {
      uid[i] = it->second;
      rep[i] = urow[it->second];
    }
  }
  const int nU = (int)urow.size();
  DBuf<long> d_rp(n_rows + 1);
  DBuf<int> d_cols(nnz), d_pidx(nnz), d_rep(n_rows), d_bad(n_rows), d_uid(n_rows);
  DBuf<float> d_kp(nnz);
  d_rp.up(rp.data(), n_rows + 1, s);
  d_rep.up(rep.data(), n_rows, s);
  DR_HIP(hipMemsetAsync(d_bad.p, 0, n_rows * sizeof(int), s));
  hipLaunchKernelGGL(k_pattern, dim3(gblk), dim3(256), 0, s, g, n_rows, d_sijk.p, d_P.p, d_pres.p, d_bidx.p, 1,
                     (int*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr, (int*)nullptr, d_rp.p, d_cols.p,
                     d_pidx.p, d_kp.p);
  DR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_verify, dim3(gblk), dim3(256), 0, s, n_rows, d_rp.p, d_rep.p, d_pidx.p, d_kp.p, d_bad.p);
  DR_HIP(hipGetLastError());
  std::vector<int> bad(n_rows);
  d_bad.down(bad.data(), n_rows, s);
  DR_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < n_rows; ++i)
    DR_REQUIRE(!bad[i], DR_ERUNTIME, "kriging_build: hash collision between distinct kriging systems");

  // unique systems sorted by size, batched by padded size; Cholesky first,
  // then LU with partial pivoting for the systems whose Cholesky failed
  std::vector<int> order(nU);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return cnt[urow[a]] > cnt[urow[b]]; });
  std::vector<long> usol(nU);
  long solsz = 0;
  for (int u = 0; u < nU; ++u) { usol[u] = solsz; solsz += cnt[urow[u]]; }
  DBuf<double> d_sol(solsz);
  DBuf<float> d_K((size_t)nP * nP);
  d_K.up(K.data(), (size_t)nP * nP, s);
  const double budget = 6.0e9;  // bytes of factor scratch per batch
  DBuf<double> d_A, d_b;
  DBuf<SysDesc> d_sd;
  DBuf<int> d_fail;
  size_t capA = 0, capB = 0, capS = 0;
  auto solve_all = [&](const std::vector<int>& uids, bool lu) {
    std::vector<int> failed;
    size_t pos = 0;
    while (pos < uids.size()) {
      const int n0 = cnt[urow[uids[pos]]];
      const int np = ((n0 + NB - 1) / NB) * NB;
      const long per = (long)np * np * 8;
      const int bmax = (int)std::max<long>(1, std::min<long>(4096, (long)(budget / per)));
      std::vector<SysDesc> sd;
      std::vector<int> bu;
      while (pos < uids.size() && (int)sd.size() < bmax) {
        const int u = uids[pos];
        sd.push_back(SysDesc{rp[urow[u]], cnt[urow[u]], usol[u]});
        bu.push_back(u);
        ++pos;
      }
      const int B = (int)sd.size();
      if ((size_t)B * np * np > capA) { capA = (size_t)B * np * np; d_A.alloc(capA); }
      if ((size_t)B * np > capB) { capB = (size_t)B * np; d_b.alloc(capB); }
      if ((size_t)B > capS) { capS = B; d_sd.alloc(capS); d_fail.alloc(capS); }
      d_sd.up(sd.data(), B, s);
      DR_HIP(hipMemsetAsync(d_fail.p, 0, B * sizeof(int), s));
      const int gx = std::max(1, std::min(cdiv((long)np * np, 256), 256));
      hipLaunchKernelGGL(k_assemble, dim3(gx, B), dim3(256), 0, s, d_sd.p, np, d_K.p, nP, d_pidx.p, d_kp.p, d_A.p,
                         d_b.p, lu ? 0 : 1);
      DR_HIP(hipGetLastError());
      const int nt = np / NB;
      for (int k = 0; k < nt; ++k) {
        const int m = nt - k - 1;
        if (lu) {
          hipLaunchKernelGGL(k_getrf_panel, dim3(B), dim3(256), 0, s, d_A.p, d_b.p, np, k, d_fail.p);
          if (m > 0) {
            hipLaunchKernelGGL(k_trsm_lu, dim3(B, m), dim3(256), 0, s, d_A.p, np, k);
            hipLaunchKernelGGL(k_gemm_lu_mfma, dim3(B, m * m), dim3(256), 0, s, d_A.p, np, k, m);
          }
        } else {
          hipLaunchKernelGGL(k_chol_diag, dim3(B), dim3(256), 0, s, d_A.p, np, k, d_fail.p);
          if (m > 0) hipLaunchKernelGGL(k_chol_panel, dim3(B, m), dim3(256), 0, s, d_A.p, np, k);
        }
        DR_HIP(hipGetLastError());
      }
      if (lu)
        hipLaunchKernelGGL(k_solve_lu, dim3(B), dim3(256), np * sizeof(double), s, d_A.p, np, d_sd.p, d_b.p,
                           d_fail.p, d_sol.p);
      else
        hipLaunchKernelGGL(k_solve_chol, dim3(B), dim3(256), np * sizeof(double), s, d_A.p, np, d_sd.p, d_b.p,
                           d_sol.p);
      DR_HIP(hipGetLastError());
      std::vector<int> fail(B);
      d_fail.down(fail.data(), B, s);
      DR_HIP(hipStreamSynchronize(s));
      if (!lu)
        for (int i = 0; i < B; ++i)
          if (fail[i]) failed.push_back(bu[i]);
    }
    return failed;
  };
  const std::vector<int> indefinite = solve_all(order, false);
  if (!indefinite.empty()) (void)solve_all(indefinite, true);
  std::vector<int> uidd(n_rows);
  for (int i = 0; i < n_rows; ++i) uidd[i] = uid[i] < 0 ? 0 : uid[i];
  d_uid.up(uidd.data(), n_rows, s);
  DBuf<long> d_usol(std::max(nU, 1));
  if (nU) d_usol.up(usol.data(), nU, s);
  DBuf<double> d_vals(nnz);
  DBuf<int> d_nan(n_rows);
  hipLaunchKernelGGL(k_gather_rows, dim3(gblk), dim3(256), 0, s, n_rows, d_rp.p, d_uid.p, d_usol.p, d_sol.p,
                     d_vals.p, d_nan.p);
  DR_HIP(hipGetLastError());
  std::vector<int> nanrow(n_rows);
  d_nan.down(nanrow.data(), n_rows, s);
  DR_HIP(hipStreamSynchronize(s));
  // drop rows whose weight sum is NaN (kriging_matrix.h:320-327): one NaN
  // weight makes the sum NaN
  KrigingCSR out;
  out.nrow = n_rows;
  out.ncol = nnz_hr;
  out.n_unique = nU;
  out.n_lu = (long)indefinite.size();
  out.rp.assign(n_rows + 1, 0);
  long dropped = 0;
  for (int i = 0; i < n_rows; ++i) {
    const bool drop = cnt[i] > 0 && nanrow[i];
    dropped += drop ? 1 : 0;
    out.rp[i + 1] = out.rp[i] + (drop ? 0 : cnt[i]);
  }
  out.nnz = out.rp[n_rows];
  if (dropped == 0) {
    out.cols = std::move(d_cols);
    out.vals = std::move(d_vals);
  } else {
    DBuf<long> d_rpn(n_rows + 1);
    d_rpn.up(out.rp.data(), n_rows + 1, s);
    out.cols.alloc(out.nnz);
    out.vals.alloc(out.nnz);
    hipLaunchKernelGGL(k_compact_rows, dim3(gblk), dim3(256), 0, s, n_rows, d_rp.p, d_rpn.p, d_cols.p, d_vals.p,
                       out.cols.p, out.vals.p);
    DR_HIP(hipGetLastError());
    DR_HIP(hipStreamSynchronize(s));
  }
  (void)oobv;
  return out;
}

}  // namespace drh

struct dr_csr {
  drh::KrigingCSR m;
};

extern "C" {

int dr_kriging_build(const double* hr, const int hr_dims[3], const float q_hr[16], const double* std_img,
                     const int std_dims[3], const float q_std[16], const double theta[3], double radius, int device,
                     dr_csr** out) {
  return drh::guarded([&] {
    DR_REQUIRE(out, DR_ELOGIC, "dr_kriging_build: null output");
    auto* c = new dr_csr{drh::kriging_build(hr, hr_dims, q_hr, std_img, std_dims, q_std, theta, radius, device)};
    *out = c;
  });
}

int dr_csr_info(dr_csr* m, int64_t* nrow, int64_t* ncol, int64_t* nnz, int64_t* n_unique) {
  return drh::guarded([&] {
    DR_REQUIRE(m, DR_ELOGIC, "null csr");
    *nrow = m->m.nrow;
    *ncol = m->m.ncol;
    *nnz = (int64_t)m->m.nnz;
    if (n_unique) *n_unique = m->m.n_unique;
  });
}

int dr_csr_solver_counts(dr_csr* m, int64_t* n_cholesky, int64_t* n_lu) {
  return drh::guarded([&] {
    DR_REQUIRE(m && n_cholesky && n_lu, DR_ELOGIC, "dr_csr_solver_counts: null argument");
    *n_lu = m->m.n_lu;
    *n_cholesky = m->m.n_unique - m->m.n_lu;
  });
}

int dr_csr_copy(dr_csr* m, int64_t* rowptr, int32_t* cols, double* vals) {
  return drh::guarded([&] {
    DR_REQUIRE(m, DR_ELOGIC, "null csr");
    DR_REQUIRE(rowptr, DR_ELOGIC, "dr_csr_copy: null argument");
    std::copy(m->m.rp.begin(), m->m.rp.end(), rowptr);
    if (m->m.nnz > 0) {
      DR_REQUIRE(cols && vals, DR_ELOGIC, "dr_csr_copy: null argument");
      DR_HIP(hipMemcpy(cols, m->m.cols.p, m->m.nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
      DR_HIP(hipMemcpy(vals, m->m.vals.p, m->m.nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
  });
}

int dr_csr_free(dr_csr* m) {
  return drh::guarded([&] { delete m; });
}

}  // extern "C"
