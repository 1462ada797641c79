// CPU restatement of the reference's HMC sampler (oracle; test and
// baseline infrastructure only: the GPU engine never links or calls it).
//
// MultiResParameters<T> (include/dualres/MultiResParameters.h) structure
// for structure, in C++17 with OpenMP, for timing the reference's CPU path
// on the host cores beside the device engine:
//   * complex state on the embedding torus, in real space (:96-101);
//   * every circulant product as an unnormalised inverse 3D c2c DFT, a
//     pointwise complex divide, the low-rank zeroing and a forward DFT
//     (:337-376, :690-700), i.e. 4 complex 3D DFTs per leapfrog step;
//   * the sparse products W mu and W^T v single-threaded, as the reference's
//     Eigen products run (w_threads = 1), or split over OpenMP threads;
//   * OpenMP only where the reference has pragmas (the gradient scatter
//     :369-373, the float64 energy reductions :450-470, :579-596); the other
//     array expressions run on one thread as Eigen evaluates them;
//   * serial std::mt19937 normals for the momentum (:783-815) with the
//     float build's Kahan-compensated E0, and the reference's energies,
//     accept test and Gibbs draws (:673-727, :820-859).
// The 3D DFT is this file's own: per axis, blocks of 16 lines are gathered
// into split real/imaginary buffers and transformed together by Stockham
// autosort stages of radix 4, 2, 3, 5 and 7 (FFTW is not available here).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int kLanes = 16;  // lines transformed together

// ---------------------------------------------------------------- 1D plan
template <typename T>
struct Plan1d {
  int n = 0;
  std::vector<int> radix;
  std::vector<T> twr, twi;  // exp(-2 pi i e / n), e in [0, n)
  void init(int len) {
    n = len;
    radix.clear();
    int m = len;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    for (int r : {3, 5, 7})
      while (m % r == 0) { radix.push_back(r); m /= r; }
    if (m != 1) throw std::domain_error("CPU DFT: length " + std::to_string(len) + " has a prime factor > 7");
    twr.resize(n);
    twi.resize(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int e = 0; e < n; ++e) {
      const long double a = -2.0L * pi * (long double)e / (long double)n;
      twr[e] = (T)std::cos(a);
      twi[e] = (T)std::sin(a);
    }
  }
};

// One Stockham stage of radix R on kLanes lines held as [n][kLanes] split
// arrays: butterfly j reads rows j + q n/R, multiplies row q by
// w^(q (j mod Ns) n / (Ns R)), and writes rows (j - j mod Ns) R + j mod Ns + q Ns.
// sgn = -1 forward, +1 backward (conjugated twiddles and roots).
template <typename T, int R>
void stage(const Plan1d<T>& p, int Ns, int sgn, const T* ar, const T* ai, T* br, T* bi) {
  const int n = p.n, nr = n / R;
  T cr[R][R], ci[R][R];
  for (int a = 0; a < R; ++a)
    for (int b = 0; b < R; ++b) {
      const double ang = sgn * 2.0 * M_PI * (double)((a * b) % R) / (double)R;
      cr[a][b] = (T)std::cos(ang);
      ci[a][b] = (T)std::sin(ang);
    }
  for (int j = 0; j < nr; ++j) {
    const int k = j % Ns;
    const int out = (j - k) * R + k;
    T vr[R][kLanes], vi[R][kLanes];
    for (int q = 0; q < R; ++q) {
      const T* xr = ar + (size_t)(j + q * nr) * kLanes;
      const T* xi = ai + (size_t)(j + q * nr) * kLanes;
      if (q == 0 || k == 0) {
        for (int l = 0; l < kLanes; ++l) { vr[q][l] = xr[l]; vi[q][l] = xi[l]; }
      } else {
        const int e = (int)(((long)q * k * (n / (Ns * R))) % n);
        const T wr = p.twr[e], wi = sgn < 0 ? p.twi[e] : -p.twi[e];
        for (int l = 0; l < kLanes; ++l) {
          vr[q][l] = xr[l] * wr - xi[l] * wi;
          vi[q][l] = xr[l] * wi + xi[l] * wr;
        }
      }
    }
    if constexpr (R == 2) {
      T* o0r = br + (size_t)out * kLanes; T* o0i = bi + (size_t)out * kLanes;
      T* o1r = br + (size_t)(out + Ns) * kLanes; T* o1i = bi + (size_t)(out + Ns) * kLanes;
      for (int l = 0; l < kLanes; ++l) {
        o0r[l] = vr[0][l] + vr[1][l]; o0i[l] = vi[0][l] + vi[1][l];
        o1r[l] = vr[0][l] - vr[1][l]; o1i[l] = vi[0][l] - vi[1][l];
      }
    } else if constexpr (R == 4) {
      T* o[4][2];
      for (int q = 0; q < 4; ++q) {
        o[q][0] = br + (size_t)(out + q * Ns) * kLanes;
        o[q][1] = bi + (size_t)(out + q * Ns) * kLanes;
      }
      const T s = (T)sgn;
      for (int l = 0; l < kLanes; ++l) {
        const T t0r = vr[0][l] + vr[2][l], t0i = vi[0][l] + vi[2][l];
        const T t1r = vr[0][l] - vr[2][l], t1i = vi[0][l] - vi[2][l];
        const T t2r = vr[1][l] + vr[3][l], t2i = vi[1][l] + vi[3][l];
        const T dr = vr[1][l] - vr[3][l], di = vi[1][l] - vi[3][l];
        const T t3r = -s * di, t3i = s * dr;  // (s i) (v1 - v3)
        o[0][0][l] = t0r + t2r; o[0][1][l] = t0i + t2i;
        o[2][0][l] = t0r - t2r; o[2][1][l] = t0i - t2i;
        o[1][0][l] = t1r + t3r; o[1][1][l] = t1i + t3i;
        o[3][0][l] = t1r - t3r; o[3][1][l] = t1i - t3i;
      }
    } else {
      for (int q = 0; q < R; ++q) {
        T* orr = br + (size_t)(out + q * Ns) * kLanes;
        T* oi = bi + (size_t)(out + q * Ns) * kLanes;
        for (int l = 0; l < kLanes; ++l) {
          T sr = vr[0][l], si = vi[0][l];
          for (int m = 1; m < R; ++m) {
            sr += vr[m][l] * cr[m][q] - vi[m][l] * ci[m][q];
            si += vr[m][l] * ci[m][q] + vi[m][l] * cr[m][q];
          }
          orr[l] = sr;
          oi[l] = si;
        }
      }
    }
  }
}

// kLanes transforms of one length; the result ends in (ar, ai) or (br, bi),
// returned as the pointer pair that holds it
template <typename T>
std::pair<T*, T*> run_lines(const Plan1d<T>& p, int sgn, T* ar, T* ai, T* br, T* bi) {
  int Ns = 1;
  for (int r : p.radix) {
    switch (r) {
      case 4: stage<T, 4>(p, Ns, sgn, ar, ai, br, bi); break;
      case 2: stage<T, 2>(p, Ns, sgn, ar, ai, br, bi); break;
      case 3: stage<T, 3>(p, Ns, sgn, ar, ai, br, bi); break;
      case 5: stage<T, 5>(p, Ns, sgn, ar, ai, br, bi); break;
      default: stage<T, 7>(p, Ns, sgn, ar, ai, br, bi); break;
    }
    Ns *= r;
    std::swap(ar, br);
    std::swap(ai, bi);
  }
  return {ar, ai};
}

// ---------------------------------------------------------------- 3D c2c DFT
// Unnormalised, x fastest (FFTW conventions of sized_fft.h): forward
// exp(-2 pi i k.f / M), backward exp(+2 pi i k.f / M), in place.
template <typename T>
struct Dft3d {
  int M[3] = {0, 0, 0};
  Plan1d<T> plan[3];
  int threads = 1;
  void init(const int m[3], int nthreads) {
    for (int a = 0; a < 3; ++a) {
      M[a] = m[a];
      plan[a].init(m[a]);
    }
    threads = nthreads;
  }
  void run(std::complex<T>* X, int sgn) const {
    axis_x(X, sgn);
    axis_strided(X, 1, sgn);
    axis_strided(X, 2, sgn);
  }
  void forward(std::complex<T>* X) const { run(X, -1); }
  void inverse(std::complex<T>* X) const { run(X, +1); }

 private:
  // rows of length M0: gather kLanes rows transposed
  void axis_x(std::complex<T>* X, int sgn) const {
    const int n = M[0];
    const long rows = (long)M[1] * M[2];
    const long blocks = (rows + kLanes - 1) / kLanes;
#pragma omp parallel num_threads(threads)
    {
      std::vector<T> buf(4 * (size_t)n * kLanes);
      T *ar = buf.data(), *ai = ar + (size_t)n * kLanes, *br = ai + (size_t)n * kLanes, *bi = br + (size_t)n * kLanes;
#pragma omp for schedule(static)
      for (long b = 0; b < blocks; ++b) {
        const long r0 = b * kLanes;
        const int nl = (int)std::min<long>(kLanes, rows - r0);
        for (int l = 0; l < kLanes; ++l) {
          const T* row = reinterpret_cast<const T*>(X + (size_t)(r0 + std::min(l, nl - 1)) * n);
          for (int j = 0; j < n; ++j) {
            ar[(size_t)j * kLanes + l] = row[2 * j];
            ai[(size_t)j * kLanes + l] = row[2 * j + 1];
          }
        }
        auto res = run_lines(plan[0], sgn, ar, ai, br, bi);
        for (int l = 0; l < nl; ++l) {
          T* row = reinterpret_cast<T*>(X + (size_t)(r0 + l) * n);
          for (int j = 0; j < n; ++j) {
            row[2 * j] = res.first[(size_t)j * kLanes + l];
            row[2 * j + 1] = res.second[(size_t)j * kLanes + l];
          }
        }
      }
    }
  }
  // y (a = 1) or z (a = 2): kLanes consecutive x columns at a time
  void axis_strided(std::complex<T>* X, int a, int sgn) const {
    const int n = M[a];
    const long stride = a == 1 ? (long)M[0] : (long)M[0] * M[1];
    const long outer = a == 1 ? (long)M[2] : (long)M[1];
    const long outer_stride = a == 1 ? (long)M[0] * M[1] : (long)M[0];
    const long xb = (M[0] + kLanes - 1) / kLanes;
#pragma omp parallel num_threads(threads)
    {
      std::vector<T> buf(4 * (size_t)n * kLanes);
      T *ar = buf.data(), *ai = ar + (size_t)n * kLanes, *br = ai + (size_t)n * kLanes, *bi = br + (size_t)n * kLanes;
#pragma omp for schedule(static)
      for (long t = 0; t < outer * xb; ++t) {
        const long o = t / xb, x0 = (t % xb) * kLanes;
        const int nl = (int)std::min<long>(kLanes, M[0] - x0);
        std::complex<T>* base = X + o * outer_stride + x0;
        for (int j = 0; j < n; ++j) {
          const T* src = reinterpret_cast<const T*>(base + (size_t)j * stride);
          T* dr = ar + (size_t)j * kLanes;
          T* di = ai + (size_t)j * kLanes;
          if (nl == kLanes) {
#pragma omp simd
            for (int l = 0; l < kLanes; ++l) {
              dr[l] = src[2 * l];
              di[l] = src[2 * l + 1];
            }
          } else {
            for (int l = 0; l < kLanes; ++l) {
              dr[l] = src[2 * std::min(l, nl - 1)];
              di[l] = src[2 * std::min(l, nl - 1) + 1];
            }
          }
        }
        auto res = run_lines(plan[a], sgn, ar, ai, br, bi);
        for (int j = 0; j < n; ++j) {
          T* dst = reinterpret_cast<T*>(base + (size_t)j * stride);
          const T* sr = res.first + (size_t)j * kLanes;
          const T* si = res.second + (size_t)j * kLanes;
#pragma omp simd
          for (int l = 0; l < nl; ++l) {
            dst[2 * l] = sr[l];
            dst[2 * l + 1] = si[l];
          }
        }
      }
    }
  }
};

// ---------------------------------------------------------------- the chain
struct ChainBase {
  virtual ~ChainBase() = default;
  virtual void set_sigma(const double* s) = 0;
  virtual void set_state(const double* mr, const double* mi, const double* pr, const double* pi) = 0;
  virtual void get_state(double* mr, double* mi, double* pr, double* pi) const = 0;
  virtual void gradient(double* gr, double* gi) = 0;
  virtual void leapfrog(double eps, int L, double* out) = 0;
  virtual void update(double eps, int L, double u, const double* zr, const double* zi, double* out) = 0;
  virtual void add_data_to_mu() = 0;
  virtual long nonpositive() const = 0;
};

// circulant_base_3d (impl/circulant_base.inl:29-89): float distance of the
// torus offset (d(i), d(j), d(k)), d(t) = min(t, M - t), through the qform
template <typename T>
std::vector<T> torus_base(const int M[3], const float q[16]) {
  auto seq = [](int m) {
    std::vector<float> s(m);
    const int grid = m / 2 + 1;
    float i = 0.0f;
    for (int l = 0; l < m; ++l) {
      i += l < grid ? 1.0f : -1.0f;
      s[l] = i - 1.0f;
    }
    return s;
  };
  const auto s0 = seq(M[0]), s1 = seq(M[1]), s2 = seq(M[2]);
  std::vector<T> b((size_t)M[0] * M[1] * M[2]);
#pragma omp parallel for
  for (int k = 0; k < M[2]; ++k)
    for (int j = 0; j < M[1]; ++j)
      for (int i = 0; i < M[0]; ++i) {
        float c[3];
        for (int r = 0; r < 3; ++r) {
          volatile float v = q[4 * r] * s0[i];
          v = v + q[4 * r + 1] * s1[j];
          v = v + q[4 * r + 2] * s2[k];
          c[r] = v;
        }
        volatile float x2 = c[0] * c[0], y2 = c[1] * c[1], z2 = c[2] * c[2];
        volatile float ss = x2 + y2;
        ss = ss + z2;
        b[i + (size_t)M[0] * (j + (size_t)M[1] * k)] = (T)std::sqrt((float)ss);
      }
  return b;
}

template <typename T>
struct CpuChain : ChainBase {
  typedef std::complex<T> C;
  int M[3];
  long N = 0;
  int nd = 1;
  Dft3d<T> dft;
  std::vector<C> lam, lam_mass, mom, mu, mu_star, tmp;
  std::vector<long> Z;
  std::vector<long> I;
  std::vector<T> Yh, Ys, real_mu, real_sub_grad, resid_s;
  std::vector<int64_t> rp;
  std::vector<int32_t> cols, rmin, rmax;  // per-row smallest / largest column
  std::vector<T> vals;
  T sig[2] = {1, 1};
  double initial_energy = 0;
  int threads = 1, w_threads = 1;
  std::mt19937 rng;

  CpuChain(int n_datasets, const double th[3], const float q[16], const int m[3], const int32_t* ijk, long n_h,
           const double* yh, const double* ys, long n_s, const int64_t* rp_, const int32_t* cols_,
           const double* vals_, int nthreads, int wthreads, uint32_t seed)
      : rng(seed) {
    nd = n_datasets;
    threads = std::max(1, nthreads);
    w_threads = std::max(1, wthreads);
    for (int a = 0; a < 3; ++a) M[a] = m[a];
    N = (long)M[0] * M[1] * M[2];
    dft.init(M, threads);
    // spectrum (MultiResParameters.h:232-253): lambda = F(rbf(base))
    const std::vector<T> base = torus_base<T>(M, q);
    lam.resize(N);
    const T tau2 = (T)th[0], psi = (T)th[1], nu = (T)th[2];
    for (long i = 0; i < N; ++i) lam[i] = C(tau2 * std::exp(-psi * std::pow(std::abs(base[i]), nu)), 0);
    dft.forward(lam.data());
    for (long i = 0; i < N; ++i)
      if (lam[i].real() <= 0) Z.push_back(i);
    if ((long)Z.size() >= N / 2) throw std::domain_error("Too many negative eigen values");
    I.resize(n_h);
    for (long i = 0; i < n_h; ++i)
      I[i] = (long)ijk[3 * i + 2] * M[0] * M[1] + (long)ijk[3 * i + 1] * M[0] + ijk[3 * i];
    Yh.assign(yh, yh + n_h);
    real_mu.assign(n_h, 0);
    real_sub_grad.assign(n_h, 0);
    if (nd == 2) {
      Ys.assign(ys, ys + n_s);
      rp.assign(rp_, rp_ + n_s + 1);
      cols.assign(cols_, cols_ + rp[n_s]);
      vals.resize(rp[n_s]);
      for (long i = 0; i < rp[n_s]; ++i) vals[i] = (T)vals_[i];
      resid_s.assign(n_s, 0);
      rmin.assign(n_s, INT32_MAX);
      rmax.assign(n_s, -1);
      for (long r = 0; r < n_s; ++r)
        for (int64_t t = rp[r]; t < rp[r + 1]; ++t) {
          rmin[r] = std::min(rmin[r], cols[t]);
          rmax[r] = std::max(rmax[r], cols[t]);
        }
    }
    lam_mass.assign(N, C(0, 0));
    mom.assign(N, C(0, 0));
    mu.assign(N, C(0, 0));
    mu_star.assign(N, C(0, 0));
    tmp.assign(N, C(0, 0));
    // _initialize_mu (:743-759)
    std::normal_distribution<T> G(0, 1);
    for (long i = 0; i < N; ++i) {
      const T a = G(rng), b = G(rng);
      mu[i] = C(a, b);
    }
    dft.forward(mu.data());
    for (long i = 0; i < N; ++i) mu[i] *= std::sqrt(lam[i]) / ((T)10 * (T)N);
    low_rank(mu);
    dft.inverse(mu.data());
  }

  long nonpositive() const override { return (long)Z.size(); }
  void low_rank(std::vector<C>& a) const {
    for (long z : Z) a[z] = C(0, 0);
  }
  void gather_real(const std::vector<C>& m) {
    for (size_t i = 0; i < I.size(); ++i) real_mu[i] = m[I[i]].real();
  }
  // W x (row-major CSR)
  void w_times(const T* x, T* out) const {
    const long ns = (long)Ys.size();
#pragma omp parallel for num_threads(w_threads) schedule(static)
    for (long r = 0; r < ns; ++r) {
      T acc = 0;
      for (int64_t t = rp[r]; t < rp[r + 1]; ++t) acc += vals[t] * x[cols[t]];
      out[r] = acc;
    }
  }
  // out += s * W^T v: rows scattered into columns, one thread per column
  // range when w_threads > 1 (each thread scans all rows for its columns)
  void wt_times_add(const T* v, T s, T* out) const {
    const long ns = (long)Ys.size();
    if (w_threads == 1) {
      for (long r = 0; r < ns; ++r) {
        const T f = s * v[r];
        for (int64_t t = rp[r]; t < rp[r + 1]; ++t) out[cols[t]] += f * vals[t];
      }
      return;
    }
    const long nc = (long)real_mu.size();
#pragma omp parallel num_threads(w_threads)
    {
      const int nt = omp_get_num_threads(), id = omp_get_thread_num();
      const long c0 = nc * id / nt, c1 = nc * (id + 1) / nt;
      for (long r = 0; r < ns; ++r) {
        const int64_t a = rp[r], b = rp[r + 1];
        if (b == a || rmax[r] < c0 || rmin[r] >= c1) continue;
        const T f = s * v[r];
        for (int64_t t = a; t < b; ++t) {
          const int32_t c = cols[t];
          if (c >= c0 && c < c1) out[c] += f * vals[t];
        }
      }
    }
  }

  // _compute_gradient (:337-376): result in tmp
  const std::vector<C>& compute_gradient(const std::vector<C>& ms) {
    std::copy(ms.begin(), ms.end(), tmp.begin());
    dft.inverse(tmp.data());
    const T nN = -(T)N;
    for (long i = 0; i < N; ++i) tmp[i] = cdiv(tmp[i], nN * lam[i]);
    low_rank(tmp);
    dft.forward(tmp.data());
    gather_real(ms);
    const long nh = (long)I.size();
    for (long i = 0; i < nh; ++i) real_sub_grad[i] = sig[0] * (Yh[i] - real_mu[i]);
    if (nd == 2) {
      w_times(real_mu.data(), resid_s.data());
      for (size_t r = 0; r < resid_s.size(); ++r) resid_s[r] = Ys[r] - resid_s[r];
      wt_times_add(resid_s.data(), sig[1], real_sub_grad.data());
    }
#pragma omp parallel for num_threads(threads)
    for (long i = 0; i < nh; ++i) tmp[I[i]] += C(real_sub_grad[i], 0);
    return tmp;
  }
  // Eigen's packet complex division: a conj(b) / |b|^2
  static C cdiv(C a, C b) {
    const T d = b.real() * b.real() + b.imag() * b.imag();
    return C((a.real() * b.real() + a.imag() * b.imag()) / d, (a.imag() * b.real() - a.real() * b.imag()) / d);
  }

  // (Re + Im) sum of conj(t) t / d over the kept modes, as the reference
  // reduces it: serial Kahan in complex double (float build) or an OpenMP
  // reduction in T (double build)
  double energy_sum(const std::vector<C>& t, const std::vector<C>& d, bool by_real_part) const {
    if (std::is_same<T, float>::value) {
      std::complex<double> acc(0, 0), cor(0, 0);
      for (long i = 0; i < N; ++i) {
        const bool keep = by_real_part ? d[i].real() > 0 : d[i] != C(0, 0);
        if (!keep) continue;
        const std::complex<double> tv(t[i]), dv(d[i]);
        const std::complex<double> A = std::conj(tv) * tv / dv - cor;
        const std::complex<double> B = acc + A;
        cor = (B - acc) - A;
        acc = B;
      }
      return acc.real() + acc.imag();
    }
    T re = 0, im = 0;
#pragma omp parallel for num_threads(threads) reduction(+ : re, im)
    for (long i = 0; i < N; ++i) {
      const bool keep = by_real_part ? d[i].real() > 0 : d[i] != C(0, 0);
      if (!keep) continue;
      const C v = cdiv(std::conj(t[i]) * t[i], d[i]);
      re += v.real();
      im += v.imag();
    }
    return (double)re + (double)im;
  }
  double log_prior(const std::vector<C>& ms) {
    std::copy(ms.begin(), ms.end(), tmp.begin());
    dft.inverse(tmp.data());
    return -0.5 * energy_sum(tmp, lam, true) / (double)N;
  }
  double log_likelihood(const std::vector<C>& ms) {
    gather_real(ms);
    double a = 0;
    for (size_t i = 0; i < I.size(); ++i) {
      const double d = (double)(Yh[i] - real_mu[i]);
      a += d * d;
    }
    double ll = (double)sig[0] * a;
    if (nd == 2) {
      w_times(real_mu.data(), resid_s.data());
      double b = 0;
      for (size_t r = 0; r < resid_s.size(); ++r) {
        const double d = (double)(Ys[r] - resid_s[r]);
        b += d * d;
      }
      ll += (double)sig[1] * b;
    }
    return -0.5 * ll;
  }
  double potential_energy() {
    std::copy(mom.begin(), mom.end(), tmp.begin());
    dft.inverse(tmp.data());
    return -0.5 * energy_sum(tmp, lam_mass, false) / (double)N;
  }
  void mass_matrix() {  // :734-738
    for (long i = 0; i < N; ++i) lam_mass[i] = (T)1 / lam[i] + sig[0];
    low_rank(lam_mass);
  }
  // _sample_momentum (:783-815); injected normals when zr is given
  void sample_momentum(const double* zr, const double* zi) {
    std::normal_distribution<T> G(0, 1);
    if (std::is_same<T, float>::value) {
      double e = 0, cor = 0;
      for (long i = 0; i < N; ++i) {
        if (zr) {
          mom[i] = C((T)zr[i], (T)zi[i]);
        } else {
          const T a = G(rng), b = G(rng);
          mom[i] = C(a, b);
        }
        const double A = (std::conj(std::complex<double>(mom[i])) * std::complex<double>(mom[i])).real() - cor;
        const double B = e + A;
        cor = (B - e) - A;
        e = B;
      }
      initial_energy = -0.5 * e;
    } else {
      double e = 0;
      for (long i = 0; i < N; ++i) {
        if (zr) {
          mom[i] = C((T)zr[i], (T)zi[i]);
        } else {
          const T a = G(rng), b = G(rng);
          mom[i] = C(a, b);
        }
        e += (std::conj(mom[i]) * mom[i]).real();
      }
      initial_energy = -0.5 * e;
    }
    dft.inverse(mom.data());
    for (long i = 0; i < N; ++i) mom[i] *= std::sqrt(lam_mass[i]) / (T)N;
    low_rank(mom);
    dft.forward(mom.data());
  }
  // the leapfrog of update (:690-700) from (mu_star, mom)
  void integrate(T eps, int L) {
    T k = 0.5;
    {
      const auto& g = compute_gradient(mu_star);
      for (long i = 0; i < N; ++i) mom[i] += k * eps * g[i];
    }
    for (int step = 0; step < L; ++step) {
      k = step == L - 1 ? (T)0.5 : (T)1;
      std::copy(mom.begin(), mom.end(), tmp.begin());
      dft.inverse(tmp.data());
      const T f = (T)N / eps;
      for (long i = 0; i < N; ++i) tmp[i] = cdiv(tmp[i], lam_mass[i] * f);
      low_rank(tmp);
      dft.forward(tmp.data());
      for (long i = 0; i < N; ++i) mu_star[i] += tmp[i];
      const auto& g = compute_gradient(mu_star);
      for (long i = 0; i < N; ++i) mom[i] += k * eps * g[i];
    }
  }

  void set_sigma(const double* s) override {
    for (int i = 0; i < nd; ++i) sig[i] = (T)s[i];
  }
  void set_state(const double* mr, const double* mi, const double* pr, const double* pi) override {
    for (long i = 0; i < N; ++i) {
      if (mr) mu[i] = C((T)mr[i], mi ? (T)mi[i] : 0);
      if (pr) mom[i] = C((T)pr[i], pi ? (T)pi[i] : 0);
    }
  }
  void get_state(double* mr, double* mi, double* pr, double* pi) const override {
    for (long i = 0; i < N; ++i) {
      if (mr) mr[i] = mu[i].real();
      if (mi) mi[i] = mu[i].imag();
      if (pr) pr[i] = mom[i].real();
      if (pi) pi[i] = mom[i].imag();
    }
  }
  void gradient(double* gr, double* gi) override {
    const auto& g = compute_gradient(mu);
    for (long i = 0; i < N; ++i) {
      gr[i] = g[i].real();
      gi[i] = g[i].imag();
    }
  }
  // out: log posterior and kinetic energy before and after
  void leapfrog(double eps, int L, double* out) override {
    mass_matrix();
    out[0] = log_prior(mu) + log_likelihood(mu);
    out[1] = potential_energy();
    mu_star = mu;
    integrate((T)eps, L);
    mu = mu_star;
    out[2] = log_prior(mu) + log_likelihood(mu);
    out[3] = potential_energy();
  }
  // MultiResParameters::update (:673-727) and _update_sigma_sq_inv
  // (:820-859).  u < 0 draws the accept uniform from the chain's generator.
  // out: H0, H1, R, accepted
  void update(double eps, int L, double u, const double* zr, const double* zi, double* out) override {
    std::uniform_real_distribution<double> U(0, 1);
    mu_star = mu;
    mass_matrix();
    sample_momentum(zr, zi);
    const double lp_initial = log_prior(mu) + log_likelihood(mu);
    initial_energy += lp_initial;
    integrate((T)eps, L);
    const double lp_proposal = log_prior(mu_star) + log_likelihood(mu_star);
    const double proposed = lp_proposal + potential_energy();
    double R = std::exp(proposed - initial_energy);
    if (std::isnan(R)) R = 0;
    const double uu = u < 0 ? U(rng) : u;
    const bool acc = uu < R;
    if (acc) mu = mu_star;
    gather_real(mu);
    {  // Gibbs draws of the noise precisions
      T shape = (T)0.5 * (T)Yh.size() + 1;
      T ss = 0;
      for (size_t i = 0; i < Yh.size(); ++i) ss += (Yh[i] - real_mu[i]) * (Yh[i] - real_mu[i]);
      std::gamma_distribution<T> Ga(shape, 1 / ((T)0.5 * ss));
      sig[0] = Ga(rng);
      if (nd == 2) {
        w_times(real_mu.data(), resid_s.data());
        T s2 = 0;
        for (size_t r = 0; r < Ys.size(); ++r) s2 += (Ys[r] - resid_s[r]) * (Ys[r] - resid_s[r]);
        shape = (T)0.5 * (T)Ys.size() + 1;
        std::gamma_distribution<T> Gb(shape, 1 / ((T)0.5 * s2));
        sig[1] = Gb(rng);
      }
    }
    out[0] = initial_energy;
    out[1] = proposed;
    out[2] = R;
    out[3] = acc ? 1 : 0;
  }
  void add_data_to_mu() override {  // :631-641
    for (size_t i = 0; i < I.size(); ++i) mu[I[i]] += C(Yh[i], 0);
  }
};

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::domain_error& e) {
    g_err = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 3;
  }
}

}  // namespace

extern "C" {

const char* drc_last_error(void) { return g_err.c_str(); }

// precision: 0 float64 (dualgpm), 1 float32 (dualgpmf).  theta: (tau^2, psi,
// nu).  M: embedding torus.  ijk: n_h x 3 data indices.  W: n_s x n_h CSR
// (n_datasets == 2).  threads: OpenMP threads of the DFTs and the reference's
// pragma sites; w_threads: threads of the W products (1 = as the reference).
int drc_create(int precision, int n_datasets, const double* theta, const float* qform, const int* M,
               const int32_t* ijk, int64_t n_h, const double* Yh, const double* Ys, int64_t n_s, const int64_t* rp,
               const int32_t* cols, const double* vals, int threads, int w_threads, uint32_t seed, void** out) {
  return guard([&] {
    if (precision == 0)
      *out = new CpuChain<double>(n_datasets, theta, qform, M, ijk, n_h, Yh, Ys, n_s, rp, cols, vals, threads,
                                  w_threads, seed);
    else
      *out = new CpuChain<float>(n_datasets, theta, qform, M, ijk, n_h, Yh, Ys, n_s, rp, cols, vals, threads,
                                 w_threads, seed);
  });
}
int drc_destroy(void* c) {
  delete static_cast<ChainBase*>(c);
  return 0;
}
int drc_nonpositive(void* c, int64_t* n) {
  *n = static_cast<ChainBase*>(c)->nonpositive();
  return 0;
}
int drc_set_sigma(void* c, const double* s) {
  return guard([&] { static_cast<ChainBase*>(c)->set_sigma(s); });
}
int drc_set_state(void* c, const double* mr, const double* mi, const double* pr, const double* pi) {
  return guard([&] { static_cast<ChainBase*>(c)->set_state(mr, mi, pr, pi); });
}
int drc_get_state(void* c, double* mr, double* mi, double* pr, double* pi) {
  return guard([&] { static_cast<ChainBase*>(c)->get_state(mr, mi, pr, pi); });
}
int drc_gradient(void* c, double* gr, double* gi) {
  return guard([&] { static_cast<ChainBase*>(c)->gradient(gr, gi); });
}
int drc_leapfrog(void* c, double eps, int L, double* out4) {
  return guard([&] { static_cast<ChainBase*>(c)->leapfrog(eps, L, out4); });
}
int drc_update(void* c, double eps, int L, double u, const double* zr, const double* zi, double* out4) {
  return guard([&] { static_cast<ChainBase*>(c)->update(eps, L, u, zr, zi, out4); });
}
int drc_add_data_to_mu(void* c) {
  return guard([&] { static_cast<ChainBase*>(c)->add_data_to_mu(); });
}

}  // extern "C"
