// Host-layer checks built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_host_sanitize.py compiles this file with g++ and runs it).
// Only CPU code is exercised: the NIfTI-1 reader and writer, the image index
// helpers, HMCParameters' dual averaging and the minimum-contrast RBF fit.
//
//   host_sanitize <in.nii[.gz]> <scratch dir> <truncated.nii> <short-data.nii.gz>
//
// Prints "ok" and exits 0 when every check holds; a sanitizer report or a
// failed check exits non-zero.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../dualres_amd/host/nifti.hpp"
#include "../dualres_amd/host/rbf_fit.hpp"
#include "../dualres_amd/host/sampler.hpp"

using namespace drhost;

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(3);                                                \
    }                                                              \
  } while (0)

static bool same(double x, double y) { return x == y || (std::isnan(x) && std::isnan(y)); }

static bool throws_on_read(const std::string& path) {
  try {
    (void)read_nifti(path);
  } catch (const std::exception&) {
    return true;
  }
  return false;
}

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const std::string in = argv[1], dir = argv[2];

  // 1. read, write back as float32 and float64, gzip and plain, read again
  const Image a = read_nifti(in);
  CHECK(a.nvox() == (long)a.data.size() && a.nvox() > 0);
  for (int dt : {16, 64})
    for (const char* ext : {".nii", ".nii.gz"}) {
      const std::string out = dir + "/rt" + std::to_string(dt) + ext;
      write_nifti(out, a, a.data, dt);
      const Image b = read_nifti(out);
      CHECK(b.nx == a.nx && b.ny == a.ny && b.nz == a.nz);
      for (long i = 0; i < a.nvox(); ++i)
        CHECK(same(b.data[i], dt == 64 ? a.data[i] : (double)(float)a.data[i]));
      for (int k = 0; k < 16; ++k) CHECK(std::fabs(b.qto_xyz[k] - a.qto_xyz[k]) <= 1e-5f * (1 + std::fabs(a.qto_xyz[k])));
    }

  // 2. malformed files are refused, not read past their end
  CHECK(throws_on_read(argv[3]));
  CHECK(throws_on_read(argv[4]));
  CHECK(throws_on_read(dir + "/does_not_exist.nii"));

  // 3. index helpers: nonzero voxels, bounding box, data round trip
  const auto ijk = nonzero_indices(a);
  const auto y = nonzero_data(a);
  CHECK(ijk.size() == y.size());
  const BBox bb = bounding_box(a);
  const auto ijk_b = nonzero_indices_bounded_by_box(a, bb);
  CHECK(ijk_b.size() == ijk.size());
  Image c = a;
  std::vector<double> doubled(y);
  for (double& v : doubled) v *= 2;
  emplace_nonzero_data(c, doubled);
  for (long i = 0; i < a.nvox(); ++i) CHECK(present(a.data[i]) ? c.data[i] == 2 * a.data[i] : same(c.data[i], a.data[i]));
  const auto vd = voxel_dimensions(a.qto_xyz);
  CHECK(vd[0] > 0 && vd[1] > 0 && vd[2] > 0);

  // 4. dual averaging over a long run of acceptance rates
  HMCParameters<double> h(300, 200, 2, 25, 0.04, 0.65);
  std::mt19937 g(7);
  std::uniform_real_distribution<double> U(0, 1);
  int saves = 0;
  while (h.iteration() <= h.max_iterations()) {
    h.update(U(g));
    saves += h.save_iteration() ? 1 : 0;
    CHECK(std::isfinite(h.eps_value()) && h.eps_value() > 0);
  }
  CHECK(saves == 200);

  // 5. minimum-contrast fit recovers the parameters of an exact RBF curve
  std::vector<float> cov, dist;
  std::vector<int> np;
  for (int i = 0; i < 400; ++i) {
    const double d = 0.05 * i;
    dist.push_back((float)d);
    cov.push_back((float)(0.9 * std::exp(-0.3 * std::pow(d, 1.4))));
    np.push_back(1000 - i);
  }
  std::vector<double> th{0.8 * cov[0], 0.6, 1.5};
  CHECK(compute_rbf_parameters(th, cov, dist, np, false) == 0);
  CHECK(std::fabs(th[0] - 0.9) < 2e-3 && std::fabs(th[1] - 0.3) < 2e-3 && std::fabs(th[2] - 1.4) < 2e-3);

  std::printf("ok\n");
  return 0;
}
