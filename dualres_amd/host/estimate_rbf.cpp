// estimate_rbf: minimum-contrast estimation of the RBF covariance parameters
// of an image (src/estimate_rbf.cpp:19-186,
// impl/CommandParser/EstimRbfCommandParser.inl:60-206).  The per-offset
// covariance summary is computed on the GPU (dr_mce_summary); --bins N adds
// the binned empirical covariance (N bins up to --maxdist mm).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "engine.hpp"
#include "nifti.hpp"
#include "rbf_fit.hpp"

using namespace drhost;

namespace {
void usage() {
  std::cerr << "Estimate voxel-wise radial basis function covariance parameters:\nUsage:\n"
            << "\testimate_rbf path/to/img <options>\n\n";
}
void help() {
  usage();
  std::cerr << "Options:\n"
            << "\t--bandwidth float  fixes correlation bandwidth parameter\n"
            << "\t--constraint       imposes a bandwidth <= exponent constraint\n"
            << "\t--exponent  float  fixes correlation exponent parameter\n"
            << "\t--mask   path/to/mask_img  mask for input image\n"
            << "\t--output ofile/basename    output file for MCE summary data\n"
            << "\t--variance  float  fixes the marginal variance parameter\n"
            << "\t--xtol      float  set the numerical tolerance (default 1e-5)\n"
            << "\t--bins N [--maxdist mm]   also report N binned covariances up to maxdist (default 20)\n"
            << "\t--device n         GPU ordinal\n";
}
}  // namespace

int main(int argc, char** argv) {
  std::string image, mask, output;
  bool constraint = false, helpf = false, err = argc < 2;
  double par[3] = {-1, -1, -1}, xtol = 1e-5, maxdist = 20.0;
  int bins = 0, device = 0;
  for (int i = 1; i < argc && !err && !helpf; ++i) {
    const std::string a = argv[i];
    auto num = [&](double& dst, const char* what) {
      if (i + 1 < argc) {
        try { dst = std::stod(argv[++i]); return; } catch (...) {}
      }
      std::cerr << "\nWarning: " << what << " option requires 1 numeric argument\n";
      err = true;
    };
    if (a == "-h" || a == "--help") helpf = true;
    else if (a == "--constrained" || a == "--constraint") constraint = true;
    else if (a == "--variance" || a == "-var") num(par[0], "--variance");
    else if (a == "--bandwidth" || a == "-bw") num(par[1], "--bandwidth");
    else if (a == "--exponent" || a == "-exp") num(par[2], "--exponent");
    else if (a == "--xtol") {
      num(xtol, "--xtol");
      if (!err && xtol <= 0) { std::cerr << "\nError: --xtol should be > 0\n"; err = true; }
      if (!err && xtol >= 0.01) std::cerr << "\nWarning: numeric tolerance is high (default is 1e-5)\n";
    }
    else if (a == "--output" || a == "-o") { if (i + 1 < argc) output = argv[++i]; else err = true; }
    else if (a == "--mask" || a == "-m") { if (i + 1 < argc) mask = argv[++i]; else err = true; }
    else if (a == "--bins") { double b = 0; num(b, "--bins"); bins = (int)b; }
    else if (a == "--maxdist") num(maxdist, "--maxdist");
    else if (a == "--device") { double d = 0; num(d, "--device"); device = (int)d; }
    else {
      image = a;
      std::ifstream f(image);
      if (!f.is_open()) { std::cerr << "\nUnrecognized argument " << image << "\n"; err = true; }
    }
  }
  if (helpf) { help(); return 0; }
  if (err) { usage(); std::cerr << "\nSee estimate_rbf --help for more information\n"; return 1; }
  std::vector<float> cov, dist;
  std::vector<int> np;
  Image nii;
  try {
    nii = read_nifti(image);
    if (!mask.empty()) {
      Image m = read_nifti(mask);
      if (!same_orientation(nii, m)) throw std::domain_error("Error: Image/mask orientation mismatch");
      apply_mask(nii, m);
    }
    std::cout << "Computing covariances across the image... " << std::flush;
    int64_t n;
    check(dr_mce_offsets(&n, nullptr));
    cov.resize(n);
    dist.resize(n);
    np.resize(n);
    const int dims[3] = {nii.nx, nii.ny, nii.nz};
    check(dr_mce_summary(nii.data.data(), dims, nii.qto_xyz.data(), device, cov.data(), dist.data(), np.data()));
    std::cout << "Done!" << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "Exception caught with message:\n'" << e.what() << "'\n" << std::endl;
    return 1;
  }
  if (!output.empty()) {
    std::ofstream csv(output);
    if (csv.is_open()) {
      csv << "Radial_Distance,Covariance,N_Pairs\n";
      for (size_t i = 0; i < dist.size(); ++i) {
        csv << dist[i] << "," << cov[i] << "," << np[i];
        if (i != dist.size() - 1) csv << "\n";
      }
    } else {
      std::cerr << "\nWarning: unable to open '" << output << "' for writing. Data will not be saved";
    }
  }
  if (bins > 0) {
    std::vector<double> bc(bins), bd(bins);
    std::vector<int64_t> bn(bins);
    const int dims[3] = {nii.nx, nii.ny, nii.nz};
    check(dr_mce_binned(nii.data.data(), dims, nii.qto_xyz.data(), device, bins, maxdist, bc.data(), bd.data(),
                        bn.data()));
    std::cout << "Binned empirical covariance (distance, covariance, pairs):\n";
    for (int b = 0; b < bins; ++b) std::cout << "  " << bd[b] << "\t" << bc[b] << "\t" << bn[b] << "\n";
  }
  std::vector<double> theta{cov[0] * 0.8, 0.6, 1.5};
  std::cout << "Estimating smoothness (radial basis function approximation)... " << std::flush;
  try {
    compute_rbf_parameters(theta, cov, dist, np, constraint, 1e6, par[0], par[1], par[2], xtol);
  } catch (const std::exception& e) {
    std::cerr << "\nException caught with message:\n'" << e.what() << "'\n" << std::endl;
    return 1;
  }
  const double fwhm = 2.0 * std::pow(std::log(2.0) / theta[1], 1.0 / theta[2]);
  std::cout << "Done!\n  Marg. Var. = " << theta[0] << "\n  Bandwidth  = " << theta[1] << "\n  Exponent   = "
            << theta[2] << "\n  <FWHM      = " << fwhm << " (mm)>\n" << std::endl;
  std::cout << "You can pass these parameters to dualgpm[f] with the option:\n\t--covariance " << theta[0] << " "
            << theta[1] << " " << theta[2] << "\n" << std::endl;
  return 0;
}
