// dualgpm / dualgpmf: Bayesian denoising of a high-resolution statistic
// image, optionally fused with a standard-resolution image.  Same options,
// defaults, outputs and error behaviour as the reference tool
// (src/dualgpm.cpp:45-464, impl/CommandParser/GPMCommandParser.inl:196-545);
// the sampler runs on the GPU through include/dualres_hip.h.
//
// Built twice: -DDUALRES_SINGLE_PRECISION gives dualgpmf (float32 gradients,
// float64 energies and accept ratio), otherwise dualgpm (float64).
// Additional options: --embedding pow2|2x|smooth, --state compat|real,
// --device N.
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"
#include "nifti.hpp"
#include "rbf_fit.hpp"
#include "sampler.hpp"

#ifdef DUALRES_SINGLE_PRECISION
typedef float scalar_type;
static const int kPrecision = DR_FP32;
#else
typedef double scalar_type;
static const int kPrecision = DR_FP64;
#endif

using namespace drhost;

namespace {

struct Options {
  std::string highres, stdres, hmask, smask, omask, output_base;
  std::vector<double> theta;
  double neighborhood = -1;
  int burnin = 1000, nsave = 1000, thin = 3, L = 25;
  double mhtarget = 0.65;
  unsigned seed = 1;
  int threads = 0;
  bool monitor = false, samples = false, rq = false, help = false, error = false, update_tau = false;
  std::string embedding = "pow2", state = "compat";
  int device = 0;
};

void usage(const std::string& caller) {
  std::cerr << "Usage:\n\t" << caller << " --highres path/to/img <options>\n\n";
}

void help(const std::string& caller) {
  usage(caller);
  std::cerr << "Options:\n"
            << "  --highres      path  high-resolution image (required)\n"
            << "  --stdres       path  standard-resolution image\n"
            << "  --hmask/-hm    path  high-resolution mask\n"
            << "  --smask/-sm    path  standard-resolution mask\n"
            << "  --omask/-om    path  output mask\n"
            << "  --covariance/--theta  f f f  kernel (variance, bandwidth, exponent)\n"
            << "  --neighborhood f     kriging neighbourhood radius (mm)\n"
            << "  --burnin n  --nsave/-n n  --thin n  --leapfrog/-L n  --mhtarget f\n"
            << "  --seed n  --threads n  --output/-o base  --samples  --monitor  --debug\n"
            << "  --rational-quadratic\n"
            << "  --update-tau                Gibbs-update the prior variance (compiled out of the reference)\n"
            << "  --embedding pow2|2x|smooth  circulant embedding rule (default pow2)\n"
            << "  --state compat|real         complex-state energies (default) or real state\n"
            << "  --device n                  GPU ordinal\n"
            << "\nDefaults: --burnin 1000 --nsave 1000 --thin 3 --leapfrog 25 --mhtarget 0.65\n";
}

Options parse(int argc, char** argv) {
  Options o;
  const auto now = std::chrono::high_resolution_clock::now().time_since_epoch();
  o.seed = std::max((unsigned)std::chrono::duration_cast<std::chrono::milliseconds>(now).count(), 1u);
  std::ostringstream ss;
  ss << "dualgpm_mcmc_" << o.seed << "_";
  o.output_base = ss.str();
  if (argc < 2) o.error = true;
  auto need = [&](int& i, int k, const char* msg) {
    if (i + k < argc) return true;
    std::cerr << msg;
    o.error = true;
    return false;
  };
  for (int i = 1; i < argc && !o.error && !o.help; ++i) {
    const std::string a = argv[i];
    try {
      if (a == "-h" || a == "--help") o.help = true;
      else if (a == "--burnin") { if (need(i, 1, "\nWarning: --burnin option requires 1 positive integer argument\n")) o.burnin = std::max(std::abs(std::stoi(argv[++i])), 1); }
      else if (a == "--covariance" || a == "--theta") {
        if (need(i, 3, "\nWarning: --covariance option requires 3 numeric arguments\n")) {
          o.theta.resize(3);
          for (int j = 0; j < 3; ++j) o.theta[j] = (scalar_type)std::stod(argv[++i]);
        }
      }
      else if (a == "--debug") { o.burnin = 50; o.nsave = 10; o.thin = 1; o.monitor = true; o.samples = true; }
      else if (a == "--highres") { if (need(i, 1, "\nWarning: --highres option requires one argument\n")) o.highres = argv[++i]; }
      else if (a == "--hmask" || a == "-hm") { if (need(i, 1, "\nWarning: --hmask option requires one argument\n")) o.hmask = argv[++i]; }
      else if (a == "--leapfrog" || a == "-L") { if (need(i, 1, "\nWarning: --leapfrog option requires 1 positive integer argument\n")) o.L = std::max(std::abs(std::stoi(argv[++i])), 1); }
      else if (a == "--mhtarget") {
        if (need(i, 1, "\nWarning: --mhtarget option requires 1 numeric argument on (0, 1)\n")) {
          o.mhtarget = std::stod(argv[++i]);
          if (o.mhtarget <= 0 || o.mhtarget >= 1) { std::cerr << "\nWarning: --mhtarget option requires 1 numeric argument on (0, 1)\n"; o.error = true; }
        }
      }
      else if (a == "--monitor") o.monitor = true;
      else if (a == "--neighborhood") { if (need(i, 1, "\nWarning: --neighborhood option requires 1 numeric argument\n")) o.neighborhood = std::fabs((scalar_type)std::stod(argv[++i])); }
      else if (a == "--nsave" || a == "-n") { if (need(i, 1, "\nWarning: --nsave option requires 1 positive integer argument\n")) o.nsave = std::max(std::abs(std::stoi(argv[++i])), 1); }
      else if (a == "--output" || a == "-o") { if (need(i, 1, "\nWarning: --output option requires 1 string argument\n")) o.output_base = argv[++i]; }
      else if (a == "--omask" || a == "-om") { if (need(i, 1, "\nWarning: --omask option requires one argument\n")) o.omask = argv[++i]; }
      else if (a == "--rational-quadratic") o.rq = true;
      else if (a == "--update-tau") o.update_tau = true;
      else if (a == "--samples") o.samples = true;
      else if (a == "--seed") { if (need(i, 1, "\nWarning: --seed option requires 1 integer argument\n")) o.seed = (unsigned)std::abs(std::stoi(argv[++i])); }
      else if (a == "--smask" || a == "-sm") { if (need(i, 1, "\nWarning: --smask option requires one argument\n")) o.smask = argv[++i]; }
      else if (a == "--stdres") { if (need(i, 1, "\nWarning: --stdres option requires one argument\n")) o.stdres = argv[++i]; }
      else if (a == "--thin") { if (need(i, 1, "\nWarning: --thin option requires 1 positive integer argument\n")) o.thin = std::max(std::abs(std::stoi(argv[++i])), 1); }
      else if (a == "--threads") { if (need(i, 1, "\nWarning: --threads option requires 1 integer argument\n")) o.threads = std::abs(std::stoi(argv[++i])); }
      else if (a == "--embedding") { if (need(i, 1, "\nWarning: --embedding requires pow2, 2x or smooth\n")) o.embedding = argv[++i]; }
      else if (a == "--state") { if (need(i, 1, "\nWarning: --state requires compat or real\n")) o.state = argv[++i]; }
      else if (a == "--device") { if (need(i, 1, "\nWarning: --device requires 1 integer argument\n")) o.device = std::stoi(argv[++i]); }
      else { std::cerr << "\nWarning: Unrecognized option: " << a << std::endl; o.error = true; }
    } catch (...) {
      std::cerr << "\nWarning: bad argument for option " << a << "\n";
      o.error = true;
    }
  }
  for (const std::string* f : {&o.highres, &o.hmask, &o.omask, &o.stdres, &o.smask})
    if (!f->empty() && !is_nifti_file(*f)) {
      std::cerr << "\nWarning: input file does not appear to conform to the nifti-1 standard\n\t" << *f << std::endl;
      o.error = true;
    }
  if (!o.help && o.highres.empty()) {
    if (!o.error) std::cerr << "\nError: User must supply the --highres argument\n";
    o.error = true;
  }
  if (o.embedding != "pow2" && o.embedding != "2x" && o.embedding != "smooth") o.error = true;
  if (o.state != "compat" && o.state != "real") o.error = true;
  if (o.hmask.empty()) o.hmask = o.highres;
  if (o.omask.empty()) o.omask = o.highres;
  return o;
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

bool compute_covar_parameters_if_needed(std::vector<double>& theta, const Image& img, int device) {
  if (!theta.empty()) return true;
  std::vector<double> th{1, 0.6, 1.5};
  std::cout << "Estimating covariance parameters... " << std::flush;
  int64_t n;
  check(dr_mce_offsets(&n, nullptr));
  std::vector<float> cov(n), dist(n);
  std::vector<int> np(n);
  const int dims[3] = {img.nx, img.ny, img.nz};
  check(dr_mce_summary(img.data.data(), dims, img.qto_xyz.data(), device, cov.data(), dist.data(), np.data()));
  if (compute_rbf_parameters(th, cov, dist, np) == 1) {
    std::cerr << "Could not estimate covar parameters.\nTry re-running with the --covariance argument supplied"
              << std::endl;
    return false;
  }
  theta = th;
  std::cout << "Done!" << std::endl;
  return true;
}

bool valid_covar_parameters(const std::vector<double>& t) {
  bool ok = true;
  if (t[0] <= 0) { std::cerr << "First covariance parameter (marginal variance) must be > 0\n"; ok = false; }
  if (t[1] <= 0) { std::cerr << "Second covariance parameter (bandwidth) must be > 0\n"; ok = false; }
  if (t[2] <= 0 || t[2] > 2) { std::cerr << "Third covariance parameter (exponent) must be in (0, 2]\n"; ok = false; }
  return ok;
}

}  // namespace

int main(int argc, char* argv[]) {
  const std::string caller = argc > 0 ? std::string(argv[0]).substr(std::string(argv[0]).find_last_of('/') + 1) : "dualgpm";
  Options in = parse(argc, argv);
  if (in.help) {
    help(caller);
    return 0;
  }
  if (in.error) {
    usage(caller);
    std::cerr << "\nSee dualgpm --help for more information\n";
    return 1;
  }
  rng().seed(in.seed);
  const bool dual = !in.stdres.empty();
  bool error_status = false;
  std::ofstream samples;
  if (in.samples) samples.open(in.output_base + "_samples.dat");
  try {
    Image hr = read_nifti(in.highres);
    std::cout << "High-resolution file: " << in.highres << std::endl;
    Image hmask = read_nifti(in.hmask);
    Image omask = read_nifti(in.omask);
    if (same_orientation(hr, hmask) && hr.nvox() == hmask.nvox()) {
      apply_mask(hr, hmask);
      std::cout << "\t[Masked by: " << in.hmask << "]" << std::endl;
    } else {
      std::cerr << "  *** WARNING ***  High-resolution image/mask mismatch. Image not masked" << std::endl;
      throw std::domain_error("Improper mask");
    }
    if (!same_orientation(hr, omask) || hr.nvox() != omask.nvox()) {
      std::cerr << "  *** WARNING ***  High-resolution/output mask mismatch." << std::endl;
      throw std::domain_error("Improper mask");
    }
    if (in.samples && !samples) throw std::runtime_error("Error: will not be able to write output to\n  " +
                                                         in.output_base + "_samples.dat");
    MultiResData data;
    data.theta = in.theta;
    double neighborhood = in.neighborhood;
    Image sr;
    if (dual) {
      sr = read_nifti(in.stdres);
      std::cout << "Standard-resolution file: " << in.stdres << std::endl;
      if (sr.datatype != hr.datatype)
        throw std::domain_error("Error: " + in.highres + " and " + in.stdres + " have different data types\n");
      if (!in.smask.empty()) {
        Image sm = read_nifti(in.smask);
        if (same_orientation(sr, sm) && sr.nvox() == sm.nvox()) {
          apply_mask(sr, sm);
          std::cout << "\t[Masked by: " << in.smask << "]" << std::endl;
        } else {
          std::cerr << "  *** WARNING ***  Standard-resolution image/mask mismatch. Image not masked" << std::endl;
        }
      }
      if (!compute_covar_parameters_if_needed(data.theta, sr, in.device))
        throw std::runtime_error("Error: could not compute covariance parameters from " + in.stdres + "\n");
      if (neighborhood <= 0) {
        const auto vd = voxel_dimensions(hr.qto_xyz);
        neighborhood = 3 * (scalar_type)std::max({vd[0], vd[1], vd[2]});
      }
      std::cout << "Neighborhood: " << neighborhood << " (mm)\n";
    }
    if (!compute_covar_parameters_if_needed(data.theta, hr, in.device))
      throw std::runtime_error("Error: could not compute covariance parameters from " + in.highres + "\n");
    if (!valid_covar_parameters(data.theta)) {
      std::ostringstream e;
      e << "Error: covariance parameters are not valid (" << data.theta[0] << ", " << data.theta[1] << ", "
        << data.theta[2] << ")\n";
      throw std::domain_error(e.str());
    }
    std::cout << "Covariance parameters: (" << data.theta[0] << ", " << data.theta[1] << ", " << data.theta[2] << ")"
              << std::endl;
    HMCParameters<scalar_type> hmc(in.burnin, in.nsave, in.thin, in.L, (scalar_type)(1.0 / in.L), in.mhtarget);
    std::cout << "\nMCMC settings:\n-----------------\nBurnin   = " << in.burnin << "\nNSave    = " << in.nsave
              << "\nThin     = " << in.thin << "\nLeapfrog = " << in.L << "\n-----------------" << std::endl;
    data.Yh = nonzero_data(hr);
    if (dual) {
      data.n_datasets = 2;
      const int hd[3] = {hr.nx, hr.ny, hr.nz}, sd[3] = {sr.nx, sr.ny, sr.nz};
      const double th[3] = {data.theta[0], data.theta[1], data.theta[2]};
      data.W = kriging_build(hr.data, hd, hr.qto_xyz, sr.data, sd, sr.qto_xyz, th, neighborhood, in.device);
      data.Ys = nonzero_data(sr);
    }
    add_to(hmask, omask);  // global input mask
    FitOptions fo;
    fo.precision = kPrecision;
    fo.state_mode = in.state == "real" ? DR_STATE_REAL : DR_STATE_COMPAT;
    fo.covfun = in.rq ? DR_COV_RQ : DR_COV_RBF;
    fo.device = in.device;
    fo.seed = in.seed;
    fo.monitor = in.monitor;
    fo.output_samples = in.samples;
    fo.update_tau = in.update_tau;
    if (in.embedding != "pow2") {
      const BBox bb = bounding_box(hmask);
      const auto ijk = nonzero_indices_bounded_by_box(hr, bb);
      int mx[3] = {0, 0, 0};
      for (const auto& v : ijk)
        for (int a = 0; a < 3; ++a) mx[a] = std::max(mx[a], v[a]);
      const int img[3] = {hr.nx, hr.ny, hr.nz};
      for (int a = 0; a < 3; ++a)
        fo.embed[a] = in.embedding == "2x" ? std::max(2 * img[a], 2 * mx[a]) : smooth_even(2 * mx[a]);
    }
    McmcSummary<scalar_type> out = fit_model(hr, hmask, omask, data, hmc, samples, fo);
    if (samples.is_open()) {
      samples.close();
      std::cout << in.output_base << "_samples.dat written\n";
    }
    const std::string f_mean = in.output_base + "_posterior_mean.nii.gz";
    const std::string f_var = in.output_base + "_posterior_variance.nii.gz";
    const std::string f_act = in.output_base + "_posterior_activation.nii.gz";
    const std::string f_res = in.output_base + "_residual.nii.gz";
    Image o = omask;
    if (f_mean != in.highres) {
      emplace_nonzero_data(o, out.mode_mu());
      write_nifti(f_mean, o, o.data);
      std::cout << f_mean << " written\n";
    } else {
      std::cerr << "  *** WARNING *** posterior mean image would overwrite data. File not written\n";
    }
    emplace_nonzero_data(o, out.var_mu());
    write_nifti(f_var, o, o.data);
    std::cout << f_var << " written\n";
    emplace_nonzero_data(o, out.activation());
    write_nifti(f_act, o, o.data);
    std::cout << f_act << " written\n";
    std::vector<double> neg = out.mode_mu();
    for (double& v : neg) v = -v;
    emplace_nonzero_data(o, neg);
    apply_mask(hr, o);
    add_to(o, hr);
    apply_mask(o, hr);
    write_nifti(f_res, o, o.data);
    std::cout << f_res << " written\n";
    std::cout << "Post burnin sampling took " << out.sampling_time << " (sec)\nMetropolis-Hastings rate was "
              << (out.mh_rate * 100) << "%\n" << std::endl;
    if (in.samples) {
      std::ofstream st(in.output_base + "_stats.csv");
      if (st.is_open()) st << "MHrate,SamplingTime,StepSize\n" << out.mh_rate << "," << out.sampling_time << ","
                           << hmc.eps_value() << "\n";
    }
  } catch (const std::exception& e) {
    error_status = true;
    std::cerr << "\nException caught with message:\n'" << e.what() << "'\n" << std::endl;
  } catch (...) {
    error_status = true;
    std::cerr << "\nUnknown error\n" << std::endl;
  }
  return error_status ? 1 : 0;
}
