// Shared command line and setup of the kriging utilities `interpolate` and
// `write_kriging_matrix` (src/interpolate.cpp:19-56,
// src/write_kriging_matrix.cpp:20-56, impl/CommandParser/KrigingCommandParser.inl).
// Both krige the nonzero voxels of an input image onto the nonzero voxels of a
// `--space` image with the RBF kernel (variance 1, bandwidth from --fwhm and
// --exponent), building the sparse weight matrix on the device.
#pragma once

#include <algorithm>
#include <cmath>
#include <iostream>
#include <string>

#include "engine.hpp"
#include "nifti.hpp"

namespace drhost {

struct KrigingArgs {
  enum Status { ok, error, help };
  Status status = ok;
  std::string caller, image, space;
  double exponent = 1.99999, fwhm = 6.0, radius = -1.0;  // KrigingCommandParser.inl:58-61
  int device = 0;
};

inline void kriging_usage(const std::string& caller) {
  std::cerr << "\nUsage:\n\t" << caller << " path/to/img --space path/to/img2 <options>\n\n";
}

inline void kriging_help(const std::string& caller) {
  kriging_usage(caller);
  std::cerr << "Options:\n"
            << "--exponent  float  kernel exponent\n"
            << "--fwhm      float  kernel full width (mm) at half max\n"
            << "--radius    float  extent (mm) of the local kriging interpolation\n"
            << "--device    int    GPU ordinal for the kriging build\n"
            << "\nimg is a valid NIfTI file. Kriging runs with a radial basis kernel of unit\n"
            << "variance. Default parameter values are [fwhm = 6] and [exponent = 2]; any exponent\n"
            << "in (0, 2] is allowed. The default radius is 3x the largest voxel dimension of the\n"
            << "'--space' image. The output is masked by the '--space' image.\n\n";
}

// KrigingCommandParser<T>::KrigingCommandParser(argc, argv) (KrigingCommandParser.inl:66-176):
// same flags, defaults, messages and precedence; a flag missing its value is an error.
inline KrigingArgs parse_kriging_args(int argc, char** argv) {
  KrigingArgs a;
  a.caller = argv[0];
  const std::string bad = a.caller + ": Warning: bad parameter value ";
  if (argc < 2) {
    kriging_usage(a.caller);
    a.status = KrigingArgs::error;
  }
  for (int i = 1; i < argc && a.status == KrigingArgs::ok; ++i) {
    const std::string arg = argv[i];
    auto number = [&](double& dst, const char* what) {
      if (i + 1 < argc) {
        ++i;
        try {
          dst = std::stod(argv[i]);
          return true;
        } catch (...) {
        }
      }
      std::cerr << bad << "(" << what << " = " << (i < argc ? argv[i] : "") << ")\n";
      a.status = KrigingArgs::error;
      return false;
    };
    if (arg == "-h" || arg == "--help") {
      a.status = KrigingArgs::help;
    } else if (arg == "--exponent" || arg == "-e") {
      if (number(a.exponent, "exponent") && (a.exponent <= 0 || a.exponent > 2)) {
        std::cerr << bad << "(exponent = " << argv[i] << ")\n";
        a.status = KrigingArgs::error;
      }
    } else if (arg == "--fwhm" || arg == "-f") {
      if (number(a.fwhm, "fwhm")) a.fwhm = std::fabs(a.fwhm);
    } else if (arg == "--radius" || arg == "--neighborhood" || arg == "-r") {
      if (number(a.radius, "radius")) a.radius = std::fabs(a.radius);
    } else if (arg == "--device") {
      double d = 0;
      if (number(d, "device")) a.device = (int)d;
    } else if (arg == "--space") {
      if (i + 1 < argc && is_nifti_file(argv[i + 1])) {
        a.space = argv[++i];
      } else {
        std::cerr << bad << "(output space image is not a NIfTI file)\n";
        a.status = KrigingArgs::error;
      }
    } else if (is_nifti_file(arg)) {
      a.image = arg;
    } else {
      std::cerr << a.caller << ": Warning: Unrecognized option: " << arg << "\n";
      a.status = KrigingArgs::error;
    }
  }
  if (a.status == KrigingArgs::help) {
    kriging_help(a.caller);
  } else if (a.status == KrigingArgs::ok && a.image.empty()) {
    std::cerr << "\n" << a.caller << ": Error: User must supply a NIfTI file input\n";
    a.status = KrigingArgs::error;
  } else if (a.status == KrigingArgs::ok && a.space.empty()) {
    std::cerr << "\n" << a.caller << ": Error: User must supply a --space NIfTI file input\n";
    a.status = KrigingArgs::error;
  }
  if (a.status == KrigingArgs::error) std::cerr << "\nSee " << a.caller << " --help for more information\n";
  return a;
}

// kernels::rbf_fwhm_to_bandwidth (impl/kernels.inl:45-50): psi = log 2 / (fwhm / 2)^nu
inline double rbf_fwhm_to_bandwidth(double fwhm, double exponent) {
  return std::log(2.0) / std::pow(fwhm / 2.0, exponent);
}

struct KrigingSetup {
  Image input, space;
  Csr W;  // rows: nonzero voxels of `space`; columns: nonzero voxels of `input`
};

// The common body of both tools (interpolate.cpp:37-71): read both images
// (float or double data, as the reference requires), default the radius to
// 3x the largest voxel dimension of the space image, and build W on the device.
inline KrigingSetup kriging_setup(const KrigingArgs& a, const char* tool) {
  KrigingSetup s;
  s.input = read_nifti(a.image);
  s.space = read_nifti(a.space);
  if (s.input.datatype != 16 && s.input.datatype != 64)
    throw std::runtime_error(std::string(tool) + ": input image is of unrecognized data type");
  double radius = a.radius;
  if (radius <= 0) {
    const auto v = voxel_dimensions(s.space.qto_xyz);
    radius = 3.0 * std::max({v[0], v[1], v[2]});
  }
  const double theta[3] = {1.0, rbf_fwhm_to_bandwidth(a.fwhm, a.exponent), a.exponent};
  const int hd[3] = {s.input.nx, s.input.ny, s.input.nz};
  const int sd[3] = {s.space.nx, s.space.ny, s.space.nz};
  s.W = kriging_build(s.input.data, hd, s.input.qto_xyz, s.space.data, sd, s.space.qto_xyz, theta, radius,
                      a.device);
  return s;
}

}  // namespace drhost
