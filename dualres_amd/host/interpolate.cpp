// interpolate: local kriging of an image into the grid of a `--space` image
// (src/interpolate.cpp:19-101).  W is built on the device; Y_interp = W Y over
// the nonzero voxels of the input, written into the nonzero voxels of the
// space image as <input basename>_<fwhm>mm_fwhm_kriged.nii, keeping the space
// image's float or double data type.
#include <exception>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "kriging_cli.hpp"

using namespace drhost;

int main(int argc, char** argv) {
  const KrigingArgs a = parse_kriging_args(argc, argv);
  if (a.status == KrigingArgs::error) return 1;
  if (a.status == KrigingArgs::help) return 0;
  try {
    KrigingSetup s = kriging_setup(a, "interpolate");
    if (s.space.datatype != 16 && s.space.datatype != 64)
      throw std::runtime_error("emplace_nonzero_data: unrecognized image data type");
    const std::vector<double> y = nonzero_data(s.input);
    if ((long)y.size() != s.W.ncol) throw std::logic_error("interpolate: input voxels do not match W columns");
    std::vector<double> yi(s.W.nrow, 0.0);
    for (long r = 0; r < s.W.nrow; ++r) {
      double acc = 0.0;
      for (int64_t t = s.W.rp[r]; t < s.W.rp[r + 1]; ++t) acc += s.W.vals[t] * y[s.W.cols[t]];
      yi[r] = acc;
    }
    emplace_nonzero_data(s.space, yi);
    std::ostringstream name;
    name << nifti_basename(a.image) << "_" << (int)a.fwhm << "mm_fwhm_kriged.nii";
    write_nifti(name.str(), s.space, s.space.data, s.space.datatype);
  } catch (const std::exception& e) {
    std::cerr << "Exception caught with message:\n'" << e.what() << "'\n" << std::endl;
    return 1;
  } catch (...) {
    std::cerr << "Unknown error\n";
    return 1;
  }
  return 0;
}
