// write_kriging_matrix: builds the kriging matrix of an image into the grid of
// a `--space` image and writes it to ./W.csv (src/write_kriging_matrix.cpp:19-111).
// The file layout is the reference's: a header "Data,ColumnIndex,RowCounts,NRow,Ncol",
// then one line per stored value i with the value (as float, default stream
// formatting), its column index, the i-th cumulative row count while i <= nrow,
// and NRow,Ncol on the first line only.
#include <exception>
#include <fstream>
#include <iostream>
#include <string>

#include "kriging_cli.hpp"

using namespace drhost;

int main(int argc, char** argv) {
  const KrigingArgs a = parse_kriging_args(argc, argv);
  if (a.status == KrigingArgs::error) return 1;
  if (a.status == KrigingArgs::help) return 0;
  try {
    const KrigingSetup s = kriging_setup(a, "interpolate");
    std::ofstream out("W.csv");
    if (!out) throw std::runtime_error("cannot write W.csv");
    std::cout << "Writing W:" << std::endl;
    out << "Data,ColumnIndex,RowCounts,NRow,Ncol\n";
    const size_t nnz = s.W.vals.size();
    for (size_t i = 0; i < nnz; ++i) {
      out << (float)s.W.vals[i] << "," << s.W.cols[i] << ",";
      if (i < s.W.rp.size()) out << s.W.rp[i];
      out << ",";
      if (i == 0) out << s.W.nrow << "," << s.W.ncol;
      else out << ",";
      out << "\n";
    }
    out.close();
    if (!out) throw std::runtime_error("write failed: W.csv");
    std::cout << "W.csv written" << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "Exception caught with message:\n'" << e.what() << "'\n" << std::endl;
    return 1;
  } catch (...) {
    std::cerr << "Unknown error\n";
    return 1;
  }
  return 0;
}
