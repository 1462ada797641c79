// rbf_neighborhood: prints the distance at which the RBF kernel falls to rho,
// rbf_inverse(rho; theta) (src/rbf_neighborhood.cpp:10-32,
// impl/kernels.inl:19-29, impl/CommandParser/NeighborhoodCommandParser.inl:17-70).
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
  const int K = 3;
  double theta[K] = {0, 0, 0}, rho = 0;
  if (argc != K + 3) {
    std::cerr << "\nUsage:\n\trbf_neighbhorood rho --covariance f1 f2 f3\n\n";
    return 1;
  }
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    try {
      if (a == "-h" || a == "--help") {
        std::cerr << "\nUsage:\n\trbf_neighbhorood rho --covariance f1 f2 f3\n\n";
        return 0;
      } else if (a == "--theta" || a == "--covariance") {
        if (i + K > argc - 0) throw std::invalid_argument("");
        for (int j = 0; j < K; ++j) theta[j] = std::stod(argv[++i]);
      } else {
        rho = std::stod(a);
      }
    } catch (...) {
      std::cerr << "\nUnrecognized option: " << a << "\n\nUsage:\n\trbf_neighbhorood rho --covariance f1 f2 f3\n\n";
      return 1;
    }
  }
  double r;
  try {
    if (rho <= 0 || rho >= 1) throw std::domain_error("rbf_inverse: inverse kernel only for rho between (0, 1)");
    r = std::pow(-std::log(rho / theta[0]) / theta[1], 1 / theta[2]);
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  std::cout << r << std::endl;
  if (std::isnan(r))
    std::cerr << "\nWarning: make sure that kernel parameters fall within ranges:\n  variance > 0\n  bandwidth > 0\n"
                 "  0 < exponent <= 2"
              << std::endl;
  return 0;
}
