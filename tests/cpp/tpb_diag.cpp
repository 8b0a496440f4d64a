// C++ caller of the adapter: FCIDUMP + alpha-string file -> energy, density
// and a 2-RDM checksum, the shape of the reference CLI's TPB run on one GPU.
// usage: tpb_diag FCIDUMP STRINGS [nb] [eps]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "sbdx_io.hpp"
#include "sbdx_tpb.hpp"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s FCIDUMP STRINGS [nb] [eps]\n", argv[0]);
    return 1;
  }
  try {
    const size_t bl = 20;
    auto fd = sbdx::load_fcidump(argv[1]);
    auto g = sbdx::setup_integrals(fd);
    auto adet = sbdx::load_strings(argv[2], bl, (size_t)g.L);
    const int nb = argc > 3 ? std::atoi(argv[3]) : 10;
    const double eps = argc > 4 ? std::atof(argv[4]) : 1e-8;
    double energy = 0.0;
    std::vector<double> density;
    std::vector<std::vector<double>> one, two;
    sbdx::diag(adet, adet, bl, (size_t)g.L, g.I0, g.I1, g.I2, 100, nb, eps, 1e30, true, energy,
               density, one, two);
    double tr = 0.0, n2 = 0.0;
    for (double d : density) tr += d;
    for (auto &b : two)
      for (double x : b) n2 += x * x;
    std::printf("norb %d nstr %zu\n", g.L, adet.size());
    std::printf("energy %.12f\n", energy);
    std::printf("electrons %.12f\n", tr);
    std::printf("rdm2_norm %.12e\n", std::sqrt(n2));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
  return 0;
}
