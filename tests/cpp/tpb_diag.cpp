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
    // run(hii, Wk, Wb) with the caller's hii: a diagonal shifted on every
    // 7th determinant must change H x by exactly that shift times x there
    double hdiff = 0.0;
    {
      sbdx::MultTPBHip<double> op;
      op.Init(adet, adet, bl, (size_t)g.L, g.I0, g.I1, g.I2);
      std::vector<double> hii;
      op.makeQChamDiagTerms(hii);
      const size_t n = hii.size();
      std::vector<double> x(n), y1(n, 0.0), y2(n, 0.0), h2 = hii;
      for (size_t i = 0; i < n; i++) x[i] = std::sin(0.37 * (double)i + 0.1);
      for (size_t i = 0; i < n; i += 7) h2[i] += 0.5 + 1e-3 * (double)(i % 11);
      op.run(hii, x, y1);
      op.run(h2, x, y2);
      for (size_t i = 0; i < n; i++) hdiff = std::fmax(hdiff, std::fabs((y2[i] - y1[i]) - (h2[i] - hii[i]) * x[i]));
    }
    std::printf("norb %d nstr %zu\n", g.L, adet.size());
    std::printf("hii_check %.3e\n", hdiff);
    std::printf("energy %.12f\n", energy);
    std::printf("electrons %.12f\n", tr);
    std::printf("rdm2_norm %.12e\n", std::sqrt(n2));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
  return 0;
}
