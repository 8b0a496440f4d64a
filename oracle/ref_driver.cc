// Driver that exercises the reference sbd-pub TPB path on one MPI rank and
// writes its intermediate results as raw little-endian binaries, so the
// oracle and the HIP library can be checked against the reference itself.
//
// TEST INFRASTRUCTURE ONLY: compiled by oracle/build_ref.sh against the
// reference headers where they lie; outputs go to oracle/_ref/ and to the
// fixture generator under tests/golden/.
//
// usage: ref_driver FCIDUMP STRINGS BIT_LENGTH OUTDIR VECFILE MAX_IT NB EPS DO_RDM
//   VECFILE: N doubles (alpha-major) to push through one sigma, or "-".
#include <chrono>
#include <cstdio>
#include <deque>
#include <fstream>
#include <iostream>
#include <random>

#include "mpi.h"
#include "sbd/sbd.h"

static void dump(const std::string &path, const void *p, size_t bytes) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (bytes) std::fwrite(p, 1, bytes, f);
  std::fclose(f);
}

int main(int argc, char **argv) {
  int provided;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  if (argc < 10) {
    std::fprintf(stderr, "usage: %s FCIDUMP STRINGS BL OUTDIR VEC MAXIT NB EPS RDM\n", argv[0]);
    return 1;
  }
  std::string fcf = argv[1], sf = argv[2], outdir = argv[4], vecf = argv[5];
  size_t bl = std::stoul(argv[3]);
  int max_it = std::atoi(argv[6]), nb = std::atoi(argv[7]);
  double eps = std::atof(argv[8]);
  int do_rdm = std::atoi(argv[9]);
  MPI_Comm comm = MPI_COMM_WORLD;

  sbd::FCIDump fcidump = sbd::LoadFCIDump(fcf);
  int L = 0, N = 0;
  double I0 = 0;
  sbd::oneInt<double> I1;
  sbd::twoInt<double> I2;
  sbd::SetupIntegrals(fcidump, L, N, I0, I1, I2);
  std::vector<std::vector<size_t>> adet;
  sbd::LoadAlphaDets(sf, adet, bl, (size_t)L);
  auto bdet = adet;
  size_t na = adet.size(), nbs = bdet.size(), ndet = na * nbs;
  std::printf("norb %d nelec %d nstr %zu ndet %zu\n", L, N, na, ndet);

  MPI_Comm h_comm, b_comm, t_comm, a_comm;
  sbd::TaskCommunicator(comm, 1, 1, 1, 1, h_comm, b_comm, t_comm, a_comm);
  std::vector<sbd::TaskHelpers> helper;
  auto t0 = std::chrono::steady_clock::now();
  sbd::MakeHelpers(adet, bdet, bl, L, helper, h_comm, b_comm, t_comm, 1, 1);
  sbd::RemakeHelpers(adet, bdet, bl, L, helper, h_comm, b_comm, t_comm, 1, 1);
  auto t1 = std::chrono::steady_clock::now();
  std::printf("helpers %.6f s tasks %zu\n", std::chrono::duration<double>(t1 - t0).count(),
              helper.size());

  // tables: per task, per list (S-alpha, D-alpha, S-beta, D-beta):
  // [n_bra][len per bra][ket ids][cran]
  for (size_t t = 0; t < helper.size(); t++) {
    auto &h = helper[t];
    struct L_ { const char *nm; size_t n; size_t *len; std::vector<size_t *> *ket; std::vector<int *> *cr; int w; };
    L_ lists[4] = {
        {"sa", h.braAlphaEnd - h.braAlphaStart, h.SinglesFromAlphaLen, &h.SinglesFromAlphaSM, &h.SinglesAlphaCrAnSM, 2},
        {"da", h.braAlphaEnd - h.braAlphaStart, h.DoublesFromAlphaLen, &h.DoublesFromAlphaSM, &h.DoublesAlphaCrAnSM, 4},
        {"sb", h.braBetaEnd - h.braBetaStart, h.SinglesFromBetaLen, &h.SinglesFromBetaSM, &h.SinglesBetaCrAnSM, 2},
        {"db", h.braBetaEnd - h.braBetaStart, h.DoublesFromBetaLen, &h.DoublesFromBetaSM, &h.DoublesBetaCrAnSM, 4}};
    for (auto &l : lists) {
      std::vector<int64_t> blob;
      blob.push_back((int64_t)h.taskType);
      blob.push_back((int64_t)l.n);
      for (size_t i = 0; i < l.n; i++) blob.push_back(l.len ? (int64_t)l.len[i] : 0);
      for (size_t i = 0; i < l.n; i++)
        for (size_t j = 0; l.len && j < l.len[i]; j++) blob.push_back((int64_t)(*l.ket)[i][j]);
      for (size_t i = 0; i < l.n; i++)
        for (size_t j = 0; l.len && j < l.len[i] * l.w; j++) blob.push_back((int64_t)(*l.cr)[i][j]);
      dump(outdir + "/table_t" + std::to_string(t) + "_" + l.nm + ".bin", blob.data(),
           blob.size() * sizeof(int64_t));
    }
  }

  std::vector<double> hii;
  sbd::makeQChamDiagTerms(adet, bdet, bl, (size_t)L, helper, I0, I1, I2, hii, h_comm, b_comm, t_comm);
  dump(outdir + "/hii.bin", hii.data(), hii.size() * sizeof(double));

  if (vecf != "-") {
    std::vector<double> v(ndet), s(ndet, 0.0);
    FILE *f = std::fopen(vecf.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), ndet, f) != ndet) {
      std::fprintf(stderr, "cannot read %s\n", vecf.c_str());
      return 2;
    }
    std::fclose(f);
    auto s0 = std::chrono::steady_clock::now();
    sbd::mult(hii, v, s, adet, bdet, bl, (size_t)L, 1, 1, helper, I0, I1, I2, h_comm, b_comm, t_comm);
    auto s1 = std::chrono::steady_clock::now();
    std::printf("sigma %.6f s\n", std::chrono::duration<double>(s1 - s0).count());
    dump(outdir + "/sigma.bin", s.data(), s.size() * sizeof(double));
  }

  if (max_it > 0) {
    std::vector<double> W;
    sbd::BasisInitVector(W, adet, bdet, 1, 1, h_comm, b_comm, t_comm, 0, 0);
    auto d0 = std::chrono::steady_clock::now();
    sbd::Davidson(hii, W, adet, bdet, bl, (size_t)L, 1, 1, helper, I0, I1, I2, h_comm, b_comm,
                  t_comm, max_it, nb, eps, 1.0e30);
    auto d1 = std::chrono::steady_clock::now();
    std::printf("davidson %.6f s\n", std::chrono::duration<double>(d1 - d0).count());
    std::vector<double> HW(W.size(), 0.0);
    sbd::mult(hii, W, HW, adet, bdet, bl, (size_t)L, 1, 1, helper, I0, I1, I2, h_comm, b_comm, t_comm);
    double E = 0.0;
    sbd::InnerProduct(W, HW, E, b_comm);
    std::printf("energy %.17g\n", E);
    dump(outdir + "/W.bin", W.data(), W.size() * sizeof(double));
    dump(outdir + "/energy.bin", &E, sizeof(double));
    if (do_rdm) {
      std::vector<std::vector<double>> one, two;
      sbd::Correlation(W, adet, bdet, bl, (size_t)L, 1, 1, helper, h_comm, b_comm, t_comm, one, two);
      std::vector<double> o1, o2;
      for (auto &x : one) o1.insert(o1.end(), x.begin(), x.end());
      for (auto &x : two) o2.insert(o2.end(), x.begin(), x.end());
      dump(outdir + "/rdm1.bin", o1.data(), o1.size() * sizeof(double));
      dump(outdir + "/rdm2.bin", o2.data(), o2.size() * sizeof(double));
    }
  }
  MPI_Finalize();
  return 0;
}
