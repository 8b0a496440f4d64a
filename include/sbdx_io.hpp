// sbdx_io.hpp -- header-only readers for the reference's input formats, for
// C++ callers that use the adapter without the reference's own framework.
//
//   load_fcidump     framework/fcidump.h:29-141 (header until &END, then
//                    "value i j k l" rows, 1-based; 0 0 0 0 = core energy)
//   setup_integrals  chemistry/basic/makeintegrals.h:123-174 (RHF): oneInt
//                    store (2L)^2 row-major, twoInt 8-fold compressed store
//   load_strings     chemistry/basic/makedeterminants.h:160-185 with
//                    framework/bit_manipulation.h:215-229 (rightmost char is
//                    orbital 0) and the sort/unique of :149-168, 381-395
//   save_wavefunction / load_wavefunction  the one-rank wavefunction file of
//                    chemistry/tpb/restart.h:20-67 and :414-660
//   write_rdm_text   the 1pRDM.txt / 2pRDM.txt files and
//   energy_decomposition  the zero-/one-/two-body energies of the driver
//                    apps/chemistry_tpb_selected_basis_diagonalization/
//                    main.cc:264-318
#ifndef SBDX_IO_HPP
#define SBDX_IO_HPP

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <numeric>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace sbdx {

struct FCIDump {
  std::map<std::string, std::string> header;
  std::vector<std::tuple<double, int, int, int, int>> integrals;
};

inline std::string trim(const std::string &s) {
  const auto b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return "";
  const auto e = s.find_last_not_of(" \t\r");
  return s.substr(b, e - b + 1);
}

inline FCIDump load_fcidump(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("sbdx: cannot open FCIDUMP " + path);
  FCIDump fd;
  std::string line, head;
  bool in_header = true;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '!') continue;
    if (in_header) {
      if (line.find("&END") != std::string::npos) {
        in_header = false;
        continue;
      }
      const auto p = line.find("&FCI");
      if (p != std::string::npos) line.erase(p, 4);
      head += line;
      continue;
    }
    std::stringstream ss(line);
    double v;
    int i, j, k, l;
    if (ss >> v >> i >> j >> k >> l) fd.integrals.emplace_back(v, i, j, k, l);
  }
  std::stringstream hs(head);
  std::string tok;
  while (std::getline(hs, tok, ',')) {
    const auto eq = tok.find('=');
    if (eq != std::string::npos) fd.header[trim(tok.substr(0, eq))] = trim(tok.substr(eq + 1));
  }
  if (!fd.header.count("NORB")) throw std::runtime_error("sbdx: FCIDUMP without NORB");
  return fd;
}

struct Integrals {
  struct Store {
    std::vector<double> store;
  };
  int L = 0, N = 0;
  double I0 = 0.0;
  Store I1, I2;  // same `store` layouts as the reference's oneInt / twoInt
};

inline Integrals setup_integrals(const FCIDump &fd) {
  Integrals g;
  g.L = std::atoi(fd.header.at("NORB").c_str());
  if (fd.header.count("NELEC")) g.N = std::atoi(fd.header.at("NELEC").c_str());
  const int L = g.L, L2 = 2 * L;
  const size_t B = (size_t)L * (L + 1) / 2;
  g.I1.store.assign((size_t)L2 * L2, 0.0);
  g.I2.store.assign(B * (B + 1) / 2, 0.0);
  auto tri = [](size_t a, size_t b) { return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };
  for (const auto &t : fd.integrals) {
    const double v = std::get<0>(t);
    const int i = std::get<1>(t), j = std::get<2>(t), k = std::get<3>(t), l = std::get<4>(t);
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      g.I0 = v;
    } else if (k == 0 && l == 0) {
      const int a = 2 * (i - 1), b = 2 * (j - 1);
      g.I1.store[(size_t)a * L2 + b] = v;
      g.I1.store[(size_t)(a + 1) * L2 + b + 1] = v;
      g.I1.store[(size_t)b * L2 + a] = v;
      g.I1.store[(size_t)(b + 1) * L2 + a + 1] = v;
    } else {
      g.I2.store[tri(tri(i - 1, j - 1), tri(k - 1, l - 1))] = v;
    }
  }
  return g;
}

inline std::vector<std::vector<size_t>> load_strings(const std::string &path, size_t bit_length,
                                                     size_t L) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("sbdx: cannot open string file " + path);
  const size_t nw = (L + bit_length - 1) / bit_length;
  std::vector<std::vector<size_t>> out;
  std::string s;
  while (std::getline(in, s)) {
    s = trim(s);
    if (s.empty()) continue;
    if (s.size() < L) throw std::runtime_error("sbdx: string shorter than norb: " + s);
    std::vector<size_t> w(nw, 0);
    for (size_t i = 0; i < L; i++)
      if (s[s.size() - 1 - i] == '1') w[i / bit_length] |= (size_t)1 << (i % bit_length);
    out.push_back(w);
  }
  std::sort(out.begin(), out.end(), [](const std::vector<size_t> &a, const std::vector<size_t> &b) {
    for (size_t n = a.size(); n > 0; n--)
      if (a[n - 1] != b[n - 1]) return a[n - 1] < b[n - 1];
    return false;
  });
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

// 1pRDM.txt and 2pRDM.txt in directory `dir`: "i j value" for i, j (outer
// to inner) with one[0][i+L*j] + one[1][i+L*j], and "i j k l value" with
// the four spin blocks of two[.][i+L*j+L^2*(k+L*l)] summed in block order,
// through streams of precision 16 (the driver's text format).
inline void write_rdm_text(const std::string &dir, size_t L, const std::vector<std::vector<double>> &one,
                           const std::vector<std::vector<double>> &two) {
  std::ofstream f1(dir + "/1pRDM.txt");
  f1.precision(16);
  for (size_t i = 0; i < L; i++)
    for (size_t j = 0; j < L; j++) f1 << i << " " << j << " " << one[0][i + L * j] + one[1][i + L * j] << "\n";
  std::ofstream f2(dir + "/2pRDM.txt");
  f2.precision(16);
  for (size_t i = 0; i < L; i++)
    for (size_t j = 0; j < L; j++)
      for (size_t k = 0; k < L; k++)
        for (size_t l = 0; l < L; l++) {
          const size_t x = i + L * j + L * L * (k + L * l);
          f2 << i << " " << j << " " << k << " " << l << " " << two[0][x] + two[1][x] + two[2][x] + two[3][x]
             << "\n";
        }
  if (!f1 || !f2) throw std::runtime_error("cannot write the RDM text files in " + dir);
}

struct EnergyParts {
  double zero = 0.0, one = 0.0, two = 0.0;
};

// I0; sum_ij h_ij (one[0] + one[1])[i+L*j]; 1/2 sum_ijkl (ik|jl)
// sum_s two[s][i+L*j+L^2*(k+L*l)], with h and (ik|jl) read from the
// integral stores (spin-orbitals 2i, 2k, 2j, 2l).
inline EnergyParts energy_decomposition(const Integrals &g, const std::vector<std::vector<double>> &one,
                                        const std::vector<std::vector<double>> &two) {
  const size_t L = (size_t)g.L, L2 = 2 * L;
  auto tri = [](size_t a, size_t b) { return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };
  EnergyParts e;
  e.zero = g.I0;
  for (size_t i = 0; i < L; i++)
    for (size_t j = 0; j < L; j++) e.one += g.I1.store[(2 * i) * L2 + 2 * j] * (one[0][i + L * j] + one[1][i + L * j]);
  for (size_t i = 0; i < L; i++)
    for (size_t j = 0; j < L; j++)
      for (size_t k = 0; k < L; k++)
        for (size_t l = 0; l < L; l++) {
          const size_t x = i + L * j + L * L * (k + L * l);
          const double v = g.I2.store[tri(tri(i, k), tri(j, l))];
          for (int s = 0; s < 4; s++) e.two += 0.5 * v * two[s][x];
        }
  return e;
}

// Wavefunction file of one rank (file name + "000000"): size_t numbers of
// alpha strings, beta strings and words per string; the strings' words (the
// reference's determinant words), alpha then beta; W (alpha-major).
template <typename ElemT>
inline void save_wavefunction(const std::string &file, const std::vector<std::vector<size_t>> &adet,
                              const std::vector<std::vector<size_t>> &bdet, const std::vector<ElemT> &W) {
  if (W.size() != adet.size() * bdet.size()) throw std::invalid_argument("W does not match the string sets");
  std::ofstream f(file + "000000", std::ios::binary);
  size_t head[3] = {adet.size(), bdet.size(), adet.empty() ? 0 : adet[0].size()};
  f.write(reinterpret_cast<const char *>(head), sizeof head);
  for (const auto &d : adet) f.write(reinterpret_cast<const char *>(d.data()), sizeof(size_t) * head[2]);
  for (const auto &d : bdet) f.write(reinterpret_cast<const char *>(d.data()), sizeof(size_t) * head[2]);
  f.write(reinterpret_cast<const char *>(W.data()), sizeof(ElemT) * W.size());
  if (!f) throw std::runtime_error("cannot write " + file + "000000");
}

// W over (adet, bdet) from a wavefunction file: zero, then each stored
// (alpha, beta) coefficient whose strings both occur in the sets at their
// position (restart from the intersection of the two determinant sets).
template <typename ElemT>
inline void load_wavefunction(const std::string &file, const std::vector<std::vector<size_t>> &adet,
                              const std::vector<std::vector<size_t>> &bdet, std::vector<ElemT> &W) {
  std::ifstream f(file + "000000", std::ios::binary);
  if (!f) throw std::runtime_error("failed to open input file " + file + "000000");
  size_t head[3];
  f.read(reinterpret_cast<char *>(head), sizeof head);
  std::vector<std::vector<size_t>> la(head[0], std::vector<size_t>(head[2])), lb(head[1], std::vector<size_t>(head[2]));
  for (auto &d : la) f.read(reinterpret_cast<char *>(d.data()), sizeof(size_t) * head[2]);
  for (auto &d : lb) f.read(reinterpret_cast<char *>(d.data()), sizeof(size_t) * head[2]);
  std::vector<ElemT> lw(head[0] * head[1]);
  f.read(reinterpret_cast<char *>(lw.data()), sizeof(ElemT) * lw.size());
  if (!f) throw std::runtime_error(file + "000000: truncated");
  std::map<std::vector<size_t>, size_t> ia, ib;
  for (size_t k = 0; k < adet.size(); k++) ia[adet[k]] = k;
  for (size_t k = 0; k < bdet.size(); k++) ib[bdet[k]] = k;
  W.assign(adet.size() * bdet.size(), ElemT(0));
  for (size_t a = 0; a < head[0]; a++) {
    const auto pa = ia.find(la[a]);
    if (pa == ia.end()) continue;
    for (size_t b = 0; b < head[1]; b++) {
      const auto pb = ib.find(lb[b]);
      if (pb != ib.end()) W[pa->second * bdet.size() + pb->second] += lw[a * head[1] + b];
    }
  }
}

// ---------------------------------------------------------------- carry-over
// The carry-over string selection of the reference's SQD loop on one rank
// (chemistry/tpb/rdmat.h:10-230, chemistry/tpb/extend.h:14-324, applied by
// chemistry/tpb/sbdiag.h:692-735 after the final energy).  Strings are
// masks (orbital i -> bit i), W is alpha-major (W[ia*nb + ib]).  String
// weights: D_a[ia] = sum_ib W[ia,ib]^2, D_b[ib] = sum_ia W[ia,ib]^2.
//   type 1, threshold > 0 (ratio 0): the fewest strings, heaviest first, whose
//     weights sum to more than 1 - threshold;
//   type 1, ratio > 0: the int(ratio * n) heaviest strings, and the weight
//     left out (truncated_alpha / truncated_beta);
//   type 2: either of these, then every selected string with all its single
//     excitations, sorted, without duplicates;
//   type 3: the strings of the determinants with W^2 > threshold, with all
//     their singles, sorted, without duplicates, and the summed weight of
//     those determinants (total_weight).
// Equal weights keep the lower index first (the reference's std::sort leaves
// their order unspecified); where the threshold rule finds no prefix above
// 1 - threshold (W not normalized) every string is kept (the reference
// reads past the end of its list there).
struct Carryover {
  std::vector<uint64_t> adet, bdet;
  double truncated_alpha = 0.0, truncated_beta = 0.0, total_weight = 0.0;
};

namespace carryover_detail {
inline std::vector<size_t> by_weight(const std::vector<double> &D) {
  std::vector<size_t> o(D.size());
  std::iota(o.begin(), o.end(), size_t(0));
  std::stable_sort(o.begin(), o.end(), [&](size_t i, size_t j) { return D[i] > D[j]; });
  return o;
}
inline std::vector<uint64_t> threshold_select(const std::vector<uint64_t> &s, const std::vector<double> &D,
                                              double threshold) {
  const auto o = by_weight(D);
  double sum = 0.0;
  size_t k = 0;
  while (k < o.size()) {
    sum += D[o[k++]];
    if (sum > 1.0 - threshold) break;
  }
  std::vector<uint64_t> r(k);
  for (size_t i = 0; i < k; i++) r[i] = s[o[i]];
  return r;
}
inline std::vector<uint64_t> ratio_select(const std::vector<uint64_t> &s, const std::vector<double> &D,
                                          double ratio, double &truncated) {
  const auto o = by_weight(D);
  const size_t k = std::min(o.size(), (size_t)(ratio * (double)D.size()));
  std::vector<uint64_t> r(k);
  double sum = 0.0;
  for (size_t i = 0; i < k; i++) {
    r[i] = s[o[i]];
    sum += D[o[i]];
  }
  truncated = 1.0 - sum;
  return r;
}
// the strings and all their single excitations, ascending, unique
inline std::vector<uint64_t> singles_extend(const std::vector<uint64_t> &s, int norb) {
  std::vector<uint64_t> r(s.begin(), s.end());
  for (uint64_t m : s)
    for (int c = 0; c < norb; c++) {
      if (!((m >> c) & 1)) continue;
      for (int o = 0; o < norb; o++)
        if (!((m >> o) & 1)) r.push_back((m & ~(1ull << c)) | (1ull << o));
    }
  std::sort(r.begin(), r.end());
  r.erase(std::unique(r.begin(), r.end()), r.end());
  return r;
}
}  // namespace carryover_detail

inline Carryover carryover(const std::vector<double> &W, const std::vector<uint64_t> &adet,
                           const std::vector<uint64_t> &bdet, int norb, int type, double threshold,
                           double ratio) {
  namespace cd = carryover_detail;
  const size_t na = adet.size(), nb = bdet.size();
  if (W.size() != na * nb) throw std::invalid_argument("carryover: W size");
  Carryover r;
  if (type == 1 || type == 2) {
    std::vector<double> Da(na, 0.0), Db(nb, 0.0);
    for (size_t ia = 0; ia < na; ia++)
      for (size_t ib = 0; ib < nb; ib++) {
        const double w2 = W[ia * nb + ib] * W[ia * nb + ib];
        Da[ia] += w2;
      }
    for (size_t ib = 0; ib < nb; ib++)
      for (size_t ia = 0; ia < na; ia++) Db[ib] += W[ia * nb + ib] * W[ia * nb + ib];
    if (ratio == 0.0) {
      r.adet = cd::threshold_select(adet, Da, threshold);
      r.bdet = cd::threshold_select(bdet, Db, threshold);
    } else {
      r.adet = cd::ratio_select(adet, Da, ratio, r.truncated_alpha);
      r.bdet = cd::ratio_select(bdet, Db, ratio, r.truncated_beta);
    }
    if (type == 2) {
      r.adet = cd::singles_extend(r.adet, norb);
      r.bdet = cd::singles_extend(r.bdet, norb);
    }
  } else if (type == 3) {
    std::vector<char> ua(na, 0), ub(nb, 0);
    for (size_t ia = 0; ia < na; ia++)
      for (size_t ib = 0; ib < nb; ib++) {
        const double w2 = W[ia * nb + ib] * W[ia * nb + ib];
        if (w2 > threshold) {
          r.total_weight += w2;
          ua[ia] = ub[ib] = 1;
        }
      }
    std::vector<uint64_t> sa, sb;
    for (size_t ia = 0; ia < na; ia++)
      if (ua[ia]) sa.push_back(adet[ia]);
    for (size_t ib = 0; ib < nb; ib++)
      if (ub[ib]) sb.push_back(bdet[ib]);
    r.adet = cd::singles_extend(sa, norb);
    r.bdet = cd::singles_extend(sb, norb);
  }
  return r;
}

}  // namespace sbdx

#endif  // SBDX_IO_HPP
