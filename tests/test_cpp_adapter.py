"""The C++ side of the boundary: the adapter headers compile standalone and
against the reference-shaped integral types (CPU), and the C++ caller
tests/cpp/tpb_diag reproduces the published h2o energy on the GPU."""
import gzip
import os
import shutil
import subprocess

import pytest

from conftest import DATA, ROOT

INC = os.path.join(ROOT, "include")
BIN = os.path.join(ROOT, "tests", "cpp", "tpb_diag")


def test_headers_compile(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text("""
#include <vector>
#include "sbdx_tpb.hpp"
#include "sbdx_io.hpp"
struct OneInt { std::vector<double> store; };   // shape of sbd::oneInt
struct TwoInt { std::vector<double> store; };   // shape of sbd::twoInt
int main() {
  sbdx::MultTPBHip<double> m;
  std::vector<std::vector<size_t>> a{{7}};
  OneInt i1; TwoInt i2;
  if (false) m.Init(a, a, 20, 3, 0.0, i1, i2);
  return 0;
}
""")
    gxx = shutil.which("g++")
    assert gxx, "g++ missing"
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", INC, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_reader_matches_python(tmp_path):
    """sbdx_io.hpp's FCIDUMP + string readers give the same integrals and
    strings as the Python readers (compiled and run on the CPU)."""
    fcf = tmp_path / "fcidump.txt"
    with gzip.open(os.path.join(DATA, "n2_fcidump.txt.gz"), "rt") as g:
        fcf.write_text(g.read())
    src = tmp_path / "r.cpp"
    src.write_text("""
#include <cstdio>
#include "sbdx_io.hpp"
int main(int argc, char** argv) {
  auto fd = sbdx::load_fcidump(argv[1]);
  auto g = sbdx::setup_integrals(fd);
  auto s = sbdx::load_strings(argv[2], 20, g.L);
  double a = 0, b = 0;
  for (double x : g.I1.store) a += x;
  for (size_t i = 0; i < g.I2.store.size(); i++) b += g.I2.store[i] * (double)(i % 7 + 1);
  std::printf("%d %d %.17g %.17g %.17g %zu %zu %zu\\n", g.L, g.N, g.I0, a, b, s.size(),
              s.front()[0], s.back()[0]);
}
""")
    exe = tmp_path / "r"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(fcf), os.path.join(DATA, "n2-1em3-alpha.txt")],
                         capture_output=True, text=True, check=True).stdout.split()
    import numpy as np

    from sbdx import io
    fd = io.read_fcidump(str(fcf))
    I0, I1, I2 = io.setup_integrals(fd)
    s = io.read_strings(os.path.join(DATA, "n2-1em3-alpha.txt"), fd["norb"])
    assert int(out[0]) == fd["norb"] and int(out[1]) == fd["nelec"]
    assert float(out[2]) == I0
    assert abs(float(out[3]) - I1.sum()) < 1e-12 * abs(I1).sum()
    w = (np.arange(I2.size) % 7 + 1).astype(float)
    assert abs(float(out[4]) - (I2 * w).sum()) < 1e-12 * abs(I2 * w).sum()
    assert int(out[5]) == s.size and int(out[6]) == int(s[0]) and int(out[7]) == int(s[-1])


@pytest.mark.gpu
def test_cpp_caller_h2o(tmp_path):
    assert os.path.exists(BIN), "tests/cpp/tpb_diag not built (run __graft_entry__.build())"
    fcf = tmp_path / "fcidump.txt"
    with gzip.open(os.path.join(DATA, "h2o_fcidump.txt.gz"), "rt") as g:
        fcf.write_text(g.read())
    print("[cpp caller] h2o-1em3", flush=True)
    r = subprocess.run([BIN, str(fcf), os.path.join(DATA, "h2o-1em3-alpha.txt"), "10", "1e-8"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = dict(line.split()[:2] for line in r.stdout.strip().splitlines())
    assert abs(float(vals["energy"]) - (-76.23594663)) < 6e-9
    assert abs(float(vals["electrons"]) - 10.0) < 1e-9
    assert float(vals["hii_check"]) < 1e-12  # run() applies the caller's hii


REF = os.environ.get("SBD_REFERENCE", "/root/reference")
MPICXX = "/opt/conda/bin/mpicxx"


@pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "include", "sbd")) and os.path.exists(MPICXX)),
                    reason="needs the reference tree and the MPI compiler of the build image")
def test_adapter_compiles_against_reference_headers(tmp_path):
    """The drop-in branch (the SBDX_HIP lines a maintainer adds to
    chemistry/tpb/sbdiag.h) compiled together with the reference's own
    headers and integral types: sbd::oneInt<double> / sbd::twoInt<double>
    (chemistry/basic/integrals.h:11-90) and the variable types of
    sbd::tpb::diag (sbdiag.h:162-187, 822), with the reference build's
    compiler and flags (oracle/build_ref.sh) and -DSBDX_HIP."""
    src = tmp_path / "branch.cc"
    src.write_text("""
// the include sequence of the reference app
// (apps/chemistry_tpb_selected_basis_diagonalization/main.cc:1-13)
#include <iostream>
#include <fstream>
#include <chrono>
#include <random>
#include <deque>
#include <unistd.h>
#define _USE_MATH_DEFINES
#include <cmath>
#include "sbd/sbd.h"
#include "mpi.h"
#ifdef SBDX_HIP
#include "sbdx_tpb.hpp"
#endif
// the variables of sbd::tpb::diag the branch uses, with their types there
void branch(const std::vector<std::vector<size_t>> &adet, const std::vector<std::vector<size_t>> &bdet,
            size_t bit_length, int L, double I0, const sbd::oneInt<double> &I1,
            const sbd::twoInt<double> &I2, int max_it, int max_nb, double eps, double max_time,
            std::vector<double> &W, std::vector<double> &C, std::vector<std::vector<double>> &one_p_rdm,
            std::vector<std::vector<double>> &two_p_rdm) {
#ifdef SBDX_HIP
  sbdx::MultTPBHip<double> device_mult;
  device_mult.Init(adet, bdet, bit_length, static_cast<size_t>(L), I0, I1, I2);
  std::vector<double> hii;
  device_mult.makeQChamDiagTerms(hii);
  sbdx::Davidson(hii, W, device_mult, max_it, max_nb, eps, max_time);
  device_mult.run(hii, W, C);
  device_mult.correlation(W, one_p_rdm, two_p_rdm);
#endif
}
int main() { return 0; }
""")
    env = dict(os.environ, MPICH_CXX="g++")
    r = subprocess.run([MPICXX, "-std=c++17", "-fopenmp", "-w", "-DSBDX_HIP", "-I", os.path.join(REF, "include"),
                        "-I", "/opt/conda/include", "-I", INC, "-c", str(src), "-o", str(tmp_path / "branch.o")],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
