"""Wavefunction files of the reference (chemistry/tpb/restart.h: one-rank
save :20-67, load :414-660): byte layout, round trip, restart on changed
string sets (coefficients of the shared determinants), and the C++ header
against the Python mirror."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")


def _sets(norb=22, seed=3):
    from sbdx import instances
    a = instances.make_strings(norb, 4, 30, seed)
    b = instances.make_strings(norb, 4, 25, seed + 1)
    return a, b


def test_layout_and_round_trip(tmp_path):
    from sbdx import io
    L = 22  # two words per string at bit_length 20
    a, b = _sets(L)
    W = np.random.default_rng(0).standard_normal(a.size * b.size)
    f = str(tmp_path / "wf")
    io.save_wavefunction(f, L, a, b, W)
    raw = open(f + "000000", "rb").read()
    head = np.frombuffer(raw[:24], dtype=np.uint64)
    assert list(head) == [a.size, b.size, 2]
    assert len(raw) == 24 + 8 * 2 * (a.size + b.size) + 8 * a.size * b.size
    # first alpha string's words: orbital i at bit i%20 of word i//20
    w0 = np.frombuffer(raw[24:40], dtype=np.uint64)
    assert int(w0[0]) | (int(w0[1]) << 20) == int(a[0])
    assert np.array_equal(np.frombuffer(raw[-8 * W.size:], dtype=np.float64), W)
    assert np.array_equal(io.load_wavefunction(f, L, a, b), W)


def test_load_onto_changed_sets(tmp_path):
    from sbdx import io
    L = 22
    a, b = _sets(L)
    W = np.arange(a.size * b.size, dtype=np.float64) + 1.0
    f = str(tmp_path / "wf")
    io.save_wavefunction(f, L, a, b, W)
    # new sets: alpha without its first 5 strings plus 3 strings not saved,
    # beta only its first 10 strings
    extra = np.array([(1 << 4) | (1 << 9) | (1 << 13) | (1 << 21), (1 << 2) | (1 << 3) | (1 << 17) | (1 << 20),
                      (1 << 0) | (1 << 6) | (1 << 11) | (1 << 19)], dtype=np.uint64)
    a2 = io.sort_unique(np.concatenate([a[5:], extra]))
    b2 = b[:10].copy()
    W2 = io.load_wavefunction(f, L, a2, b2).reshape(a2.size, b2.size)
    Wm = W.reshape(a.size, b.size)
    pos_a = {int(s): k for k, s in enumerate(a)}
    for i, s in enumerate(a2):
        for j in range(b2.size):
            want = Wm[pos_a[int(s)], j] if int(s) in pos_a else 0.0
            assert W2[i, j] == want


def test_cpp_matches_python(tmp_path):
    """C++ load of a Python-written file, saved again by C++: identical bytes;
    and the C++ loader on changed sets agrees with the Python one."""
    from sbdx import io
    gxx = shutil.which("g++")
    assert gxx, "g++ missing"
    L = 22
    a, b = _sets(L)
    W = np.random.default_rng(5).standard_normal(a.size * b.size)
    io.save_wavefunction(str(tmp_path / "py"), L, a, b, W)
    io.write_strings(str(tmp_path / "a.txt"), a, L)
    io.write_strings(str(tmp_path / "b.txt"), b, L)
    io.write_strings(str(tmp_path / "a2.txt"), a[3:], L)
    src = tmp_path / "w.cpp"
    src.write_text('''
#include <cstdio>
#include "sbdx_io.hpp"
int main(int argc, char **argv) {
  const size_t L = 22, bl = 20;
  auto a = sbdx::load_strings(argv[1], bl, L), b = sbdx::load_strings(argv[2], bl, L);
  auto a2 = sbdx::load_strings(argv[3], bl, L);
  std::vector<double> W, W2;
  sbdx::load_wavefunction(argv[4], a, b, W);
  sbdx::save_wavefunction(argv[5], a, b, W);
  sbdx::load_wavefunction(argv[4], a2, b, W2);
  for (double v : W2) std::printf("%.17g\\n", v);
  return 0;
}
''')
    exe = tmp_path / "w"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path / "a.txt"), str(tmp_path / "b.txt"), str(tmp_path / "a2.txt"),
                          str(tmp_path / "py"), str(tmp_path / "cpp")],
                         check=True, capture_output=True, text=True).stdout.split()
    assert (tmp_path / "cpp000000").read_bytes() == (tmp_path / "py000000").read_bytes()
    W2 = np.array([float(v) for v in out])
    assert np.array_equal(W2, io.load_wavefunction(str(tmp_path / "py"), L, a[3:], b))


@pytest.mark.gpu
def test_diag_restart_from_saved_wavefunction(tmp_path):
    """sbdx.diag saves its solution; a second diag that loads it starts at
    the converged vector and needs only a few steps for the same energy."""
    import sbdx
    from conftest import DATA
    fcid = os.path.join(DATA, "h2o_fcidump.txt.gz")
    strs = os.path.join(DATA, "h2o-1em3-alpha.txt")
    f = str(tmp_path / "wf")
    print("[wavefunction] h2o-1em3 diag, save", flush=True)
    r1 = sbdx.diag(fcid, strs, eps=1e-9, do_rdm=False, savename=f)
    assert r1["converged"] and os.path.exists(f + "000000")
    print("[wavefunction] h2o-1em3 diag, load", flush=True)
    r2 = sbdx.diag(fcid, strs, eps=1e-9, do_rdm=False, loadname=f)
    assert r2["converged"]
    assert abs(r2["energy"] - r1["energy"]) < 1e-9
    assert r2["steps"] <= 3, (r1["steps"], r2["steps"])
