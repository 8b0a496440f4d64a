"""Carry-over string selection (chemistry/tpb/rdmat.h, extend.h as applied by
sbdiag.h:692-735) in the Python mirror (sbdx.carryover) and the C++ header
(include/sbdx_io.hpp), against the strings the reference app wrote for the
h2o-1em3 data set (tests/golden/carryover_h2o.npz, made by
tests/golden/make_carryover.py from the reference built from its sources).
The selection is evaluated on the oracle's ground state (Davidson to 1e-10,
as the reference app was run), so string weights agree with the
reference's to ~1e-10: strings of equal weight may come in another order
(the reference sorts with std::sort), which the weight-sequence comparison
allows for."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, molecule

L = 24


@pytest.fixture(scope="module")
def h2o_ground(oracle_mod):
    import sbdx
    inst = molecule("h2o", "1em3")
    orc = oracle_mod.Oracle.from_instance(inst)
    res = orc.davidson(W=sbdx.basis_init_vector(len(inst["astr"]) ** 2), nb=12, eps=1e-10, max_it=300)
    orc.close()
    assert res["converged"]
    return inst, np.asarray(res["W"])


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "carryover_h2o.npz"))


def _same_by_weight(ours, theirs, strings, D, tol=1e-9):
    # same length, the same weight at every position, and the same strings
    # wherever a weight is not shared (within tol) by another string
    assert ours.size == theirs.size
    pos = {int(s): i for i, s in enumerate(strings)}
    wo = np.array([D[pos[int(s)]] for s in ours])
    wt = np.array([D[pos[int(s)]] for s in theirs])
    assert np.max(np.abs(wo - wt)) < tol
    lone = [i for i in range(wt.size) if np.sum(np.abs(D - wt[i]) < tol) == 1]
    assert all(int(ours[i]) == int(theirs[i]) for i in lone)


@pytest.mark.parametrize("case", ["t1_thr", "t1_ratio"])
def test_selection_matches_reference(h2o_ground, ref, case):
    from sbdx import carryover as co
    inst, W = h2o_ground
    a, b = inst["astr"], inst["bstr"]
    Da, Db = co.string_weights(W, a.size, b.size)
    if case == "t1_thr":
        ca, cb, info = co.carryover(W, a, b, L, 1, threshold=1e-3)
    else:
        ca, cb, info = co.carryover(W, a, b, L, 1, ratio=0.2)
        assert abs(info["truncated_alpha"] - ref[f"{case}_w"][0]) < 1e-9
        assert abs(info["truncated_beta"] - ref[f"{case}_w"][1]) < 1e-9
    _same_by_weight(ca, ref[f"{case}_a"], a, Da)
    _same_by_weight(cb, ref[f"{case}_b"], b, Db)


@pytest.mark.parametrize("case", ["t2_thr", "t3_thr"])
def test_singles_extension_matches_reference(h2o_ground, ref, case):
    from sbdx import carryover as co
    inst, W = h2o_ground
    a, b = inst["astr"], inst["bstr"]
    if case == "t2_thr":
        ca, cb, _ = co.carryover(W, a, b, L, 2, threshold=1e-2)
    else:
        ca, cb, info = co.carryover(W, a, b, L, 3, threshold=1e-6)
        assert abs(info["total_weight"] - ref[f"{case}_w"][0]) < 1e-9
    assert np.array_equal(ca, ref[f"{case}_a"]) and np.array_equal(cb, ref[f"{case}_b"])


def test_singles_closed_form():
    """n(norb - n) singles per string, each one bit moved; the extension of
    one string is it plus its singles."""
    from sbdx import carryover as co
    s = 0b1011001
    ex = co.singles(s, 9)
    assert len(ex) == 4 * 5 and len(set(ex)) == 20
    assert all(bin(e).count("1") == 4 and bin(e ^ s).count("1") == 2 for e in ex)
    assert np.array_equal(co.singles_extend([s], 9), np.unique(np.array([s] + ex, dtype=np.uint64)))


def test_cpp_carryover_matches_python(tmp_path, h2o_ground):
    """include/sbdx_io.hpp's carry-over functions give the Python mirror's
    strings, in the same order, and the same weights."""
    from sbdx import carryover as co
    inst, W = h2o_ground
    a = np.asarray(inst["astr"], dtype=np.uint64)
    gxx = shutil.which("g++")
    assert gxx, "g++ missing"
    W.astype(np.float64).tofile(tmp_path / "w.bin")
    a.tofile(tmp_path / "a.bin")
    src = tmp_path / "c.cpp"
    src.write_text("""
#include <cstdio>
#include <fstream>
#include "sbdx_io.hpp"
int main(int argc, char **argv) {
  const size_t n = std::stoul(argv[1]);
  std::vector<double> W(n * n);
  std::vector<uint64_t> a(n);
  std::ifstream(argv[2], std::ios::binary).read((char *)W.data(), sizeof(double) * n * n);
  std::ifstream(argv[3], std::ios::binary).read((char *)a.data(), sizeof(uint64_t) * n);
  auto dump = [](const char *tag, const std::vector<uint64_t> &v) {
    std::printf("%s %zu", tag, v.size());
    for (uint64_t x : v) std::printf(" %llu", (unsigned long long)x);
    std::printf("\\n");
  };
  const struct { int type; double thr, ratio; } cases[] = {{1, 1e-3, 0.0}, {1, 0.0, 0.2}, {2, 1e-2, 0.0},
                                                           {3, 1e-6, 0.0}};
  for (const auto &cs : cases) {
    auto r = sbdx::carryover(W, a, a, 24, cs.type, cs.thr, cs.ratio);
    dump("a", r.adet);
    dump("b", r.bdet);
    std::printf("w %.17g %.17g %.17g\\n", r.truncated_alpha, r.truncated_beta, r.total_weight);
  }
}
""")
    exe = tmp_path / "c"
    subprocess.run([gxx, "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), str(a.size), str(tmp_path / "w.bin"), str(tmp_path / "a.bin")],
                         check=True, capture_output=True, text=True).stdout.splitlines()
    cases = [(1, 1e-3, 0.0), (1, 0.0, 0.2), (2, 1e-2, 0.0), (3, 1e-6, 0.0)]
    for k, (t, thr, ratio) in enumerate(cases):
        ca, cb, info = co.carryover(W, a, a, L, t, threshold=thr, ratio=ratio)
        la, lb, lw = out[3 * k].split(), out[3 * k + 1].split(), out[3 * k + 2].split()
        assert [int(x) for x in la[2:]] == [int(x) for x in ca]
        assert [int(x) for x in lb[2:]] == [int(x) for x in cb]
        w = [float(x) for x in lw[1:]]
        assert abs(w[0] - info.get("truncated_alpha", 0.0)) < 1e-15
        assert abs(w[2] - info.get("total_weight", 0.0)) < 1e-12


@pytest.mark.gpu
def test_diag_carryover_on_gpu(ref, tmp_path):
    """sbdx.diag on the GPU returns the reference's carry-over strings for
    the singles extension of the dominant determinants (type 3)."""
    import gzip
    import sbdx
    from conftest import DATA
    fcf = tmp_path / "fcidump.txt"
    with gzip.open(os.path.join(DATA, "h2o_fcidump.txt.gz"), "rt") as g:
        fcf.write_text(g.read())
    print("[carryover] h2o-1em3 diag type 3", flush=True)
    out = sbdx.diag(str(fcf), os.path.join(DATA, "h2o-1em3-alpha.txt"), eps=1e-10, do_rdm=False,
                    carryover_type=3, carryover_threshold=1e-6)
    assert out["converged"]
    assert np.array_equal(out["co_adet"], ref["t3_thr_a"]) and np.array_equal(out["co_bdet"], ref["t3_thr_b"])
    assert abs(out["carryover"]["total_weight"] - ref["t3_thr_w"][0]) < 1e-9
