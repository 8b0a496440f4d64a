"""RDM text outputs and the energy decomposition of the reference driver
(apps/chemistry_tpb_selected_basis_diagonalization/main.cc:264-318), in the
Python mirror (sbdx.io) and the C++ header (include/sbdx_io.hpp)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import bruteforce as B
from conftest import ROOT
from test_oracle_pins import tiny

INC = os.path.join(ROOT, "include")


def _fake_rdms(L):
    # every value encodes its own block and indices, so a swapped index or a
    # missing block changes the file
    one = np.zeros((2, L * L))
    two = np.zeros((4, L ** 4))
    for s in range(2):
        one[s] = np.arange(L * L) * 1.0 + 0.25 * s
    for s in range(4):
        two[s] = np.arange(L ** 4) * 1.0 + 1e-3 * (s + 1)
    return one, two


def test_rdm_text_layout(tmp_path):
    from sbdx import io
    L = 3
    one, two = _fake_rdms(L)
    io.write_rdm_text(str(tmp_path), L, one, two)
    lines1 = open(tmp_path / "1pRDM.txt").read().splitlines()
    lines2 = open(tmp_path / "2pRDM.txt").read().splitlines()
    assert len(lines1) == L * L and len(lines2) == L ** 4
    # outer index first, value one[0][i+L*j] + one[1][i+L*j]
    assert lines1[1].split()[:2] == ["0", "1"]
    assert float(lines1[1].split()[2]) == (0 + L * 1) * 2 + 0.25
    # "i j k l" with the value at i + L*j + L^2*(k + L*l), four blocks summed
    i, j, k, l, v = lines2[5].split()
    assert (i, j, k, l) == ("0", "0", "1", "2")
    x = 0 + L * 0 + L * L * (1 + L * 2)
    assert abs(float(v) - (4 * x + 1e-3 * 10)) < 1e-9
    g1, g2 = io.read_rdm_text(str(tmp_path), L)
    assert np.array_equal(g1, one[0] + one[1])
    assert np.allclose(g2, two.sum(axis=0), rtol=1e-15, atol=0)


@pytest.mark.parametrize("seed", [11, 4])
def test_energy_decomposition_reproduces_ground_energy(oracle_mod, seed):
    from sbdx import io
    L = 5
    inst = tiny(L, 2, 8, 7, seed)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    Hb = B.hamiltonian(L, inst["h"], inst["eri"], inst["I0"], inst["astr"], inst["bstr"])
    w, V = np.linalg.eigh(Hb)
    one, two = orc.rdm(V[:, 0])
    e = io.energy_decomposition(inst["I0"], inst["h"], inst["eri"], one, two)
    assert abs(e["total"] - w[0]) < 1e-12
    assert e["zero"] == inst["I0"]
    orc.close()


def test_cpp_writer_matches_python(tmp_path, oracle_mod):
    """The C++ writer and decomposition (sbdx_io.hpp, reading the integrals
    from an FCIDUMP) against the Python ones on the same RDMs: identical
    files, equal energies."""
    from sbdx import io
    gxx = shutil.which("g++")
    assert gxx, "g++ missing"
    L = 5
    inst = tiny(L, 2, 8, 7, 11)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    Hb = B.hamiltonian(L, inst["h"], inst["eri"], inst["I0"], inst["astr"], inst["bstr"])
    w, V = np.linalg.eigh(Hb)
    one, two = orc.rdm(V[:, 0])
    orc.close()
    one = np.asarray(one).reshape(2, -1)
    two = np.asarray(two).reshape(4, -1)
    fcf = tmp_path / "FCIDUMP"
    io.write_fcidump(str(fcf), L, 4, 0, inst["h"], inst["eri"], inst["I0"])
    np.concatenate([one.ravel(), two.ravel()]).tofile(str(tmp_path / "rdm.bin"))
    src = tmp_path / "w.cpp"
    src.write_text('''
#include <cstdio>
#include <fstream>
#include "sbdx_io.hpp"
int main(int argc, char **argv) {
  const size_t L = 5;
  auto g = sbdx::setup_integrals(sbdx::load_fcidump(argv[1]));
  std::vector<std::vector<double>> one(2, std::vector<double>(L * L)), two(4, std::vector<double>(L * L * L * L));
  std::ifstream f(argv[2], std::ios::binary);
  for (auto &b : one) f.read(reinterpret_cast<char *>(b.data()), sizeof(double) * b.size());
  for (auto &b : two) f.read(reinterpret_cast<char *>(b.data()), sizeof(double) * b.size());
  sbdx::write_rdm_text(argv[3], L, one, two);
  const auto e = sbdx::energy_decomposition(g, one, two);
  std::printf("%.17g %.17g %.17g\\n", e.zero, e.one, e.two);
  return 0;
}
''')
    exe = tmp_path / "w"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", INC, str(src), "-o", str(exe)], check=True)
    (tmp_path / "cpp").mkdir()
    (tmp_path / "py").mkdir()
    out = subprocess.run([str(exe), str(fcf), str(tmp_path / "rdm.bin"), str(tmp_path / "cpp")],
                         check=True, capture_output=True, text=True).stdout.split()
    io.write_rdm_text(str(tmp_path / "py"), L, one, two)
    for name in ("1pRDM.txt", "2pRDM.txt"):
        assert (tmp_path / "cpp" / name).read_bytes() == (tmp_path / "py" / name).read_bytes(), name
    e = io.energy_decomposition(inst["I0"], inst["h"], inst["eri"], one, two)
    zero, e1, e2 = (float(x) for x in out)
    assert abs(zero - e["zero"]) < 1e-12
    assert abs(e1 - e["one"]) < 1e-10 and abs(e2 - e["two"]) < 1e-10
    assert abs(zero + e1 + e2 - w[0]) < 1e-10


@pytest.mark.gpu
def test_diag_writes_rdm_text(tmp_path):
    """The driver path on the GPU: FCIDUMP + strings -> energy and RDMs; the
    text files' spin-summed 1-RDM has trace N and the decomposition of the
    device RDMs reproduces <W|HW>."""
    import sbdx
    from conftest import DATA
    print("[rdm text] h2o-1em3 diag", flush=True)
    out = sbdx.diag(os.path.join(DATA, "h2o_fcidump.txt.gz"), os.path.join(DATA, "h2o-1em3-alpha.txt"),
                    eps=1e-9, rdm_dir=str(tmp_path))
    assert out["converged"]
    assert abs(out["energy_parts"]["total"] - out["energy"]) < 1e-8
    g1, _ = sbdx.io.read_rdm_text(str(tmp_path), out["norb"])
    L = out["norb"]
    assert abs(sum(g1[i + L * i] for i in range(L)) - 10) < 1e-8  # h2o: 10 electrons


# ---- against the reference driver's own files (tests/golden/rdm_text_h2o.npz,
# made by tests/golden/make_rdm_text.py from the reference app built from its
# sources, apps/chemistry_tpb_selected_basis_diagonalization/main.cc:264-318,
# on the h2o-1em3 data set)
def _ref_rdm_text():
    from conftest import GOLDEN
    d = np.load(os.path.join(GOLDEN, "rdm_text_h2o.npz"))
    lines1 = d["lines1"].tobytes()
    lines2 = d["lines2"].tobytes().split(b"\n")[:-1]
    return d, lines1, dict(zip(d["idx2"].tolist(), lines2))


def test_rdm_text_bytes_match_reference(tmp_path):
    """The writer, fed the reference's printed values, reproduces the
    reference's files byte for byte: 1pRDM.txt whole (its SHA-256), and
    every sampled line of 2pRDM.txt (index order, separators, 16
    significant digits, exponent form) with the same line count."""
    import hashlib
    from sbdx import io
    d, lines1, sample2 = _ref_rdm_text()
    L = 24
    one = np.zeros((2, L * L))
    two = np.zeros((4, L ** 4))
    for ln in lines1.decode().splitlines():
        i, j, v = ln.split()
        one[0][int(i) + L * int(j)] = float(v)
    for ln in sample2.values():
        i, j, k, l, v = ln.decode().split()
        two[0][int(i) + L * int(j) + L * L * (int(k) + L * int(l))] = float(v)
    io.write_rdm_text(str(tmp_path), L, one, two)
    b1 = open(tmp_path / "1pRDM.txt", "rb").read()
    assert b1 == lines1
    assert hashlib.sha256(b1).hexdigest().encode() == d["sha1p"].tobytes()
    out2 = open(tmp_path / "2pRDM.txt", "rb").read().split(b"\n")[:-1]
    assert len(out2) == int(d["nlines2"][0])
    bad = [n for n, ln in sample2.items() if out2[n] != ln]
    assert not bad, (bad[:5], [out2[n] for n in bad[:3]], [sample2[n] for n in bad[:3]])


def test_rdm_of_oracle_ground_state_matches_reference(oracle_mod):
    """The oracle's ground state of h2o-1em3 (Davidson to 1e-10, as the
    reference app was run) gives, through the RDM and the energy
    decomposition of this build, the reference driver's printed energies
    (1e-9) and its 1pRDM.txt / sampled 2pRDM.txt values (1e-8)."""
    import sbdx
    from conftest import molecule
    from sbdx import io
    d, lines1, sample2 = _ref_rdm_text()
    inst = molecule("h2o", "1em3")
    L = inst["norb"]
    orc = oracle_mod.Oracle.from_instance(inst)
    res = orc.davidson(W=sbdx.basis_init_vector(len(inst["astr"]) ** 2), nb=12, eps=1e-10, max_it=300)
    assert res["converged"]
    one, two = orc.rdm(res["W"])
    orc.close()
    one = np.asarray(one).reshape(2, -1)
    two = np.asarray(two).reshape(4, -1)
    E = io.energy_decomposition(inst["I0"], io.spatial_h(L, inst["I1"]), io.dense_eri(L, inst["I2"]), one, two)
    Etot, E0, E1, E2 = d["energies"]
    assert abs(res["E"] - Etot) < 1e-9
    assert E["zero"] == E0
    assert abs(E["one"] - E1) < 1e-8 and abs(E["two"] - E2) < 1e-8
    g1 = one[0] + one[1]
    for ln in lines1.decode().splitlines():
        i, j, v = ln.split()
        assert abs(g1[int(i) + L * int(j)] - float(v)) < 1e-8
    g2 = ((two[0] + two[1]) + two[2]) + two[3]
    err = max(abs(g2[int(i) + L * int(j) + L * L * (int(k) + L * int(l))] - float(v))
              for i, j, k, l, v in (ln.decode().split() for ln in sample2.values()))
    assert err < 1e-8
