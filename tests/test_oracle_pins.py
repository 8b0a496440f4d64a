"""Pins of the CPU oracle (CPU only).

* brute-force second quantization (tests/bruteforce.py) on tiny spaces:
  every H element, the Davidson energy, the 1-RDM, the same-spin 2-RDM
  blocks and the sum of the two mixed blocks;
* closed forms for the subspace eigensolver and the parity rule;
* invariants of the excitation tables;
* the published TPB energies of the reference's h2o and n2 data sets
  (data/h2o/README.md:6, data/n2/README.md:6);
* the alpha-first decomposition used by the HIP kernels.
Golden vectors written by the reference itself are checked in
test_golden_reference.py.
"""
import numpy as np
import pytest

import bruteforce as B
from conftest import molecule, random_vector


def tiny(norb=5, n=2, na=8, nb=7, seed=11, nchol=6):
    from sbdx import instances as I
    inst = I.make_instance(norb=norb, nelec_half=n, nstr=na, nchol=nchol, seed=seed)
    inst["bstr"] = I.make_strings(norb, n, nb, seed + 3)
    return inst


def dense_from_oracle(orc):
    hii = orc.diag()
    eye = np.eye(orc.ndet)
    return np.array([orc.sigma(eye[k], hii) for k in range(orc.ndet)]).T


@pytest.mark.parametrize("seed,norb,n,na,nb", [(11, 5, 2, 8, 7), (3, 6, 3, 12, 9), (8, 6, 2, 10, 15)])
def test_hamiltonian_matches_bruteforce(oracle_mod, seed, norb, n, na, nb):
    inst = tiny(norb, n, na, nb, seed)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    H = dense_from_oracle(orc)
    Hb = B.hamiltonian(norb, inst["h"], inst["eri"], inst["I0"], inst["astr"], inst["bstr"])
    assert np.max(np.abs(H - Hb)) < 1e-12 * np.max(np.abs(Hb))
    # single-element dispatch agrees too
    a, b = inst["astr"], inst["bstr"]
    for (i, j) in [(0, 1), (1, 0), (2, 3), (len(a) - 1, 0)]:
        for (k, l) in [(0, 0), (0, 1), (1, 2)]:
            bra, ket = i * len(b) + k, j * len(b) + l
            assert abs(orc.hij(a[i], b[k], a[j], b[l]) - Hb[bra, ket]) < 1e-12


def test_davidson_matches_eigh(oracle_mod):
    inst = tiny(6, 3, 14, 12, 21)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    Hb = B.hamiltonian(6, inst["h"], inst["eri"], inst["I0"], inst["astr"], inst["bstr"])
    w, V = np.linalg.eigh(Hb)
    r = orc.davidson(nb=6, eps=1e-10, max_it=200)
    assert r["converged"]
    assert abs(r["E"] - w[0]) < 1e-10
    assert abs(abs(r["W"] @ V[:, 0]) - 1.0) < 1e-8
    # trajectory energies are variational upper bounds, non-increasing within a restart
    assert np.all(r["trace"][:, 0] >= w[0] - 1e-10)


@pytest.mark.parametrize("seed", [11, 4])
def test_rdm_matches_bruteforce(oracle_mod, seed):
    L = 5
    inst = tiny(L, 2, 8, 7, seed)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    Hb = B.hamiltonian(L, inst["h"], inst["eri"], inst["I0"], inst["astr"], inst["bstr"])
    w, V = np.linalg.eigh(Hb)
    W = V[:, 0]
    one, two = orc.rdm(W)
    one = one.reshape(2, -1)
    two = two.reshape(4, -1)
    p1, p2 = B.rdms(L, inst["astr"], inst["bstr"], W)
    assert np.max(np.abs(one - p1)) < 1e-13
    assert np.max(np.abs(two[0] - p2[0])) < 1e-13
    assert np.max(np.abs(two[3] - p2[3])) < 1e-13
    # the mixed blocks split each element between blocks 1 and 2 differently
    # from <a+a+aa>, but their sum is the physical one
    assert np.max(np.abs((two[1] + two[2]) - (p2[1] + p2[2]))) < 1e-13
    # energy identity E = I0 + sum h*gamma + 1/2 sum (ik|jl) Gamma[ijkl]
    h, eri = inst["h"], inst["eri"]
    E1 = np.einsum("pq,spq->", h, one.reshape(2, L, L).transpose(0, 2, 1))
    G = two.reshape(4, L, L, L, L)  # [b, l, k, j, i]
    E2 = 0.5 * np.einsum("ikjl,blkji->", eri, G)
    assert abs(inst["I0"] + E1 + E2 - w[0]) < 1e-12
    # traces: sum_i gamma_ii = N_s ; sum_ij Gamma[s+2t][ijij] = N_s (N_t - delta_st)
    nel = 2
    for s in range(2):
        assert abs(np.trace(one[s].reshape(L, L)) - nel) < 1e-12
    for s in range(2):
        for t in range(2):
            tr = sum(G[s + 2 * t, j, i, j, i] for i in range(L) for j in range(L))
            assert abs(tr - nel * (nel - (1 if s == t else 0))) < 1e-12


def test_syev_closed_forms(oracle_mod):
    w, V = oracle_mod.syev(np.array([[2.0, 1.0], [1.0, 2.0]]))
    assert np.allclose(w, [1.0, 3.0], atol=1e-14)
    assert abs(abs(V[0, 0]) - 2 ** -0.5) < 1e-14
    w, _ = oracle_mod.syev(np.diag([3.0, -1.0, 2.0]))
    assert np.allclose(w, [-1.0, 2.0, 3.0], atol=0)
    rng = np.random.default_rng(0)
    A = rng.standard_normal((12, 12))
    A = A + A.T
    w, V = oracle_mod.syev(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
    assert np.allclose(A @ V, V * w, atol=1e-11)


def test_parity_rule(oracle_mod):
    rng = np.random.default_rng(1)
    for _ in range(200):
        d = int(rng.integers(0, 2 ** 40))
        s, e = sorted(rng.integers(0, 40, size=2))
        cnt = sum((d >> k) & 1 for k in range(s, e)) + ((d >> s) & 1)
        assert oracle_mod.lib().orc_parity(d, 0, int(s), int(e)) == (-1.0 if cnt % 2 else 1.0)


def test_table_invariants(oracle_mod):
    from sbdx import instances as I
    inst = I.make_instance(norb=12, nelec_half=4, nstr=200, nchol=8, seed=2)
    orc = oracle_mod.Oracle.from_instance(inst, 4)
    s = inst["astr"].astype(np.int64)
    n = s.size
    for kind, dist in ((1, 2), (2, 4)):
        off, ket, cran = orc.table(0, kind)
        pairs = set()
        for i in range(n):
            row = ket[off[i]:off[i + 1]]
            assert np.all(np.diff(row.astype(np.int64)) > 0)  # ascending
            brute = [j for j in range(n) if bin(int(s[i] ^ s[j])).count("1") == dist]
            assert list(row) == brute
            for j in row:
                pairs.add((i, int(j)))
        assert all((j, i) in pairs for (i, j) in pairs)  # symmetric
        # orbitals: cr in bra only, an in ket only, alpha spin-orbital labels
        w = 2 * kind
        c = cran.reshape(-1, w)
        assert np.all(c % 2 == 0)


PUBLISHED = [("h2o", "1em3", -76.23594663), ("n2", "1em3", -109.04162110)]


@pytest.mark.parametrize("mol,subset,energy", PUBLISHED)
def test_published_energy(oracle_mod, mol, subset, energy):
    m = molecule(mol, subset)
    orc = oracle_mod.Oracle(m["norb"], m["I0"], m["I1"], m["I2"], m["astr"], m["bstr"])
    r = orc.davidson(nb=10, eps=1e-8)
    assert r["converged"]
    assert abs(r["E"] - energy) < 6e-9, (r["E"], energy)


def test_alpha_first_model(oracle_mod):
    import alphafirst_model as M
    inst = tiny(8, 3, 25, 20, 5)
    orc = oracle_mod.Oracle.from_instance(inst, 2)
    hii = orc.diag()
    C = random_vector(orc.ndet, 0)
    s1 = orc.sigma(C, hii)
    s2 = M.sigma_alpha_first(inst["astr"], inst["bstr"], inst["h"], inst["eri"], hii, C)
    assert np.max(np.abs(s1 - s2)) < 1e-13 * np.max(np.abs(s1))
