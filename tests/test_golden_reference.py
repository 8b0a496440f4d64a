"""Checks against golden vectors written by the reference implementation
itself (tests/golden/make_golden.py runs the reference CPU path built from
its own sources).  The oracle is checked on the CPU, libsbdx on the GPU."""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, molecule

CASES = ["h2o_1em3", "n2_1em3", "n2_3em4", "syn10", "tiny16"]
# Cases whose Davidson result is compared, with (nb, max_it).  syn10's
# ground state has exactly zero Hartree-Fock weight (another spin-symmetry
# sector): from the HF start vector the reference reaches it only through
# rounding noise after ~250 steps, so its solver result is not a
# reproducible target; its tables, diagonal and sigma still are.  tiny16
# (norb 16, 5 electrons per spin, 500 x 500 strings) has its ground state in
# the odd spin-flip sector too, and the HF vector is a poor start for the even
# one (weight 3.5e-4): the reference's Davidson from it is chaotic -- the
# preconditioner 1/(E0 - h_ii) meets diagonal elements close to E0, and
# rounding differences change the trajectory (the oracle's departs from the
# reference's at step 8, the GPU's summation order at step 5).  Its first
# 4 steps, as the reference prints them (6 significant digits), are the
# fixture; its ground state is pinned
# by tests/golden/sectors_tiny16.json instead.
SOLVE = {"h2o_1em3": (10, 100), "n2_1em3": (10, 100), "n2_3em4": (10, 100), "tiny16": (12, 1)}
TRAJ_PREFIX = {"tiny16": 4}


def load(name):
    return dict(np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False))


def instance(name):
    from sbdx import instances
    if name == "syn10":
        return instances.make_instance(norb=10, nelec_half=3, nstr=80, nchol=10, seed=77)
    if name == "tiny16":
        return instances.workload("tiny16")
    mol, sub = name.split("_")
    return molecule(mol, sub)


def seeded_vector(g):
    v = np.random.default_rng(int(g["seed"])).standard_normal(int(g["ndet"]))
    return v / np.linalg.norm(v)


def table_sha(off, ket, cran):
    blob = np.concatenate([np.diff(off).astype(np.int64), ket.astype(np.int64), cran.astype(np.int64)])
    return np.frombuffer(hashlib.sha256(blob.tobytes()).digest(), dtype=np.uint8)


def check_all(g, tables, hii, sigma, davidson, rdm, tol_e=1e-10):
    for nm, (spin, kind) in {"sa": (0, 1), "da": (0, 2), "sb": (1, 1), "db": (1, 2)}.items():
        off, ket, cran = tables(spin, kind)
        assert int(g[f"table_{nm}_count"]) == ket.size, nm
        assert np.array_equal(table_sha(off, ket, cran), g[f"table_{nm}_sha"]), nm
    h = hii()
    assert np.max(np.abs(h[g["hii_idx"]] - g["hii_val"])) < 1e-12 * np.abs(g["hii_val"]).max()
    assert abs(h.sum() - g["hii_sum"]) < 1e-11 * abs(g["hii_sum"])
    v = seeded_vector(g)
    s = sigma(v)
    scale = np.abs(g["sigma_val"]).max()
    assert np.max(np.abs(s[g["sigma_idx"]] - g["sigma_val"])) < 1e-12 * scale
    assert abs(np.linalg.norm(s) - g["sigma_norm"]) < 1e-12 * g["sigma_norm"]
    assert abs(v @ s - g["sigma_vhv"]) < 1e-11 * max(1.0, abs(g["sigma_vhv"]))
    if davidson is None:
        return
    if "prefix" in g:
        trace = davidson()
        k = int(g["prefix"])
        ref = g["trajectory"][:k]
        # the reference prints E0 and |r| with 6 significant digits: half a
        # unit of the 6th digit, plus the float rounding of the parse
        for col in (0, 1):
            half = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(ref[:, col]))) - 5)
            assert np.all(np.abs(trace[:k, col] - ref[:, col]) <= half * (1 + 1e-9)), (col, trace[:k], ref)
        return
    E, W, steps = davidson()
    assert abs(E - float(g["energy"])) < tol_e, (E, float(g["energy"]))
    traj = g["trajectory"]
    assert abs(steps - traj.shape[0]) <= 2, (steps, traj.shape)
    assert abs((W * W).sum() - g["W_sq"]) < 1e-12
    assert np.max(np.abs(np.abs(W[g["W_idx"]]) - g["W_abs"])) < 1e-6
    if rdm is not None and "rdm1" in g:
        one, two = rdm(W)
        assert np.max(np.abs(one - g["rdm1"])) < 1e-9
        two = two.reshape(4, -1)
        assert np.max(np.abs(two[:, g["rdm2_idx"]] - g["rdm2_val"])) < 1e-9
        assert np.max(np.abs(two.sum(axis=1) - g["rdm2_sum"])) < 1e-8
        assert np.max(np.abs(np.linalg.norm(two, axis=1) - g["rdm2_norm"])) < 1e-8


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(oracle_mod, name):
    g = load(name)
    inst = instance(name)
    orc = oracle_mod.Oracle.from_instance(inst)
    hii = orc.diag()

    def dav():
        nb, max_it = SOLVE[name]
        r = orc.davidson(nb=nb, eps=1e-8, max_it=max_it, hii=hii)
        if name in TRAJ_PREFIX:
            return r["trace"]
        return r["E"], r["W"], r["steps"]

    def rdm(W):
        o1, o2 = orc.rdm(W)
        return o1.ravel(), o2.ravel()

    solve = name in ("h2o_1em3", "n2_1em3", "tiny16")  # keep the CPU suite short
    if name in TRAJ_PREFIX:
        g["prefix"] = TRAJ_PREFIX[name]
    check_all(g, orc.table, lambda: hii, lambda v: orc.sigma(v, hii), dav if solve else None,
              rdm if solve else None)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_reference(name):
    import sbdx
    g = load(name)
    inst = instance(name)
    print(f"[golden {name}] ndet={int(g['ndet'])}", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)

    def dav():
        nb, max_it = SOLVE[name]
        E, W, steps, conv = sbdx.davidson(op, nb=nb, eps=1e-8, max_it=max_it)
        if name in TRAJ_PREFIX:
            return op.trace()
        assert conv
        return op.energy(W), W, steps

    if name in TRAJ_PREFIX:
        g["prefix"] = TRAJ_PREFIX[name]

    check_all(g, op.table, op.make_diag, op.sigma, dav if name in SOLVE else None,
              op.correlation)
    op.close()
