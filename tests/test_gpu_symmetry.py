"""Spin-flip symmetric sigma and solves (equal alpha/beta string sets).

A vector of the symmetric sector is W = S o W' with W' a symmetric
na x na matrix and S(Ia,Ib) = (-1)^#{p in Ia, r in Ib: r < p} the sign
between the reference's interleaved order and the alpha-first order.  The
symmetric-form product must equal the general product (FP64 to 1e-12, FP32
to 2e-6), and a mixed-precision thick-restart solve in symmetric form from
the Hartree-Fock vector must reach the oracle's Davidson energy."""
import os

import numpy as np
import pytest

from conftest import molecule

pytestmark = pytest.mark.gpu


def basis_sign(astr, bstr, norb):
    a = np.asarray(astr, dtype=np.uint64)
    b = np.asarray(bstr, dtype=np.uint64)
    cnt = np.zeros((a.size, b.size), dtype=np.int64)
    for p in range(norb):
        ap = ((a >> np.uint64(p)) & np.uint64(1)).astype(np.int64)
        below = b & np.uint64((1 << p) - 1)
        pc = np.array([bin(int(x)).count("1") for x in below], dtype=np.int64)
        cnt += np.outer(ap, pc)
    return np.where(cnt % 2 == 1, -1.0, 1.0)


def symmetric_part(inst, w):
    """The spin-flip symmetric part of a vector in the reference convention."""
    n = len(inst["astr"])
    S = basis_sign(inst["astr"], inst["bstr"], inst["norb"])
    X = (S * w.reshape(n, n))
    return (S * (X + X.T) / 2).reshape(-1)


def symmetric_vector(inst, seed):
    n = len(inst["astr"])
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n))
    X = X + X.T
    S = basis_sign(inst["astr"], inst["bstr"], inst["norb"])
    return (S * X).reshape(-1)


def antisymmetric_vector(inst, seed):
    """A random vector of the odd sector: W = S o X with X = -X^T."""
    n = len(inst["astr"])
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n))
    X = X - X.T
    S = basis_sign(inst["astr"], inst["bstr"], inst["norb"])
    return (S * X).reshape(-1)


def exchange(inst, w):
    """(P W)[a,b] = (-1)^(n_e^2 + popcount(a & b)) W[b,a], the exchange of the
    alpha and beta strings (n_e electrons each) in the reference's
    interleaved convention; the Hartree-Fock vector is even."""
    a = np.asarray(inst["astr"], dtype=np.uint64)
    n = a.size
    key = a.tobytes()
    if key not in _XSIGN:
        ne = bin(int(a[0])).count("1")
        pc = np.array([[bin(int(x) & int(y)).count("1") for y in a] for x in a]) + ne * ne
        _XSIGN[key] = np.where(pc % 2 == 0, 1.0, -1.0)
    return (_XSIGN[key] * w.reshape(n, n).T).reshape(-1)


_XSIGN = {}
_SECTORS = {}


def _cases():
    from test_gpu_parity import CASES
    return {"tiny_same": CASES["tiny_same"], "h2o_1em3": CASES["h2o_1em3"]}


class _env:
    """Set one environment variable for the duration of a with-block."""

    def __init__(self, name, value):
        self.name, self.value = name, value

    def __enter__(self):
        self.old = os.environ.get(self.name)
        os.environ[self.name] = self.value

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ[self.name]
        else:
            os.environ[self.name] = self.old


# split "0": FP32 SGEMM in the symmetric same-spin term; "1": the
# split-FP16x3 GEMM (FP16 hi/lo halves with power-of-two scales, three FP16
# MFMA GEMMs accumulating in FP32), each checked against the FP64 product.
@pytest.mark.parametrize("split", ["0", "1"])
@pytest.mark.parametrize("case", ["tiny_same", "h2o_1em3"])
def test_symmetric_sigma_matches_general(case, split):
    import sbdx
    inst = _cases()[case]()
    assert np.array_equal(inst["astr"], inst["bstr"])
    print(f"[symmetry] sigma {case} split={split}", flush=True)
    with _env("SBDX_SPLIT_GEMM", split):
        op = sbdx.TPBOperator.from_instance(inst)
        v = symmetric_vector(inst, 5)
        s = op.sigma(v)
        ss = op.sigma_sym(v)
        assert np.max(np.abs(ss - s)) < 1e-12 * np.max(np.abs(s))
        s32 = op.sigma_sym(v, fp32=True)
        # a vector with components over many decades: the scale and the
        # absolute floor of the split keep the small ones accurate too
        w = v * np.exp(-np.abs(np.random.default_rng(6).standard_normal(v.size)) * 8)
        w = symmetric_part(inst, w)
        w32 = op.sigma_sym(w, fp32=True)
        forms = op.same_spin_forms()["fp32"]
    assert forms["sym_gemm"] == ("split_fp16x3" if split == "1" else "sgemm")
    assert np.linalg.norm(s32 - s) / np.linalg.norm(s) < 2e-6
    sw = op.sigma(w)
    assert np.linalg.norm(w32 - sw) / np.linalg.norm(sw) < 2e-6
    op.close()


def test_symmetric_sigma_rejects_asymmetric_vector():
    import sbdx
    inst = _cases()["tiny_same"]()
    op = sbdx.TPBOperator.from_instance(inst)
    v = np.random.default_rng(1).standard_normal(op.ndet)
    with pytest.raises(RuntimeError):
        op.sigma_sym(v)
    op.close()


@pytest.mark.parametrize("precond,split", [("diag", "0"), ("diag", "1"), ("mlp", "1")])
def test_symmetric_mixed_solve_matches_oracle(precond, split, oracle_mod):
    import sbdx
    inst = molecule("h2o", "1em3")
    with _env("SBDX_SPLIT_GEMM", split):
        op = sbdx.TPBOperator.from_instance(inst)
        op.sigma_sym(sbdx.basis_init_vector(op.ndet), fp32=True)  # settles the GEMM form
    orc = oracle_mod.Oracle.from_instance(inst)
    W0 = sbdx.basis_init_vector(op.ndet)
    ref = orc.davidson(W=W0, nb=12, eps=1e-9, max_it=300)
    assert ref["converged"]
    op.configure("mixed")
    op.set_symmetry(True)
    if precond == "mlp":
        w = np.load(os.path.join(os.path.dirname(sbdx.__file__), "data", "precond_mlp.npy"))
        op.mlp_set_weights(w)
        op.set_preconditioner(1)
    print(f"[symmetry] h2o mixed solve, {precond}, split={split}", flush=True)
    op.set_vector(W0)
    E, steps, conv = op.solve(max_it=300, nb=12, eps=1e-9)
    assert conv
    assert abs(E - ref["E"]) < 1e-9
    W = op.get_vector()
    r = orc.sigma(W) - E * W
    assert np.linalg.norm(r) < 5e-9
    op.close()
    orc.close()


@pytest.mark.parametrize("split", ["0", "1"])
@pytest.mark.parametrize("case", ["tiny_same", "h2o_1em3"])
def test_odd_sector_sigma_matches_general(case, split):
    """Odd sector (X = -X^T): y = hii*x + (G - G^T) and the mixed triangle
    mirrored with a minus sign equal the general product; the result lies in
    the odd sector exactly (bitwise antisymmetric, zero diagonal)."""
    import sbdx
    inst = _cases()[case]()
    n = len(inst["astr"])
    S = basis_sign(inst["astr"], inst["bstr"], inst["norb"])
    print(f"[symmetry] odd sigma {case} split={split}", flush=True)
    with _env("SBDX_SPLIT_GEMM", split):
        op = sbdx.TPBOperator.from_instance(inst)
        v = antisymmetric_vector(inst, 7)
        assert np.allclose(exchange(inst, v), -v)
        s = op.sigma(v)
        ss = op.sigma_sym(v)
        s32 = op.sigma_sym(v, fp32=True)
    assert np.max(np.abs(ss - s)) < 1e-12 * np.max(np.abs(s))
    assert np.linalg.norm(s32 - s) / np.linalg.norm(s) < 2e-6
    for y in (ss, s32):
        X = S * y.reshape(n, n)
        assert np.array_equal(X, -X.T)
    op.close()


def _sector_energies(inst, oracle_mod):
    """Lowest eigenvalue of each spin-flip sector (and of H) by Lanczos
    (scipy eigsh) on the oracle's sigma, projected with (1 +- P)/2."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    orc = oracle_mod.Oracle.from_instance(inst)
    hii = orc.diag()
    N = orc.ndet
    out = {}
    for key, par in (("even", 1.0), ("odd", -1.0), ("ground", 0.0)):
        def mv(x, par=par):
            x = np.asarray(x).ravel()
            if par:
                x = 0.5 * (x + par * exchange(inst, x))
            y = orc.sigma(x, hii)
            return 0.5 * (y + par * exchange(inst, y)) if par else y
        op = LinearOperator((N, N), matvec=mv, dtype=np.float64)
        w = eigsh(op, k=1, which="SA", tol=1e-13, return_eigenvectors=False,
                  v0=np.random.default_rng(2).standard_normal(N))
        out[key] = float(w[0])
    orc.close()
    return out


@pytest.mark.parametrize("mode", ["even", "odd", "lowest"])
def test_sector_solves_match_dense(mode, oracle_mod):
    """Each sector's mixed-precision thick solve reaches that sector's lowest
    eigenvalue (Lanczos on the oracle's sigma, 1e-9), and 'lowest' the ground
    state (norb 64, 3 electrons per spin: the ground state is odd)."""
    import sbdx
    from sbdx import instances as I
    inst = I.make_instance(norb=64, nelec_half=3, nstr=60, nchol=10, seed=1)
    if "norb64" not in _SECTORS:
        _SECTORS["norb64"] = _sector_energies(inst, oracle_mod)
    ref = _SECTORS["norb64"]
    assert min(ref["even"], ref["odd"]) == pytest.approx(ref["ground"], abs=1e-9)
    op = sbdx.TPBOperator.from_instance(inst)
    op.configure("mixed")
    op.set_symmetry(mode)
    W0 = sbdx.basis_init_vector(op.ndet)
    if mode == "odd":
        W0 = antisymmetric_vector(inst, 3)
        W0 /= np.linalg.norm(W0)
    print(f"[symmetry] sector solve {mode}", flush=True)
    op.set_vector(W0)
    E, steps, conv = op.solve(max_it=300, nb=8, eps=1e-9)
    st = op.sector_stats()
    assert conv
    target = ref["ground"] if mode == "lowest" else ref[mode]
    assert abs(E - target) < 1e-9, (E, target, st)
    if mode == "lowest":
        assert st["chosen"] == (1 if ref["even"] <= ref["odd"] else 2)
        assert abs(st["energies"][0] - ref["even"]) < 1e-9
    op.close()
