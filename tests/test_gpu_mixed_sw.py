"""The signed-triangle mixed kernel (k_mixed_sw, staging width 16).

Its staged ERI rows hold the r >= s triangle twice, with both signs, and a
beta single's sign is folded into the address of its ERI value.  Pinned with
SBDX_MIXED_EB=16 (the autotuner then times only this kernel): the general
FP64 product against the oracle (1e-12), the FP32 product against the FP64
one (2e-6), and the spin-flip symmetric products (even and odd sectors)
against the general product, on distinct and equal string sets and the
water instance.  The tuned width must read 16, so the new kernel is the one
that ran."""
import numpy as np
import pytest

from conftest import random_vector
from test_gpu_symmetry import _env, antisymmetric_vector, basis_sign, symmetric_vector

pytestmark = pytest.mark.gpu


def _cases():
    from test_gpu_parity import CASES
    return CASES


@pytest.mark.parametrize("case", ["tiny_ab", "tiny_same", "h2o_1em3"])
def test_general_product_matches_oracle(case, oracle_mod):
    import sbdx
    inst = _cases()[case]()
    print(f"[mixed_sw] general {case}", flush=True)
    orc = oracle_mod.Oracle.from_instance(inst)
    hii = orc.diag()
    with _env("SBDX_MIXED_EB", "16"):
        op = sbdx.TPBOperator.from_instance(inst)
        v = random_vector(op.ndet, 3)
        s = op.sigma(v)
        s32 = op.sigma_fp32(v)
        forms = op.same_spin_forms()
    ref = orc.sigma(v, hii)
    orc.close()
    assert forms["fp64"]["mixed_width"] == 16 and forms["fp32"]["mixed_width"] == 16, forms
    assert np.max(np.abs(s - ref)) < 1e-12 * np.max(np.abs(ref))
    assert np.linalg.norm(s32 - s) / np.linalg.norm(s) < 2e-6
    op.close()


@pytest.mark.parametrize("case", ["tiny_same", "h2o_1em3"])
def test_symmetric_products_match_general(case):
    import sbdx
    inst = _cases()[case]()
    n = len(inst["astr"])
    S = basis_sign(inst["astr"], inst["bstr"], inst["norb"])
    print(f"[mixed_sw] symmetric {case}", flush=True)
    with _env("SBDX_MIXED_EB", "16"):
        op = sbdx.TPBOperator.from_instance(inst)
        out = {}
        for name, v in (("even", symmetric_vector(inst, 11)), ("odd", antisymmetric_vector(inst, 12))):
            out[name] = (op.sigma(v), op.sigma_sym(v), op.sigma_sym(v, fp32=True))
        forms = op.same_spin_forms()
    assert forms["fp64"]["sym_mixed_width"] == 16 and forms["fp32"]["sym_mixed_width"] == 16, forms
    for name, (s, ss, s32) in out.items():
        assert np.max(np.abs(ss - s)) < 1e-12 * np.max(np.abs(s)), name
        assert np.linalg.norm(s32 - s) / np.linalg.norm(s) < 2e-6, name
        sign = 1.0 if name == "even" else -1.0
        for y in (ss, s32):
            X = S * y.reshape(n, n)
            assert np.array_equal(X, sign * X.T), name
    op.close()


def test_solve_with_signed_triangle_kernel(oracle_mod):
    """A mixed-precision symmetric solve with only the new kernel reaches the
    energy of the default configuration's solve (1e-9)."""
    import sbdx
    inst = _cases()["h2o_1em3"]()
    print("[mixed_sw] h2o-1em3 solve", flush=True)
    energies = {}
    for eb in ("16", "4"):
        with _env("SBDX_MIXED_EB", eb):
            op = sbdx.TPBOperator.from_instance(inst)
            op.configure("mixed")
            op.set_symmetry("even")
            op.set_vector(sbdx.basis_init_vector(op.ndet))
            E, n, ok = op.solve(max_it=200, nb=12, eps=1e-8)
            assert ok, (eb, E, n)
            energies[eb] = E
            op.close()
    assert abs(energies["16"] - energies["4"]) < 1e-9, energies
