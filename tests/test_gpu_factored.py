"""The mixed alpha-beta term as a dense (pq|rs) contraction on the matrix
cores (set_mixed_form 1: one DGEMM or SGEMM per batch of beta rows; 2: the
split-FP16x3 GEMM for FP32 products) against the oracle: products on
distinct and equal string sets and the edge sets, one batch and several
batches with a ragged last one (SBDX_FACTORED_ROWS), block products, and a
mixed-precision solve in the factored form."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

from conftest import random_vector

pytestmark = pytest.mark.gpu


@contextmanager
def _env(key, val):
    old = os.environ.get(key)
    if val is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = val
    try:
        yield
    finally:
        if old is None:
            os.environ.pop(key, None)
        else:
            os.environ[key] = old


def _case(name):
    from test_gpu_edge import CASES as EDGE
    from test_gpu_parity import CASES
    return (CASES.get(name) or EDGE[name])()


@pytest.mark.parametrize("rows", [None, "7"])
@pytest.mark.parametrize("case", ["tiny_ab", "tiny_same", "h2o_1em3", "no_singles", "one_det", "norb64"])
def test_factored_sigma_matches_oracle(case, rows, oracle_mod):
    import sbdx
    inst = _case(case)
    print(f"[factored] {case} rows={rows}", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    v = random_vector(op.ndet, 23)
    r = orc.sigma(v)
    with _env("SBDX_FACTORED_ROWS", rows):
        op.set_mixed_form(1)
        s = op.sigma(v)
        s32 = op.sigma_fp32(v)
        op.set_mixed_form(2)
        s16 = op.sigma_fp32(v)
        s2 = op.sigma(v)  # FP64 products stay on the DGEMM in form 2
        if np.array_equal(inst["astr"], inst["bstr"]):
            # the symmetric entry point runs the general product in this form
            from test_gpu_symmetry import symmetric_part
            w = symmetric_part(inst, v)
            rw = orc.sigma(w)
            op.set_mixed_form(1)
            assert np.max(np.abs(op.sigma_sym(w) - rw)) < 1e-12 * np.max(np.abs(rw))
    op.set_mixed_form(0)
    s0 = op.sigma(v)
    scale = np.max(np.abs(r))
    assert np.max(np.abs(s - r)) < 1e-12 * scale
    assert np.max(np.abs(s2 - r)) < 1e-12 * scale
    assert np.max(np.abs(s - s0)) < 1e-12 * scale
    nr = np.linalg.norm(r)
    assert np.linalg.norm(s32 - r) / nr < 2e-6
    assert np.linalg.norm(s16 - r) / nr < 1e-5
    op.close()
    orc.close()


def test_factored_block_products(oracle_mod):
    import sbdx
    inst = _case("h2o_1em3")
    print("[factored] block products", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    V = np.stack([random_vector(op.ndet, 40 + k) for k in range(3)])
    op.set_mixed_form(1)
    S = op.sigma_block(V)
    S32 = op.sigma_block(V, fp32=True)
    for k in range(3):
        r = orc.sigma(V[k])
        assert np.max(np.abs(S[k] - r)) < 1e-12 * np.max(np.abs(r))
        assert np.linalg.norm(S32[k] - r) / np.linalg.norm(r) < 2e-6
    op.close()
    orc.close()


# Mixed-precision thick solve with the split-FP16x3 factored products: the
# same energy as the oracle's FP64 Davidson.
def test_factored_mixed_solve(oracle_mod):
    import sbdx
    inst = _case("tiny_same")
    print("[factored] mixed solve", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    W0 = sbdx.basis_init_vector(op.ndet)
    ref = orc.davidson(W=W0, nb=12, eps=1e-10, max_it=400)
    assert ref["converged"]
    op.configure("mixed")
    op.set_mixed_form(2)
    op.set_vector(W0)
    E, steps, conv = op.solve(max_it=400, nb=12, eps=1e-9)
    print(f"[factored] products {steps}, dE {E - ref['E']:.2e}", flush=True)
    assert conv
    assert abs(E - ref["E"]) < 1e-9
    W = op.get_vector()
    assert np.linalg.norm(orc.sigma(W) - E * W) < 1e-8
    op.close()
    orc.close()
