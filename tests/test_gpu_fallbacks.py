"""The mixed-term kernel variants for rows that do not fit in LDS (row
parts, then the global-memory gather), forced on a small instance by
lowering the LDS budget with SBDX_LDS_LIMIT, against the oracle."""
import os

import numpy as np
import pytest

from conftest import molecule, random_vector

pytestmark = pytest.mark.gpu


# h2o-1em3: 275 strings, K = 576.  FP64 one-pair staging needs 6816 bytes:
# 6000 forces the two-part kernel, 4000 the global gather.
@pytest.mark.parametrize("limit", [6000, 4000])
def test_mixed_fallbacks_match_oracle(limit, oracle_mod):
    import sbdx
    m = molecule("h2o", "1em3")
    old = os.environ.get("SBDX_LDS_LIMIT")
    os.environ["SBDX_LDS_LIMIT"] = str(limit)
    print(f"[fallback] LDS limit {limit}", flush=True)
    try:
        op = sbdx.TPBOperator.from_instance(m)
    finally:
        if old is None:
            del os.environ["SBDX_LDS_LIMIT"]
        else:
            os.environ["SBDX_LDS_LIMIT"] = old
    orc = oracle_mod.Oracle.from_instance(m)
    v = random_vector(op.ndet, 31)
    s = op.sigma(v)
    r = orc.sigma(v)
    assert np.max(np.abs(s - r)) < 1e-12 * np.max(np.abs(r))
    s32 = op.sigma_fp32(v)
    assert np.linalg.norm(s32 - r) / np.linalg.norm(r) < 2e-6
    op.close()


# The same-spin terms as dense GEMMs or as sparse row kernels (normally the
# faster form is picked on first use), each against the oracle, on distinct
# alpha/beta string sets and on a molecule.
@pytest.mark.parametrize("form", ["sparse", "dense"])
@pytest.mark.parametrize("case", ["tiny_ab", "h2o_1em3"])
def test_same_spin_forms_match_oracle(form, case, oracle_mod):
    import sbdx
    from test_gpu_parity import CASES
    inst = CASES[case]()
    old = os.environ.get("SBDX_SAME_SPIN")
    os.environ["SBDX_SAME_SPIN"] = form
    print(f"[same-spin {form}] {case}", flush=True)
    try:
        op = sbdx.TPBOperator.from_instance(inst)
        v = random_vector(op.ndet, 17)
        s = op.sigma(v)
        s32 = op.sigma_fp32(v)
    finally:
        if old is None:
            del os.environ["SBDX_SAME_SPIN"]
        else:
            os.environ["SBDX_SAME_SPIN"] = old
    orc = oracle_mod.Oracle.from_instance(inst)
    r = orc.sigma(v)
    assert np.max(np.abs(s - r)) < 1e-12 * np.max(np.abs(r))
    assert np.linalg.norm(s32 - r) / np.linalg.norm(r) < 2e-6
    op.close()
    orc.close()


# The spin-flip symmetric product when one staged pair does not fit in LDS:
# the row-parts kernel over the same triangle of slots (lowered LDS budget on
# h2o-1em3, equal string sets), FP64 and FP32, against the general product.
def test_symmetric_parts_match_general():
    import sbdx
    from test_gpu_symmetry import _env, symmetric_vector
    m = molecule("h2o", "1em3")
    print("[fallback] symmetric row parts, LDS limit 6000", flush=True)
    with _env("SBDX_LDS_LIMIT", "6000"):
        op = sbdx.TPBOperator.from_instance(m)
    v = symmetric_vector(m, 12)
    s = op.sigma(v)
    ss = op.sigma_sym(v)
    assert np.max(np.abs(ss - s)) < 1e-12 * np.max(np.abs(s))
    s32 = op.sigma_sym(v, fp32=True)
    assert np.linalg.norm(s32 - s) / np.linalg.norm(s) < 2e-6
    op.close()
