"""hipGraph replay of the FP32 symmetric sigma core: every replay, on
different input vectors and with other eager work between replays, must
give the eager result; a solve with graphs on must reach the oracle's
energy."""
import numpy as np
import pytest

from conftest import molecule

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("split", ["0", "1"])
def test_graph_replays_match_eager(split):
    import sbdx
    from test_gpu_symmetry import _env, symmetric_vector
    inst = molecule("h2o", "1em3")
    with _env("SBDX_SPLIT_GEMM", split):
        op = sbdx.TPBOperator.from_instance(inst)
        op.sigma_sym(symmetric_vector(inst, 19), fp32=True)  # settles the GEMM form
    vs = [symmetric_vector(inst, 20 + k) for k in range(5)]
    eager = [op.sigma_sym(v, fp32=True) for v in vs]
    op.set_graphs(True)
    rng = np.random.default_rng(4)
    print(f"[graph] h2o replays split={split}", flush=True)
    for k, v in enumerate(vs):
        op.sigma(rng.standard_normal(op.ndet))  # unrelated eager work between replays
        g = op.sigma_sym(v, fp32=True)
        assert np.max(np.abs(g - eager[k])) <= 1e-6 * np.max(np.abs(eager[k])), k
    for k, v in enumerate(vs):  # replay again, same order
        g = op.sigma_sym(v, fp32=True)
        assert np.max(np.abs(g - eager[k])) <= 1e-6 * np.max(np.abs(eager[k])), k
    op.close()


def test_graph_solve_matches_oracle(oracle_mod):
    import sbdx
    inst = molecule("h2o", "1em3")
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst)
    W0 = sbdx.basis_init_vector(op.ndet)
    ref = orc.davidson(W=W0, nb=12, eps=1e-9, max_it=300)
    op.configure("mixed")
    op.set_symmetry(True)
    op.set_graphs(True)
    print("[graph] h2o mixed symmetric solve", flush=True)
    op.set_vector(W0)
    E, steps, conv = op.solve(max_it=300, nb=12, eps=1e-9)
    assert conv and abs(E - ref["E"]) < 1e-9
    op.close()
    orc.close()
