"""Edge cases against the oracle: a string set with no single excitations
at all (every slice of the sliced-ELL table empty; only dressing and
spectator terms in the mixed kernels), a one-determinant space, and the
largest orbital count one 64-bit word holds."""
import numpy as np
import pytest

from conftest import random_vector

pytestmark = pytest.mark.gpu


def _no_singles():
    from sbdx import instances as I
    inst = I.make_instance(norb=16, nelec_half=2, nstr=10, nchol=8, seed=5)
    s = np.sort(np.array([(1 << (2 * k)) | (1 << (2 * k + 1)) for k in range(8)], dtype=np.uint64))
    inst["astr"], inst["bstr"] = s, s.copy()  # pairwise disjoint: every pair differs by a double
    return inst


def _one_det():
    inst = _no_singles()
    inst["astr"], inst["bstr"] = inst["astr"][:1].copy(), inst["astr"][:1].copy()
    return inst


def _norb64():
    from sbdx import instances as I
    return I.make_instance(norb=64, nelec_half=3, nstr=60, nchol=10, seed=1)


CASES = {"no_singles": _no_singles, "one_det": _one_det, "norb64": _norb64}


@pytest.mark.parametrize("case", list(CASES))
def test_edge_sigma_and_energy(case, oracle_mod):
    import sbdx
    inst = CASES[case]()
    print(f"[edge] {case}", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    v = random_vector(op.ndet, 3)
    r = orc.sigma(v)
    s = op.sigma(v)
    assert np.max(np.abs(s - r)) <= 1e-12 * np.max(np.abs(r))
    s32 = op.sigma_fp32(v)
    assert np.linalg.norm(s32 - r) <= 2e-6 * np.linalg.norm(r)
    ref = orc.davidson(nb=8, eps=1e-9, max_it=500)
    E, W, steps, conv = sbdx.davidson(op, max_it=500, nb=8, eps=1e-9)
    assert conv and abs(E - ref["E"]) < 1e-9
    op.close()
    orc.close()


def test_edge_mixed_solve_no_singles(oracle_mod):
    """The optimized solver (mixed precision, thick restart, symmetric form)
    with every slice of the table empty."""
    import sbdx
    inst = _no_singles()
    print("[edge] mixed no_singles", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    ref = orc.davidson(nb=8, eps=1e-9, max_it=500)
    op.configure("mixed")
    op.set_symmetry(True)
    op.set_vector(sbdx.basis_init_vector(op.ndet))
    E, steps, conv = op.solve(max_it=500, nb=8, eps=1e-9)
    assert conv and abs(E - ref["E"]) < 1e-9
    op.close()
    orc.close()


def test_edge_norb64_exchange_sectors(oracle_mod):
    """norb 64 with 3 electrons per spin: the ground state here is odd under
    the exchange of the alpha and beta strings (X = -X^T in the alpha-first
    basis); the oracle's Davidson from the HF vector reaches it only through
    rounding noise.  The general mixed solve from a random start reaches the
    oracle's ground energy; the symmetric mode, restricted to the even sector
    that holds the HF vector, finds that sector's lowest state, the same in
    mixed precision and in FP64, above the ground state; with both sectors
    the symmetric mode returns the ground state."""
    import sbdx
    inst = _norb64()
    print("[edge] norb64 exchange sectors", flush=True)
    op = sbdx.TPBOperator.from_instance(inst)
    orc = oracle_mod.Oracle.from_instance(inst, 16)
    ref = orc.davidson(nb=8, eps=1e-9, max_it=800)
    assert ref["converged"]
    op.configure("mixed")
    op.set_symmetry(False)
    op.set_vector(random_vector(op.ndet, 8))
    E, steps, conv = op.solve(max_it=800, nb=8, eps=1e-9)
    assert conv and abs(E - ref["E"]) < 1e-9
    e_sym = {}
    for solver in ("mixed", "thick"):
        op.configure(solver)
        op.set_symmetry(True)
        op.set_vector(sbdx.basis_init_vector(op.ndet))
        e_sym[solver], steps, conv = op.solve(max_it=800, nb=8, eps=1e-9)
        assert conv
    assert abs(e_sym["mixed"] - e_sym["thick"]) < 1e-9
    assert e_sym["mixed"] > ref["E"] + 1e-3
    # both sectors: the even solve from the HF vector, then the odd one from
    # the antisymmetric pair on the lowest off-diagonal diagonal element; the
    # odd result is the ground state
    for solver in ("mixed", "thick"):
        op.configure(solver)
        op.set_symmetry("lowest")
        op.set_vector(sbdx.basis_init_vector(op.ndet))
        E, steps, conv = op.solve(max_it=800, nb=8, eps=1e-9)
        st = op.sector_stats()
        assert conv and abs(E - ref["E"]) < 1e-9, (solver, E, ref["E"], st)
        assert st["chosen"] == 2 and abs(st["energies"][0] - e_sym["mixed"]) < 1e-9
    op.close()
    orc.close()
