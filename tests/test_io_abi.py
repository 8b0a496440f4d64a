"""CPU tests of the host side: file formats, integral layouts, the instance
generator, and the C ABI surface of libsbdx.so (loaded, no GPU calls)."""
import os
import re

import numpy as np
import pytest

from conftest import DATA, ROOT


def test_fcidump_and_integrals_match_oracle(oracle_mod):
    from sbdx import io
    fd = io.read_fcidump(os.path.join(DATA, "n2_fcidump.txt.gz"))
    assert fd["norb"] == 18 and fd["nelec"] == 14 and fd["ms2"] == 0
    I0, I1, I2 = io.setup_integrals(fd)
    J0, J1, J2 = oracle_mod.setup_integrals(fd)
    assert I0 == J0 and np.array_equal(I1, J1) and np.array_equal(I2, J2)
    L = fd["norb"]
    eri = io.dense_eri(L, I2)
    # 8-fold symmetry of the expanded store
    assert np.array_equal(eri, eri.transpose(1, 0, 2, 3))
    assert np.array_equal(eri, eri.transpose(2, 3, 0, 1))
    h = io.spatial_h(L, I1)
    assert np.array_equal(h, h.T)
    # alpha and beta blocks of I1 are equal, cross-spin blocks are zero
    M = I1.reshape(2 * L, 2 * L)
    assert np.array_equal(M[0::2, 0::2], M[1::2, 1::2])
    assert not M[0::2, 1::2].any()


def test_fcidump_roundtrip(tmp_path):
    from sbdx import instances as I
    from sbdx import io
    inst = I.make_instance(norb=6, nelec_half=2, nstr=10, nchol=5, seed=9)
    p = tmp_path / "FCIDUMP"
    io.write_fcidump(p, 6, 4, 0, inst["h"], inst["eri"], 0.25)
    fd = io.read_fcidump(p)
    I0, I1, I2 = io.setup_integrals(fd)
    assert I0 == 0.25
    assert np.array_equal(I1, inst["I1"]) and np.array_equal(I2, inst["I2"])


def test_strings_roundtrip_and_order(tmp_path):
    from sbdx import io
    L = 36
    rng = np.random.default_rng(3)
    masks = set()
    while len(masks) < 50:
        masks.add(int(sum(1 << int(b) for b in rng.choice(L, 8, replace=False))))
    masks = list(masks)
    p = tmp_path / "a.txt"
    io.write_strings(p, list(reversed(masks)) + masks[:5], L)  # unsorted, duplicated
    s = io.read_strings(p, L)
    assert list(s) == sorted(masks)
    assert io.parse_string("0000011", 7) == 3
    assert io.format_string(3, 7) == "0000011"
    for bl in (20, 7, 64):
        w = io.masks_to_words(s, L, bl)
        assert w.shape[1] == (L + bl - 1) // bl
        assert np.array_equal(io.words_to_masks(w, L, bl), s)
        if bl < 64:
            assert np.all(w < np.uint64(1 << bl))


def test_reference_string_files_sorted():
    from sbdx import io
    s = io.read_strings(os.path.join(DATA, "h2o-1em3-alpha.txt"), 24)
    assert s.size == 275 and s[0] == (1 << 5) - 1  # HF first
    assert np.all(np.diff(s.astype(np.int64)) > 0)


def test_generator_properties():
    from sbdx import instances as I
    a = I.make_instance(norb=10, nelec_half=3, nstr=60, nchol=7, seed=4)
    b = I.make_instance(norb=10, nelec_half=3, nstr=60, nchol=7, seed=4)
    assert np.array_equal(a["I2"], b["I2"]) and np.array_equal(a["astr"], b["astr"])
    eri = a["eri"]
    K = 100
    M = eri.reshape(K, K)
    assert np.allclose(M, M.T) and np.linalg.eigvalsh(M).min() > -1e-10  # PSD
    assert np.allclose(eri, eri.transpose(1, 0, 2, 3))
    h = a["h"]
    assert np.allclose(h, h.T) and np.all(np.diag(h) <= 0) and np.all(np.diag(h) >= -2)
    s = a["astr"]
    assert s[0] == 0b111 and s.size == 60
    pc = [bin(int(x)).count("1") for x in s]
    assert set(pc) == {3}
    # HF, all singles (3*7 = 21), then doubles
    from itertools import combinations
    singles = {0b111 ^ (1 << i) ^ (1 << a) for i in range(3) for a in range(3, 10)}
    assert singles <= set(int(x) for x in s)


def test_workload_sizes():
    from sbdx import instances as I
    s = I.make_strings(36, 8, 8000, 1234)
    assert s.size == 8000 and len(set(s.tolist())) == 8000


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "sbdx.h")).read()
    return sorted(set(re.findall(r"\b(sbdx_[a-z0-9_]+)\s*\(", txt)))


def test_abi_exports_every_declared_symbol():
    import sbdx
    names = _header_functions()
    assert names, "no declarations parsed"
    assert sorted(sbdx.ABI_SYMBOLS) == names
    L = sbdx.lib()
    for n in names:
        assert hasattr(L, n), n
    assert L.sbdx_version() >= 100


def test_abi_rejects_bad_arguments():
    """Argument validation runs before any device call."""
    import ctypes as C

    import sbdx
    L = sbdx.lib()
    h = C.c_void_p()
    one = np.zeros(1, dtype=np.uint64)
    d = np.zeros(1)
    rc = L.sbdx_create(C.byref(h), 70, 20, one, 1, one, 1, 0.0, d, d, 1)
    assert rc == 3 and b"norb" in L.sbdx_last_error()
    rc = L.sbdx_create(C.byref(h), 4, 20, one, 1, one, 1, 0.0, d, d, 1)
    assert rc == 1 and b"n_I2" in L.sbdx_last_error()
    assert L.sbdx_dims(None, None, None, None) == 1
