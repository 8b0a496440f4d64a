"""sbdx_rdm_device_noalloc: the RDMs with device buffers and working memory
from the library's static scratch array, against sbdx_rdm (the entry point
that allocates), on two calls in a row, the second on a smaller instance,
and the size limit's error."""
import ctypes as C

import numpy as np
import pytest

from conftest import molecule, random_vector

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers for the test through the HIP runtime libsbdx itself
    links (the test allocates; the entry point under test must not)."""

    def __init__(self):
        self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        self.bufs.append(p)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(arr.nbytes)
        self._ok(self.rt.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1), "hipMemcpy")
        return p

    def download(self, p, n):
        out = np.empty(n)
        self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc_rdm(op, W):
    import sbdx
    L = op.norb
    hip = _Hip()
    try:
        dW = hip.upload(W)
        # garbage in the outputs: the entry point must overwrite every element
        d1 = hip.upload(np.full(2 * L * L, np.nan))
        d2 = hip.upload(np.full(4 * L ** 4, np.nan))
        rc = sbdx.lib().sbdx_rdm_device_noalloc(op.h, dW, d1, d2)
        msg = sbdx.lib().sbdx_last_error().decode() if rc else ""
        if rc:
            return rc, msg, None, None
        return rc, msg, hip.download(d1, 2 * L * L), hip.download(d2, 4 * L ** 4)
    finally:
        hip.free()


def _case(inst, seed):
    import sbdx
    op = sbdx.TPBOperator.from_instance(inst)
    W = random_vector(op.ndet, seed)
    one_ref, two_ref = op.correlation(W)
    rc, msg, one, two = _noalloc_rdm(op, W)
    assert rc == 0, msg
    for got, ref in ((one, one_ref), (two, two_ref)):
        assert np.all(np.isfinite(got))
        # same kernels, same summation per element except the order of the
        # floating-point atomics
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    op.close()


def test_noalloc_rdm_matches_rdm_twice():
    from sbdx import instances as I
    print("[no-alloc rdm] h2o-1em3 then a norb-10 instance", flush=True)
    _case(molecule("h2o", "1em3"), 11)
    _case(I.make_instance(norb=10, nelec_half=3, nstr=40, nchol=8, seed=4), 12)


def test_noalloc_rdm_rejects_oversize():
    """norb 64: the L^4 mixed-term accumulator alone exceeds the 128 MiB array."""
    import sbdx
    from sbdx import instances as I
    op = sbdx.TPBOperator.from_instance(I.make_instance(norb=64, nelec_half=3, nstr=20, nchol=6, seed=2))
    rc, msg, _, _ = _noalloc_rdm(op, random_vector(op.ndet, 1))
    assert rc == 1  # SBDX_ERR_ARG
    assert "128 MiB" in msg
    op.close()
