"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_szp = np.ctypeslib.ndpointer(dtype=np.uintp, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_setup_integrals.argtypes = [C.c_int, C.c_size_t, _dp, _ip, C.POINTER(C.c_double), _dp, _dp]
        L.orc_i2_size.argtypes = [C.c_int]
        L.orc_i2_size.restype = C.c_size_t
        L.orc_create.argtypes = [C.c_int, C.c_double, _dp, _dp, _u64p, C.c_size_t, _u64p, C.c_size_t, C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_table.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_table.restype = C.c_size_t
        L.orc_diag.argtypes = [C.c_void_p, _dp]
        L.orc_sigma.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_sigma_rows.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_size_t, C.c_size_t]
        L.orc_sigma_rowlist.argtypes = [C.c_void_p, _dp, _dp, _dp,
                                        np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                        C.c_size_t]
        L.orc_davidson.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int), _dp, C.c_int]
        L.orc_davidson.restype = C.c_int
        L.orc_rdm.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_syev.argtypes = [C.c_int, _dp, C.c_int, _dp]
        L.orc_parity.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int]
        L.orc_parity.restype = C.c_double
        L.orc_hij_bra.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_hij_bra.restype = C.c_double
        _LIB = L
    return _LIB


def setup_integrals(fd):
    L = fd["norb"]
    I1 = np.zeros(4 * L * L)
    I2 = np.zeros(lib().orc_i2_size(L))
    I0 = C.c_double(0.0)
    vals = np.ascontiguousarray(fd["vals"], dtype=np.float64)
    idx = np.ascontiguousarray(fd["idx"], dtype=np.int32)
    lib().orc_setup_integrals(L, vals.size, vals, idx, C.byref(I0), I1, I2)
    return I0.value, I1, I2


class Oracle:
    """CPU restatement of the reference TPB operator on one rank."""

    def __init__(self, norb, I0, I1, I2, astr, bstr, nthreads=0):
        self.norb = norb
        self.astr = np.ascontiguousarray(astr, dtype=np.uint64)
        self.bstr = np.ascontiguousarray(bstr, dtype=np.uint64)
        self.na, self.nb = self.astr.size, self.bstr.size
        self.ndet = self.na * self.nb
        self._I1 = np.ascontiguousarray(I1, dtype=np.float64)
        self._I2 = np.ascontiguousarray(I2, dtype=np.float64)
        self.h = lib().orc_create(norb, float(I0), self._I1, self._I2, self.astr, self.na,
                                  self.bstr, self.nb, int(nthreads))
        if not self.h:
            raise RuntimeError("oracle: unsupported instance")

    @classmethod
    def from_instance(cls, inst, nthreads=0):
        return cls(inst["norb"], inst["I0"], inst["I1"], inst["I2"], inst["astr"], inst["bstr"],
                   nthreads)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def table(self, spin, kind):
        n = self.na if spin == 0 else self.nb
        cnt = lib().orc_table(self.h, spin, kind, None, None, None)
        off = np.zeros(n + 1, dtype=np.uintp)
        ket = np.zeros(max(cnt, 1), dtype=np.uint32)
        cran = np.zeros(max(cnt, 1) * (2 if kind == 1 else 4), dtype=np.int32)
        lib().orc_table(self.h, spin, kind, off.ctypes.data, ket.ctypes.data, cran.ctypes.data)
        return off.astype(np.int64), ket[:cnt], cran[: cnt * (2 if kind == 1 else 4)]

    def diag(self):
        out = np.zeros(self.ndet)
        lib().orc_diag(self.h, out)
        return out

    def sigma(self, C_vec, hii=None, out=None):
        if hii is None:
            hii = self.diag()
        C_vec = np.ascontiguousarray(C_vec, dtype=np.float64)
        if out is None:
            out = np.zeros(self.ndet)
        lib().orc_sigma(self.h, hii, C_vec, out)
        return out

    def sigma_rows(self, C_vec, hii, lo, hi, out):
        lib().orc_sigma_rows(self.h, hii, C_vec, out, lo, hi)
        return out

    def sigma_rowlist(self, C_vec, hii, rows, out):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        lib().orc_sigma_rowlist(self.h, hii, C_vec, out, rows, rows.size)
        return out

    def davidson(self, W=None, max_it=100, nb=12, eps=1e-8, hii=None, trace_cap=4096):
        if hii is None:
            hii = self.diag()
        if W is None:
            W = np.zeros(self.ndet)
            W[0] = 1.0
        W = np.ascontiguousarray(W, dtype=np.float64).copy()
        E = C.c_double(0.0)
        steps = C.c_int(0)
        trace = np.zeros(2 * trace_cap)
        rc = lib().orc_davidson(self.h, hii, W, max_it, nb, eps, C.byref(E), C.byref(steps),
                                trace, trace_cap)
        n = min(steps.value, trace_cap)
        return dict(E=E.value, W=W, steps=steps.value, converged=(rc == 0),
                    trace=trace[: 2 * n].reshape(n, 2))

    def rdm(self, W):
        L = self.norb
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        lib().orc_rdm(self.h, np.ascontiguousarray(W, dtype=np.float64), one, two)
        return one.reshape(2, L, L), two.reshape(4, L, L, L, L)

    def hij(self, a_bra, b_bra, a_ket, b_ket):
        return lib().orc_hij_bra(self.h, int(a_bra), int(b_bra), int(a_ket), int(b_ket))


def syev(A):
    n = A.shape[0]
    M = np.asfortranarray(A.copy()).reshape(-1, order="F").copy()
    w = np.zeros(n)
    lib().orc_syev(n, M, n, w)
    return w, M.reshape(n, n, order="F")
