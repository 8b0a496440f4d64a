"""sbdx -- MI355X implementation of the sbd-pub TPB Hamiltonian path.

Python mirror of the reference's operator and solver interface, bound to
the C ABI of libsbdx.so (include/sbdx.h):

* ``TPBOperator``  ~ ``MultTPBThrust`` (chemistry/tpb/mult_thrust.h:23-95):
  ``run`` (MultBase::run, Wb += H Wk), ``make_diag`` (makeQChamDiagTerms),
  ``correlation`` (1-/2-RDM).
* ``davidson``     ~ ``sbd::Davidson(hii, W, mult, max_it, nb, eps, max_time)``
  (chemistry/basic/davidson_thrust.h:85-92).
* ``diag``         ~ ``sbd::tpb::diag`` (chemistry/tpb/sbdiag.h:139-151):
  energy, density, 1-/2-RDM from an FCIDUMP and string sets.

There is no CPU fallback: if libsbdx.so is missing or no GPU is visible, the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import io

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsbdx.so")
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

ERRORS = {1: "invalid argument", 2: "HIP failure", 3: "unsupported instance",
          4: "device allocation failed", 5: "not converged", 6: "I/O error"}

# every entry point of include/sbdx.h
ABI_SYMBOLS = [
    "sbdx_last_error", "sbdx_version", "sbdx_create", "sbdx_destroy", "sbdx_dims", "sbdx_diag",
    "sbdx_sigma", "sbdx_sigma_device", "sbdx_davidson", "sbdx_set_vector", "sbdx_solve",
    "sbdx_get_vector", "sbdx_trace", "sbdx_energy", "sbdx_rdm", "sbdx_table_count",
    "sbdx_table_export", "sbdx_profile", "sbdx_profile_count", "sbdx_profile_get",
    "sbdx_work_counts", "sbdx_set_precision", "sbdx_sigma_fp32", "sbdx_set_restart",
    "sbdx_mlp_nparam", "sbdx_mlp_nfeat", "sbdx_mlp_set_weights", "sbdx_mlp_get_weights",
    "sbdx_mlp_forward", "sbdx_mlp_gradient", "sbdx_mlp_train", "sbdx_mlp_record", "sbdx_mlp_collect",
    "sbdx_set_preconditioner", "sbdx_precond_stats", "sbdx_set_verbose",
    "sbdx_same_spin_form", "sbdx_sigma_block", "sbdx_davidson_block",
    "sbdx_mixed_staging", "sbdx_split_gemm", "sbdx_set_mixed_form", "sbdx_fp64_switch_step", "sbdx_sigma_block_fp32", "sbdx_symmetric_eig", "sbdx_set_symmetry", "sbdx_sector_stats",
    "sbdx_sigma_sym", "sbdx_mlp_apply", "sbdx_set_coarse_start", "sbdx_coarse_stats",
    "sbdx_set_graphs", "sbdx_rdm_device_noalloc", "sbdx_set_block_basis", "sbdx_block_stats",
    "sbdx_symmetric_geig",
]
PREC_FP64, PREC_MIXED = 0, 1


class SbdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sbdx error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def lib():
    """Load libsbdx.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    L.sbdx_last_error.restype = C.c_char_p
    L.sbdx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_size_t, _u64p, C.c_size_t, _u64p,
                              C.c_size_t, C.c_double, _dp, _dp, C.c_size_t]
    L.sbdx_destroy.argtypes = [C.c_void_p]
    L.sbdx_dims.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                            C.POINTER(C.c_size_t)]
    L.sbdx_diag.argtypes = [C.c_void_p, _dp]
    L.sbdx_sigma.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    L.sbdx_sigma_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    L.sbdx_davidson.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.sbdx_set_vector.argtypes = [C.c_void_p, _dp]
    L.sbdx_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.sbdx_get_vector.argtypes = [C.c_void_p, _dp]
    L.sbdx_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.sbdx_energy.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double)]
    L.sbdx_rdm.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.sbdx_table_count.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_size_t)]
    L.sbdx_table_export.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
    L.sbdx_profile.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_profile_count.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.sbdx_profile_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_size_t,
                                   C.POINTER(C.c_long), C.POINTER(C.c_double)]
    L.sbdx_work_counts.argtypes = [C.c_void_p, _dp]
    L.sbdx_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_sigma_fp32.argtypes = [C.c_void_p, _dp, _dp]
    L.sbdx_sigma_block.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
    L.sbdx_symmetric_eig.argtypes = [C.c_int, _dp, _dp, _dp]
    L.sbdx_symmetric_geig.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.sbdx_set_symmetry.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_rdm_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sbdx_sector_stats.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p, C.c_void_p]
    L.sbdx_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_set_coarse_start.argtypes = [C.c_void_p, C.c_size_t, C.c_double]
    L.sbdx_coarse_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.sbdx_mlp_apply.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
    L.sbdx_sigma_sym.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    L.sbdx_sigma_block_fp32.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.sbdx_davidson_block.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_double,
                                      C.c_double, _dp, C.POINTER(C.c_int)]
    L.sbdx_set_block_basis.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_block_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   _dp, C.c_void_p, C.c_int]
    L.sbdx_set_restart.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_set_verbose.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_same_spin_form.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.sbdx_fp64_switch_step.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.sbdx_set_mixed_form.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_split_gemm.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.sbdx_mixed_staging.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float), C.POINTER(C.c_int)]
    _u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
    L.sbdx_mlp_set_weights.argtypes = [C.c_void_p, _fp, C.c_size_t]
    L.sbdx_mlp_get_weights.argtypes = [C.c_void_p, _fp, C.c_size_t]
    L.sbdx_mlp_forward.argtypes = [C.c_void_p, _u16p, C.c_size_t, _fp]
    L.sbdx_mlp_gradient.argtypes = [C.c_void_p, _u16p, _fp, C.c_size_t, _fp, C.POINTER(C.c_double)]
    L.sbdx_mlp_train.argtypes = [C.c_void_p, _u16p, _fp, C.c_size_t, C.c_int, C.c_int, C.c_float,
                                 C.c_uint64, _dp]
    L.sbdx_mlp_record.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_mlp_collect.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64, _u16p, _fp,
                                   C.POINTER(C.c_size_t)]
    L.sbdx_set_preconditioner.argtypes = [C.c_void_p, C.c_int]
    L.sbdx_precond_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    _LIB = L
    return L


def _check(rc, allow=()):
    if rc != 0 and rc not in allow:
        raise SbdxError(rc, lib().sbdx_last_error().decode())
    return rc


class TPBOperator:
    """Device-resident TPB Hamiltonian of one (alpha strings x beta strings)
    space; builds the excitation tables and the diagonal on construction
    (MultTPBThrust::Init + MakeHelpers for one rank)."""

    def __init__(self, norb, I0, I1, I2, astr, bstr=None, bit_length=20):
        astr = np.ascontiguousarray(astr, dtype=np.uint64)
        bstr = astr if bstr is None else np.ascontiguousarray(bstr, dtype=np.uint64)
        self.norb = int(norb)
        self.na, self.nb = astr.size, bstr.size
        self.ndet = self.na * self.nb
        aw = np.ascontiguousarray(io.masks_to_words(astr, norb, bit_length).reshape(-1))
        bw = np.ascontiguousarray(io.masks_to_words(bstr, norb, bit_length).reshape(-1))
        I1 = np.ascontiguousarray(I1, dtype=np.float64).reshape(-1)
        I2 = np.ascontiguousarray(I2, dtype=np.float64).reshape(-1)
        h = C.c_void_p()
        _check(lib().sbdx_create(C.byref(h), self.norb, bit_length, aw, self.na, bw, self.nb,
                                 float(I0), I1, I2, I2.size))
        self.h = h

    @classmethod
    def from_instance(cls, inst, bit_length=20):
        return cls(inst["norb"], inst["I0"], inst["I1"], inst["I2"], inst["astr"], inst["bstr"],
                   bit_length)

    def close(self):
        if getattr(self, "h", None):
            lib().sbdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def make_diag(self):
        out = np.empty(self.ndet)
        _check(lib().sbdx_diag(self.h, out))
        return out

    def run(self, Wk, Wb=None, accumulate=True):
        """Wb += H*Wk (MultBase::run semantics); returns Wb."""
        Wk = np.ascontiguousarray(Wk, dtype=np.float64)
        if Wb is None:
            Wb = np.zeros(self.ndet)
            accumulate = False
        _check(lib().sbdx_sigma(self.h, Wk, Wb, 1 if accumulate else 0))
        return Wb

    def sigma(self, Wk):
        return self.run(Wk, None, accumulate=False)

    def sigma_fp32(self, Wk):
        """H*Wk in FP32 arithmetic (the mixed solver's correction product)."""
        out = np.zeros(self.ndet)
        _check(lib().sbdx_sigma_fp32(self.h, np.ascontiguousarray(Wk, dtype=np.float64), out))
        return out

    def set_precision(self, mode):
        """PREC_FP64 (reference algorithm) or PREC_MIXED for later solves."""
        _check(lib().sbdx_set_precision(self.h, int(mode)))

    # ---------------------------------------------------------- learned preconditioner
    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().sbdx_mlp_set_weights(self.h, w, w.size))

    def mlp_get_weights(self):
        w = np.zeros(lib().sbdx_mlp_nparam(), dtype=np.float32)
        _check(lib().sbdx_mlp_get_weights(self.h, w, w.size))
        return w

    def mlp_forward(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.uint16)
        n = feat.shape[0]
        out = np.zeros(n, dtype=np.float32)
        _check(lib().sbdx_mlp_forward(self.h, feat.reshape(-1), n, out))
        return out

    def mlp_gradient(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.uint16)
        target = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(lib().sbdx_mlp_nparam(), dtype=np.float32)
        sse = C.c_double()
        _check(lib().sbdx_mlp_gradient(self.h, feat.reshape(-1), target, target.size, g, C.byref(sse)))
        return g, sse.value

    def mlp_train(self, feat, target, epochs=10, batch=4096, lr=1e-3, seed=0):
        feat = np.ascontiguousarray(feat, dtype=np.uint16)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = np.zeros(epochs)
        _check(lib().sbdx_mlp_train(self.h, feat.reshape(-1), target, target.size, epochs, batch,
                                    lr, seed, loss))
        return loss

    def mlp_record(self, cap):
        _check(lib().sbdx_mlp_record(self.h, int(cap)))

    def mlp_collect(self, max_pairs, seed=0):
        nf = lib().sbdx_mlp_nfeat()
        feat = np.zeros(max_pairs * nf, dtype=np.uint16)
        tgt = np.zeros(max_pairs, dtype=np.float32)
        n = C.c_size_t()
        _check(lib().sbdx_mlp_collect(self.h, max_pairs, seed, feat, tgt, C.byref(n)))
        return feat[: n.value * nf].reshape(n.value, nf), tgt[: n.value]

    def mlp_apply(self, R, E):
        """The learned step for residual R (alpha-first order) at Ritz value E."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        out = np.zeros_like(R)
        _check(lib().sbdx_mlp_apply(self.h, R, float(E), out))
        return out

    def set_preconditioner(self, mode):
        """0: diagonal R/(E0-d); 1: learned MLP with diagonal fallback on stall."""
        _check(lib().sbdx_set_preconditioner(self.h, int(mode)))

    def precond_stats(self):
        a, b = C.c_long(), C.c_long()
        _check(lib().sbdx_precond_stats(self.h, C.byref(a), C.byref(b)))
        return dict(mlp=a.value, diag=b.value)

    def set_restart(self, keep):
        """0: restart from W alone (reference); k >= 1: thick restart keeping
        k Ritz vectors, exact FP64 residual at every restart."""
        _check(lib().sbdx_set_restart(self.h, int(keep)))

    def set_symmetry(self, mode):
        """Spin-flip sectors of the thick-restart and block solves (equal
        string sets): 0/False general form, 1/True/"even" the even sector
        (symmetric start vector), 2/"odd" the odd sector (antisymmetric start
        vector), 3/"lowest" both sectors (thick: the lower result returned;
        block: one basis holding both sectors' columns, the lowest roots of
        H)."""
        names = {"off": 0, "even": 1, "odd": 2, "lowest": 3}
        m = names[mode] if isinstance(mode, str) else (1 if mode is True else int(mode))
        _check(lib().sbdx_set_symmetry(self.h, m))

    def sector_stats(self):
        """Per sector of the last thick solve: dict with energies [even, odd],
        products, states (0 not run, 1 converged, 2 not converged, 3 odd stage
        stopped above the even result) and chosen (1 even, 2 odd, 0 general)."""
        e = np.zeros(2)
        p = (C.c_int * 2)()
        st = (C.c_int * 2)()
        ch = C.c_int(0)
        _check(lib().sbdx_sector_stats(self.h, e, p, st, C.byref(ch)))
        return dict(energies=e.tolist(), products=list(p), states=list(st), chosen=ch.value)

    def sigma_sym(self, W, fp32=False):
        W = np.ascontiguousarray(W, dtype=np.float64)
        out = np.zeros_like(W)
        _check(lib().sbdx_sigma_sym(self.h, W, out, 1 if fp32 else 0))
        return out

    def set_coarse_start(self, m, eps=1e-2):
        """Refine the start vector on the first m strings of each set first
        (thick-restart solves); m = 0 switches it off."""
        _check(lib().sbdx_set_coarse_start(self.h, int(m), float(eps)))

    def set_mixed_form(self, form):
        """Mixed alpha-beta term of later products: 0 direct sliced-ELL
        kernels (default), 1 the dense (pq|rs) contraction as one GEMM per
        batch of beta rows, 2 the same with the split-FP16x3 GEMM for FP32
        products."""
        _check(lib().sbdx_set_mixed_form(self.h, int(form)))

    def fp64_switch_step(self):
        """Step of the last thick solve from which its products were FP64
        (the mixed-precision floor was reached), or -1."""
        n = C.c_int()
        _check(lib().sbdx_fp64_switch_step(self.h, C.byref(n)))
        return n.value

    def coarse_stats(self):
        n, e, t = C.c_int(), C.c_double(), C.c_double()
        _check(lib().sbdx_coarse_stats(self.h, C.byref(n), C.byref(e), C.byref(t)))
        return dict(products=n.value, energy=e.value, seconds=round(t.value, 4))

    def set_graphs(self, on):
        """Replay the FP32 symmetric sigma core as a captured hipGraph."""
        _check(lib().sbdx_set_graphs(self.h, 1 if on else 0))

    def set_verbose(self, level):
        """Nonzero: one stderr line per Davidson step (E0, |r|)."""
        _check(lib().sbdx_set_verbose(self.h, int(level)))

    def configure(self, solver):
        """'plain' (reference algorithm, FP64), 'thick' (FP64 thick restart),
        'mixed' (thick restart, FP32 correction products) or 'block' (the
        precision of block Davidson: FP32 correction products)."""
        keep, prec = {"plain": (0, PREC_FP64), "thick": (4, PREC_FP64),
                      "mixed": (4, PREC_MIXED), "block": (4, PREC_MIXED)}[solver]
        self.set_restart(keep)
        self.set_precision(prec)

    def energy(self, W):
        e = C.c_double()
        _check(lib().sbdx_energy(self.h, np.ascontiguousarray(W, dtype=np.float64), C.byref(e)))
        return e.value

    def correlation(self, W):
        """(one, two): one[s*L*L + i + L*j], two[(s+2t)*L^4 + i + L*j + L^2*k + L^3*l]."""
        L = self.norb
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        _check(lib().sbdx_rdm(self.h, np.ascontiguousarray(W, dtype=np.float64), one, two))
        return one, two

    def set_vector(self, W):
        _check(lib().sbdx_set_vector(self.h, np.ascontiguousarray(W, dtype=np.float64)))

    def get_vector(self):
        out = np.empty(self.ndet)
        _check(lib().sbdx_get_vector(self.h, out))
        return out

    def solve(self, max_it=100, nb=12, eps=1e-8, max_time=1e30):
        e = C.c_double()
        n = C.c_int()
        rc = _check(lib().sbdx_solve(self.h, max_it, nb, eps, max_time, C.byref(e), C.byref(n)),
                    allow=(5,))
        return e.value, n.value, rc == 0

    def sigma_block(self, V, fp32=False):
        """Products of the k rows of V (k x ndet) in one block call (FP32
        arithmetic with fp32=True)."""
        V = np.ascontiguousarray(np.atleast_2d(V), dtype=np.float64)
        out = np.zeros_like(V)
        if fp32:
            _check(lib().sbdx_sigma_block_fp32(self.h, V.shape[0], V.reshape(-1), out.reshape(-1)))
        else:
            _check(lib().sbdx_sigma_block(self.h, V.shape[0], V.reshape(-1), out.reshape(-1), 0))
        return out

    def davidson_block(self, nroots, block, W0=None, max_it=200, eps=1e-8, max_time=1e30):
        """Lowest nroots eigenpairs by block Davidson.  W0: nroots x ndet
        start vectors (default: HF and unit vectors at the next-lowest
        diagonal elements).  Returns (energies, W, products, converged)."""
        if W0 is None:
            W0 = np.zeros((nroots, self.ndet))
            order = np.argsort(self.make_diag(), kind="stable")[:nroots]
            for j, i in enumerate(order):
                W0[j, i] = 1.0
        W = np.ascontiguousarray(W0, dtype=np.float64).reshape(nroots, self.ndet).copy()
        E = np.zeros(nroots)
        n = C.c_int()
        rc = _check(lib().sbdx_davidson_block(self.h, nroots, block, W.reshape(-1), max_it, eps,
                                              max_time, E, C.byref(n)), allow=(5,))
        return E, W, n.value, rc == 0

    def set_block_basis(self, m):
        """Basis size of the block solver (0: as large as memory allows, up
        to 64)."""
        _check(lib().sbdx_set_block_basis(self.h, int(m)))

    def block_stats(self, nroots):
        """Basis size, Ritz vectors kept per restart, restarts and final
        residual norms of the last block solve."""
        m, k, r = C.c_int(), C.c_int(), C.c_int()
        res = np.zeros(nroots)
        par = np.zeros(nroots, dtype=np.int32)
        _check(lib().sbdx_block_stats(self.h, C.byref(m), C.byref(k), C.byref(r), res,
                                      par.ctypes.data_as(C.c_void_p), nroots))
        return dict(basis=m.value, keep=k.value, restarts=r.value, resid=[float(x) for x in res],
                    sectors=[int(x) for x in par])

    def trace(self):
        n = C.c_int()
        _check(lib().sbdx_trace(self.h, None, 0, C.byref(n)))
        buf = np.zeros(2 * max(n.value, 1))
        _check(lib().sbdx_trace(self.h, buf.ctypes.data, n.value, C.byref(n)))
        return buf[: 2 * n.value].reshape(n.value, 2)

    def table(self, spin, kind):
        n = self.na if spin == 0 else self.nb
        cnt = C.c_size_t()
        _check(lib().sbdx_table_count(self.h, spin, kind, C.byref(cnt)))
        off = np.zeros(n + 1, dtype=np.uint64)
        ket = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        cran = np.zeros(max(cnt.value, 1) * 2 * kind, dtype=np.int32)
        _check(lib().sbdx_table_export(self.h, spin, kind, off.ctypes.data, ket.ctypes.data,
                                       cran.ctypes.data))
        c = cnt.value
        return off.astype(np.int64), ket[:c], cran[: c * 2 * kind]

    def work_counts(self):
        out = np.zeros(6)
        _check(lib().sbdx_work_counts(self.h, out))
        return dict(alpha_singles=out[0], alpha_doubles=out[1], beta_singles=out[2],
                    beta_doubles=out[3], mixed_terms=out[4], ndet=out[5])

    def same_spin_forms(self):
        """Per precision: the same-spin form in use ('dense' GEMM, 'sparse'
        rows, None before the first sigma) and the autotuner's timings (ms),
        and the mixed kernel's tuned staging width and slots per thread for
        the general and the spin-flip symmetric form (0 before first use);
        for FP32 also the symmetric same-spin GEMM form (SGEMM or
        split-FP16x3) and its deciding timings."""
        out = {}
        for prec, key in ((0, "fp64"), (1, "fp32")):
            f, a, b = C.c_int(), C.c_float(), C.c_float()
            _check(lib().sbdx_same_spin_form(self.h, prec, C.byref(f), C.byref(a), C.byref(b)))
            out[key] = dict(form={-1: None, 0: "sparse", 1: "dense"}[f.value],
                            ms_sparse=round(a.value, 3), ms_dense=round(b.value, 3))
            w, nt, ms, mu = C.c_int(), C.c_int(), C.c_float(), C.c_int()
            _check(lib().sbdx_mixed_staging(self.h, prec, C.byref(w), C.byref(nt), C.byref(ms),
                                            C.byref(mu)))
            out[key].update(mixed_width=w.value, mixed_slots_per_thread=nt.value,
                            mixed_ms=round(ms.value, 3), block_multi_kernel=mu.value)
            _check(lib().sbdx_mixed_staging(self.h, prec + 2, C.byref(w), C.byref(nt), C.byref(ms),
                                            C.byref(mu)))
            out[key].update(sym_mixed_width=w.value, sym_mixed_slots_per_thread=nt.value,
                            sym_mixed_ms=round(ms.value, 3))
        m, a, b = C.c_int(), C.c_float(), C.c_float()
        _check(lib().sbdx_split_gemm(self.h, C.byref(m), C.byref(a), C.byref(b)))
        out["fp32"].update(sym_gemm={-1: None, 0: "sgemm", 1: "split_fp16x3"}[m.value],
                           sym_gemm_ms_sgemm=round(a.value, 3), sym_gemm_ms_split=round(b.value, 3))
        return out

    def profile(self, enable=True):
        _check(lib().sbdx_profile(self.h, 1 if enable else 0))

    def profile_stats(self):
        n = C.c_int()
        _check(lib().sbdx_profile_count(self.h, C.byref(n)))
        out = {}
        name = C.create_string_buffer(128)
        for i in range(n.value):
            cnt = C.c_long()
            ms = C.c_double()
            _check(lib().sbdx_profile_get(self.h, i, name, 128, C.byref(cnt), C.byref(ms)))
            out[name.value.decode()] = (cnt.value, ms.value)
        return out


def symmetric_eig(A):
    """Eigenvalues (ascending) and eigenvectors (columns) of a symmetric
    matrix by the library's subspace eigensolver (host code, no GPU)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    a = np.ascontiguousarray(A.T).reshape(-1)  # column-major
    V = np.zeros(n * n)
    E = np.zeros(n)
    _check(lib().sbdx_symmetric_eig(n, a, V, E))
    return E, V.reshape(n, n).T


def symmetric_geig(H, S):
    """Eigenvalues (ascending) and S-orthonormal eigenvectors (columns) of
    H y = E S y by the block solver's Rayleigh-Ritz step (host code, no
    GPU)."""
    H = np.asarray(H, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    n = H.shape[0]
    Y = np.zeros(n * n)
    E = np.zeros(n)
    _check(lib().sbdx_symmetric_geig(n, np.ascontiguousarray(H.T).reshape(-1),
                                     np.ascontiguousarray(S.T).reshape(-1), Y, E))
    return E, Y.reshape(n, n).T


def basis_init_vector(ndet):
    """W = e_0 (the Hartree-Fock determinant is index 0 after sorting),
    BasisInitVector with init == 0 (chemistry/tpb/davidson.h:104-144)."""
    W = np.zeros(ndet)
    W[0] = 1.0
    return W


def davidson(op: TPBOperator, W=None, max_it=100, nb=12, eps=1e-8, max_time=1e30):
    """Davidson on the device; returns (energy, W, n_steps, converged)."""
    if W is None:
        W = basis_init_vector(op.ndet)
    W = np.ascontiguousarray(W, dtype=np.float64).copy()
    e = C.c_double()
    n = C.c_int()
    rc = _check(lib().sbdx_davidson(op.h, W, max_it, nb, eps, max_time, C.byref(e), C.byref(n)),
                allow=(5,))
    return e.value, W, n.value, rc == 0


def diag(fcidump_path, adet_path, bdet_path=None, max_it=100, nb=12, eps=1e-8, do_rdm=True,
         bit_length=20, rdm_dir=None, loadname="", savename="", carryover_type=0,
         carryover_threshold=0.0, carryover_ratio=0.0):
    """FCIDUMP + string files -> energy <W|HW>, density and RDMs, like
    sbd::tpb::diag (chemistry/tpb/sbdiag.h:139-763) on one GPU.  With
    rdm_dir, also the driver's 1pRDM.txt / 2pRDM.txt there and its zero-,
    one- and two-body energies in out["energy_parts"]
    (apps/chemistry_tpb_selected_basis_diagonalization/main.cc:264-318).
    loadname / savename: start from a saved wavefunction (its coefficients
    on the determinants both sets share) instead of the HF vector, and save
    the result, as sbdiag.h:285-291 and :742-746 do.  carryover_type 1-3
    with carryover_threshold / carryover_ratio: the carry-over strings of
    sbdiag.h:692-735 (sbdx.carryover) in out["co_adet"], out["co_bdet"] and
    their weights in out["carryover"]."""
    fd = io.read_fcidump(fcidump_path)
    I0, I1, I2 = io.setup_integrals(fd)
    L = fd["norb"]
    a = io.read_strings(adet_path, L)
    b = a if bdet_path is None else io.read_strings(bdet_path, L)
    op = TPBOperator(L, I0, I1, I2, a, b, bit_length)
    W0 = None
    if loadname:
        W0 = io.load_wavefunction(loadname, L, a, b, bit_length)
        nrm = np.linalg.norm(W0)
        if nrm == 0.0:
            raise ValueError(f"{loadname}: no determinant in common with the string sets")
        W0 /= nrm
    _, W, steps, conv = davidson(op, W=W0, max_it=max_it, nb=nb, eps=eps)
    if savename:
        io.save_wavefunction(savename, L, a, b, W, bit_length)
    E = op.energy(W)
    out = dict(energy=E, W=W, steps=steps, converged=conv, norb=L)
    if carryover_type:
        from . import carryover as co
        out["co_adet"], out["co_bdet"], out["carryover"] = co.carryover(
            W, a, b, L, carryover_type, threshold=carryover_threshold, ratio=carryover_ratio)
    if do_rdm:
        one, two = op.correlation(W)
        out["one"], out["two"] = one, two
        out["density"] = np.array([one[s * L * L + i + L * i] for i in range(L) for s in range(2)])
        if rdm_dir is not None:
            io.write_rdm_text(rdm_dir, L, one, two)
            out["energy_parts"] = io.energy_decomposition(I0, io.spatial_h(L, I1), io.dense_eri(L, I2),
                                                          one, two)
    op.close()
    return out
