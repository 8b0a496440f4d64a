"""Seeded synthetic TPB instances (no counterpart in the reference).

Integrals: h symmetric with h_pp ~ U(-2, 0) and h_pq ~ N(0, 0.1) (p < q);
Cholesky vectors L^k_pq symmetric in pq with entries N(0, 0.15);
(pq|rs) = sum_k L^k_pq L^k_rs, which is 8-fold symmetric and positive
semi-definite; core energy 0.

Strings: the Hartree-Fock string (lowest n orbitals), all its single
excitations, then a seeded subset of its double excitations until the set
holds ``nstr`` strings.  Singles plus doubles always exceed the requested
sizes of the standard workloads, so no random fill beyond doubles is ever
needed (if it were, random n-of-L strings are drawn).  The beta set equals
the alpha set.

Standard workloads:
  tiny16 : norb=16, n=5, 500 strings, 40 vectors    (250k determinants)
  c36    : norb=36, n=8, 8000 strings, 150 vectors  (64M determinants)
  c58    : norb=58, n=10, 20000 strings, 150 vectors (400M determinants)
"""
from __future__ import annotations

import itertools

import numpy as np

from . import io

WORKLOADS = {
    "tiny16": dict(norb=16, nelec_half=5, nstr=500, nchol=40, seed=1234),
    "c36": dict(norb=36, nelec_half=8, nstr=8000, nchol=150, seed=1234),
    "c58": dict(norb=58, nelec_half=10, nstr=20000, nchol=150, seed=1234),
}


def make_integrals(norb, nchol, seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(0.0, 0.1, size=(norb, norb))
    h = np.triu(h, 1)
    h = h + h.T
    h[np.diag_indices(norb)] = rng.uniform(-2.0, 0.0, size=norb)
    Lk = rng.normal(0.0, 0.15, size=(nchol, norb, norb))
    Lk = np.triu(Lk)
    Lk = Lk + np.transpose(np.triu(Lk, 1), (0, 2, 1))
    Lflat = Lk.reshape(nchol, norb * norb)
    eri = (Lflat.T @ Lflat).reshape(norb, norb, norb, norb)
    return h, eri


def pack_integrals(norb, h, eri, core=0.0):
    """(I0, I1, I2) in the reference layouts from dense h and (pq|rs)."""
    L2 = 2 * norb
    I1 = np.zeros((L2, L2))
    I1[0::2, 0::2] = h
    I1[1::2, 1::2] = h
    B = norb * (norb + 1) // 2
    I2 = np.zeros(io.i2_size(norb))
    p, q = np.tril_indices(norb)
    pq = io.tri(p, q)
    order = np.argsort(pq)
    p, q = p[order], q[order]
    mat = eri[p[:, None], q[:, None], p[None, :], q[None, :]]  # B x B
    a, b = np.tril_indices(B)
    I2[a * (a + 1) // 2 + b] = mat[a, b]
    return float(core), I1.reshape(-1), I2


def make_strings(norb, n, nstr, seed):
    rng = np.random.default_rng(seed + 1)
    hf = (1 << n) - 1
    occ = list(range(n))
    vir = list(range(n, norb))
    singles = [hf ^ (1 << i) ^ (1 << a) for i in occ for a in vir]
    out = [hf] + singles
    if len(out) >= nstr:
        return io.sort_unique(np.array(out[:nstr], dtype=np.uint64))
    doubles = [hf ^ (1 << i) ^ (1 << j) ^ (1 << a) ^ (1 << b)
               for i, j in itertools.combinations(occ, 2)
               for a, b in itertools.combinations(vir, 2)]
    need = nstr - len(out)
    if need <= len(doubles):
        pick = rng.choice(len(doubles), size=need, replace=False)
        out += [doubles[k] for k in sorted(pick)]
    else:
        out += doubles
        have = set(out)
        while len(out) < nstr:
            bits = rng.choice(norb, size=n, replace=False)
            v = int(sum(1 << int(b) for b in bits))
            if v not in have:
                have.add(v)
                out.append(v)
    return io.sort_unique(np.array(out, dtype=np.uint64))


def make_instance(norb, nelec_half, nstr, nchol, seed):
    h, eri = make_integrals(norb, nchol, seed)
    I0, I1, I2 = pack_integrals(norb, h, eri)
    strs = make_strings(norb, nelec_half, nstr, seed)
    return dict(norb=norb, nelec=2 * nelec_half, I0=I0, I1=I1, I2=I2,
                h=h, eri=eri, astr=strs, bstr=strs.copy(), seed=seed, nchol=nchol)


def workload(name):
    return make_instance(**WORKLOADS[name])


def write_instance(inst, fcidump_path, strings_path):
    io.write_fcidump(fcidump_path, inst["norb"], inst["nelec"], 0, inst["h"], inst["eri"],
                     inst["I0"])
    io.write_strings(strings_path, inst["astr"], inst["norb"])
