"""Brute-force second quantization on tiny orbital spaces (test code).

Determinants are bitmasks over 2L spin orbitals in the interleaved order
(alpha orbital k -> 2k, beta k -> 2k+1); a determinant is the product of
creators in ascending spin-orbital order acting on the vacuum, so
a_k and a_k^dagger carry (-1)^(number of occupied spin orbitals below k).
H = sum h_pq a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q summed over spins,
applied operator by operator.  Nothing here uses the reference's
Slater-Condon formulas, so it pins them independently.
"""
import itertools

import numpy as np


def _ann(det, k):
    if not (det >> k) & 1:
        return None, 0
    sg = -1 if bin(det & ((1 << k) - 1)).count("1") & 1 else 1
    return det & ~(1 << k), sg


def _cre(det, k):
    if (det >> k) & 1:
        return None, 0
    sg = -1 if bin(det & ((1 << k) - 1)).count("1") & 1 else 1
    return det | (1 << k), sg


def apply_ops(det, ops):
    """ops = [(kind, k)] applied right to left; kind 'c' creator, 'a' annihilator."""
    sg = 1
    for kind, k in reversed(ops):
        det, s = (_cre if kind == "c" else _ann)(det, k)
        if det is None:
            return None, 0
        sg *= s
    return det, sg


def interleave(a, b, L):
    d = 0
    for k in range(L):
        if (a >> k) & 1:
            d |= 1 << (2 * k)
        if (b >> k) & 1:
            d |= 1 << (2 * k + 1)
    return d


def hamiltonian(L, h, eri, core, astr, bstr):
    """Dense H on the tensor-product basis, alpha-major order, reference
    (interleaved) determinant convention."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    n = len(dets)
    H = np.zeros((n, n))
    for j, dj in enumerate(dets):
        H[j, j] += core
        for p in range(L):
            for q in range(L):
                if h[p, q] == 0.0:
                    continue
                for s in (0, 1):
                    d, sg = apply_ops(dj, [("c", 2 * p + s), ("a", 2 * q + s)])
                    if d is not None and d in index:
                        H[index[d], j] += h[p, q] * sg
        for p, q, r, t in itertools.product(range(L), repeat=4):
            v = eri[p, q, r, t]
            if v == 0.0:
                continue
            for s1 in (0, 1):
                for s2 in (0, 1):
                    d, sg = apply_ops(dj, [("c", 2 * p + s1), ("c", 2 * r + s2),
                                           ("a", 2 * t + s2), ("a", 2 * q + s1)])
                    if d is not None and d in index:
                        H[index[d], j] += 0.5 * v * sg
    return H


def rdms(L, astr, bstr, W):
    """Physical one[s][i + L j] = <a+_is a_js>, two[s+2t][i + L j + L^2 k + L^3 l]
    = <a+_is a+_jt a_lt a_ks> of the vector W (interleaved convention)."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    one = np.zeros((2, L * L))
    two = np.zeros((4, L ** 4))
    for j, dj in enumerate(dets):
        wj = W[j]
        if wj == 0.0:
            continue
        for s in (0, 1):
            for i in range(L):
                for k in range(L):
                    d, sg = apply_ops(dj, [("c", 2 * i + s), ("a", 2 * k + s)])
                    if d is not None and d in index:
                        one[s, i + L * k] += W[index[d]] * wj * sg
        for s, t in itertools.product((0, 1), repeat=2):
            for i, jj, k, l in itertools.product(range(L), repeat=4):
                d, sg = apply_ops(dj, [("c", 2 * i + s), ("c", 2 * jj + t),
                                       ("a", 2 * l + t), ("a", 2 * k + s)])
                if d is not None and d in index:
                    two[s + 2 * t, i + L * jj + L * L * k + L ** 3 * l] += W[index[d]] * wj * sg
    return one, two
