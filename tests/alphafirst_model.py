"""numpy model of the sigma decomposition the HIP kernels implement.

The kernels work in the "alpha-first" determinant basis, where a
determinant is (alpha creators ascending)(beta creators ascending)|0>.
The reference orders creators by interleaved spin-orbital index
(alpha k -> 2k, beta k -> 2k+1), so the two bases differ by the diagonal
sign S(Ia,Ib) = (-1)^#{(p in Ia, r in Ib): r < p}.  In the alpha-first
basis an alpha excitation's phase depends on the alpha string only and a
beta excitation's on the beta string only, so

  sigma' = hii*C' + A_same(C') + B_same(C') + AB(C'),   C' = S*C, sigma = S*sigma'

with, for bra string I and ket J of one spin,
  single p->q (p bra-only, q ket-only): s*(F + sum_{k in other} (pq|kk))
      F = h_pq + sum_{k in I} ((pq|kk) - (pk|kq))
  double (i<j bra-only, a<b ket-only): reference TwoExcite sign evaluated on
      the one-spin string, times (ai|bj) - (aj|bi)
  mixed: s_alpha * s_beta * (pq|rs).
This is test code: it pins the decomposition against the oracle on the CPU.
"""
import numpy as np


def popc(x):
    return bin(int(x)).count("1")


def parity1(s, a, b):
    """Reference parity rule on a one-spin string: bits in [lo,hi), flip if bit lo set."""
    lo, hi = min(a, b), max(a, b)
    mask = ((1 << hi) - 1) ^ ((1 << lo) - 1)
    sg = -1.0 if popc(s & mask) & 1 else 1.0
    if (s >> lo) & 1:
        sg = -sg
    return sg


def s_sign(a, b):
    n = 0
    for p in range(64):
        if (a >> p) & 1:
            n += popc(b & ((1 << p) - 1))
    return -1.0 if n & 1 else 1.0


def bits(x):
    return [k for k in range(64) if (x >> k) & 1]


def spin_tables(strs, h, eri):
    strs = [int(s) for s in strs]
    n = len(strs)
    singles = [[] for _ in range(n)]
    doubles = [[] for _ in range(n)]
    for I in range(n):
        for J in range(n):
            x = strs[I] ^ strs[J]
            d = popc(x)
            cr = strs[I] & ~strs[J]
            an = strs[J] & ~strs[I]
            if d == 2:
                p, q = bits(cr)[0], bits(an)[0]
                s = parity1(strs[I], p, q)
                F = h[p, q] + sum(eri[p, q, k, k] - eri[p, k, k, q] for k in bits(strs[I]))
                singles[I].append((J, p, q, s, s * F))
            elif d == 4:
                i, j = bits(cr)
                a, b = bits(an)
                sg = parity1(strs[I], i, a) * parity1(strs[I], j, b)
                if a > j or b < i:
                    sg = -sg
                doubles[I].append((J, sg * (eri[a, i, b, j] - eri[a, j, b, i])))
    return singles, doubles


def sigma_alpha_first(astr, bstr, h, eri, hii, C):
    na, nb = len(astr), len(bstr)
    Sa, Da = spin_tables(astr, h, eri)
    Sb, Db = spin_tables(bstr, h, eri)
    S = np.array([[s_sign(int(a), int(b)) for b in bstr] for a in astr])
    Cp = (S * C.reshape(na, nb))
    sig = hii.reshape(na, nb) * Cp
    occa = [bits(int(a)) for a in astr]
    occb = [bits(int(b)) for b in bstr]
    for I in range(na):
        for (J, p, q, s, val) in Sa[I]:
            for ib in range(nb):
                sig[I, ib] += (val + s * sum(eri[p, q, k, k] for k in occb[ib])) * Cp[J, ib]
        for (J, val) in Da[I]:
            sig[I, :] += val * Cp[J, :]
    for ib in range(nb):
        for (J, p, q, s, val) in Sb[ib]:
            for ia in range(na):
                sig[ia, ib] += (val + s * sum(eri[p, q, k, k] for k in occa[ia])) * Cp[ia, J]
        for (J, val) in Db[ib]:
            sig[:, ib] += val * Cp[:, J]
    for I in range(na):
        for (J, p, q, sa, _) in Sa[I]:
            for ib in range(nb):
                for (Jb, r, t, sb, _) in Sb[ib]:
                    sig[I, ib] += sa * sb * eri[p, q, r, t] * Cp[J, Jb]
    return (S * sig).reshape(-1)
