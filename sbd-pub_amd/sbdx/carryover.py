"""Carry-over string selection of the reference's SQD loop, one rank
(chemistry/tpb/rdmat.h:10-230, chemistry/tpb/extend.h:14-324, applied by
chemistry/tpb/sbdiag.h:692-735 after the final energy).

Strings are integer masks (orbital i -> bit i), W is alpha-major
(W[ia*nb + ib]).  The string weights are the diagonal of the one-spin
reduced density matrix: D_a[ia] = sum_ib W[ia,ib]^2 and
D_b[ib] = sum_ia W[ia,ib]^2 (rdmat.h:10-62, 139-190).

Selections (carryover_type of sbdiag.h):
  1, threshold: the fewest strings, in descending weight, whose weights sum
     to more than 1 - threshold (rdmat.h:64-101);
  1, ratio:     the int(ratio * n) heaviest strings and the weight left out,
     1 - (their sum) (rdmat.h:103-136);
  2:            either of the above, then every selected string and all its
     single excitations, sorted and without duplicates (extend.h:186-322);
  3:            the strings that occur in a determinant of weight W^2 above
     the cutoff, with all their single excitations, sorted and without
     duplicates, and the total weight of those determinants
     (extend.h:14-184).
The reference orders equal weights with std::sort (unspecified order for
ties); here ties keep the lower index first.  Where the reference's
threshold rule finds no prefix above 1 - threshold (an unnormalized W) it
reads past the end of its list; here every string is kept.
"""
import numpy as np


def string_weights(W, na, nb):
    """(D_a, D_b): summed squared coefficients per alpha and per beta string."""
    X = np.asarray(W, dtype=np.float64).reshape(na, nb)
    X2 = X * X
    return X2.sum(axis=1), X2.sum(axis=0)


def _by_weight(D):
    # descending weight, ties by index
    return np.lexsort((np.arange(D.size), -D))


def select_threshold(strings, D, threshold):
    """Heaviest strings until their cumulative weight exceeds 1 - threshold
    (the string in which it does included)."""
    order = _by_weight(D)
    S = np.cumsum(D[order])
    k = int(np.searchsorted(S, 1.0 - threshold, side="right")) + 1
    return np.asarray(strings, dtype=np.uint64)[order[:min(k, D.size)]]


def select_ratio(strings, D, ratio):
    """The int(ratio * n) heaviest strings and the truncated weight."""
    n = int(ratio * D.size)
    order = _by_weight(D)[:n]
    return np.asarray(strings, dtype=np.uint64)[order], 1.0 - float(D[order].sum())


def singles(mask, norb):
    """The single excitations of one string (closed orbital outer, open
    inner, ascending: extend.h via excitation.h:9-30)."""
    m = int(mask)
    closed = [i for i in range(norb) if (m >> i) & 1]
    opened = [i for i in range(norb) if not (m >> i) & 1]
    return [(m & ~(1 << c)) | (1 << o) for c in closed for o in opened]


def singles_extend(strings, norb):
    """The strings and all their single excitations, sorted, unique
    (extend.h:186-322)."""
    out = [int(s) for s in strings]
    for s in strings:
        out.extend(singles(s, norb))
    return np.unique(np.array(out, dtype=np.uint64))


def singles_extend_weighted(W, adet, bdet, norb, cutoff):
    """(alpha strings, beta strings, total weight): the strings of the
    determinants with W^2 > cutoff, extended by their singles, sorted,
    unique; the summed weight of those determinants (extend.h:14-184)."""
    adet = np.asarray(adet, dtype=np.uint64)
    bdet = np.asarray(bdet, dtype=np.uint64)
    X2 = np.asarray(W, dtype=np.float64).reshape(adet.size, bdet.size) ** 2
    big = X2 > cutoff
    total = float(X2[big].sum())
    ra = singles_extend(adet[big.any(axis=1)], norb)
    rb = singles_extend(bdet[big.any(axis=0)], norb)
    return ra, rb, total


def carryover(W, adet, bdet, norb, carryover_type, threshold=0.0, ratio=0.0):
    """The carry-over strings of sbdiag.h:692-735: returns (co_adet, co_bdet,
    info) with info the truncated weights (type 1/2 with a ratio) or the
    total weight (type 3)."""
    adet = np.asarray(adet, dtype=np.uint64)
    bdet = np.asarray(bdet, dtype=np.uint64)
    info = {}
    if carryover_type in (1, 2):
        Da, Db = string_weights(W, adet.size, bdet.size)
        if ratio == 0.0:
            ca, cb = select_threshold(adet, Da, threshold), select_threshold(bdet, Db, threshold)
        else:
            ca, ta = select_ratio(adet, Da, ratio)
            cb, tb = select_ratio(bdet, Db, ratio)
            info = dict(truncated_alpha=ta, truncated_beta=tb)
        if carryover_type == 2:
            ca, cb = singles_extend(ca, norb), singles_extend(cb, norb)
        return ca, cb, info
    if carryover_type == 3:
        ca, cb, tot = singles_extend_weighted(W, adet, bdet, norb, threshold)
        return ca, cb, dict(total_weight=tot)
    return np.zeros(0, dtype=np.uint64), np.zeros(0, dtype=np.uint64), info
