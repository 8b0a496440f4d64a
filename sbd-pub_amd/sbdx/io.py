"""FCIDUMP and alpha/beta string-file formats of the reference, plus the
integral layouts its TPB path consumes.

Formats (reference citations relative to the sbd-pub tree):

* FCIDUMP text: header lines until ``&END`` (``&FCI`` removed, ``KEY=VALUE``
  pairs split on commas); then ``value i j k l`` rows, 1-based orbitals;
  ``0 0 0 0`` is the core energy, ``i j 0 0`` a one-body element
  (framework/fcidump.h:29-141).  Values are written with 17 significant
  digits so a reader recovers identical doubles (fcidump.h:158-173).
* String files: one string per line, the rightmost character is orbital 0
  (framework/bit_manipulation.h:215-229); after loading the set is sorted
  numerically, most-significant word first, and made unique
  (bit_manipulation.h:149-168, 381-395; chemistry/basic/makedeterminants.h:160-185).
* Words: the reference stores a string as ``ceil(L/bit_length)`` size_t words,
  orbital i at bit ``i % bit_length`` of word ``i // bit_length``.
* Integrals (chemistry/basic/makeintegrals.h:123-174, integrals.h:10-107):
  I1 is the spin-orbital (2L)x(2L) matrix, row-major, spin orbital 2p+s;
  I2 is the 8-fold compressed spatial store of size B(B+1)/2, B = L(L+1)/2,
  indexed by tri(tri(p,q), tri(r,s)), tri(a,b) = max(a,b)(max+1)/2 + min.
"""
from __future__ import annotations

import gzip
import os

import numpy as np

MAX_NORB = 64


def _open_text(path):
    return gzip.open(path, "rt") if str(path).endswith(".gz") else open(path, "r")


def tri(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    mx = np.maximum(a, b)
    mn = np.minimum(a, b)
    return mx * (mx + 1) // 2 + mn


def i2_size(norb: int) -> int:
    B = norb * (norb + 1) // 2
    return B * (B + 1) // 2


# ---------------------------------------------------------------- FCIDUMP
def read_fcidump(path):
    """Parse an FCIDUMP file.  Returns a dict with ``header`` (str->str),
    ``norb``, ``nelec``, ``ms2``, ``vals`` (float64[n]) and ``idx`` (int32[n,4])."""
    header_text = []
    vals = []
    idx = []
    in_header = True
    with _open_text(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "!":
                continue
            if in_header:
                if "&END" in line or "/" == line.strip():
                    in_header = False
                    continue
                header_text.append(line.replace("&FCI", "", 1))
                continue
            parts = line.split()
            if len(parts) < 5:
                continue
            vals.append(float(parts[0].replace("D", "E").replace("d", "e")))
            idx.append([int(parts[1]), int(parts[2]), int(parts[3]), int(parts[4])])
    header = {}
    for tok in "".join(header_text).split(","):
        if "=" in tok:
            k, v = tok.split("=", 1)
            header[k.strip()] = v.strip()
    if "NORB" not in header:
        raise RuntimeError(f"FCIDUMP {path}: NORB missing from header")
    out = {
        "header": header,
        "norb": int(header["NORB"]),
        "nelec": int(header.get("NELEC", "0")),
        "ms2": int(header.get("MS2", "0")),
        "vals": np.asarray(vals, dtype=np.float64),
        "idx": np.asarray(idx, dtype=np.int32).reshape(-1, 4),
    }
    return out


def write_fcidump(path, norb, nelec, ms2, h, eri, core=0.0):
    """Write an FCIDUMP: every 8-fold-unique (pq|rs) with p>=q, r>=s,
    pq>=rs, then h_pq for p>=q, then the core energy, 17 significant digits."""
    with open(path, "w") as f:
        f.write(f" &FCI NORB={norb},NELEC={nelec},MS2={ms2},\n")
        f.write("  ORBSYM=" + "1," * norb + "\n")
        f.write("  ISYM=1,\n &END\n")
        lines = []
        for p in range(norb):
            for q in range(p + 1):
                pq = p * (p + 1) // 2 + q
                for r in range(norb):
                    for s in range(r + 1):
                        rs = r * (r + 1) // 2 + s
                        if rs > pq:
                            continue
                        lines.append(f"{eri[p, q, r, s]:.17g} {p+1} {q+1} {r+1} {s+1}\n")
        f.writelines(lines)
        for p in range(norb):
            for q in range(p + 1):
                f.write(f"{h[p, q]:.17g} {p+1} {q+1} 0 0\n")
        f.write(f"{core:.17g} 0 0 0 0\n")


def setup_integrals(fd):
    """Build (I0, I1, I2) from a parsed FCIDUMP, with the reference's rules:
    later rows overwrite earlier ones; one-body rows fill both spins and
    both index orders."""
    L = fd["norb"]
    L2 = 2 * L
    I1 = np.zeros((L2, L2), dtype=np.float64)
    I2 = np.zeros(i2_size(L), dtype=np.float64)
    I0 = 0.0
    vals = fd["vals"]
    idx = fd["idx"]
    core = np.all(idx == 0, axis=1)
    one = (idx[:, 2] == 0) & (idx[:, 3] == 0) & ~core
    two = ~(core | one)
    if core.any():
        I0 = float(vals[np.nonzero(core)[0][-1]])
    # one-body: apply in file order so duplicates resolve like the reference
    oi = np.nonzero(one)[0]
    for r in oi:
        i, j = idx[r, 0] - 1, idx[r, 1] - 1
        v = vals[r]
        I1[2 * i, 2 * j] = v
        I1[2 * i + 1, 2 * j + 1] = v
        I1[2 * j, 2 * i] = v
        I1[2 * j + 1, 2 * i + 1] = v
    ti = np.nonzero(two)[0]
    p, q, r, s = (idx[ti, k] - 1 for k in range(4))
    pos = tri(tri(p, q), tri(r, s))
    # numpy fancy assignment keeps the LAST write for repeated positions
    I2[pos] = vals[ti]
    return I0, I1.reshape(-1), I2


def dense_eri(norb, I2):
    """Expand the compressed store to a dense (pq|rs) array [p,q,r,s]."""
    p = np.arange(norb)
    pq = tri(p[:, None], p[None, :])
    return I2[tri(pq[:, :, None, None], pq[None, None, :, :])]


def spatial_h(norb, I1):
    I1 = np.asarray(I1).reshape(2 * norb, 2 * norb)
    return I1[0::2, 0::2].copy()


# ---------------------------------------------------------------- strings
def parse_string(s: str, norb: int) -> int:
    """Rightmost character is orbital 0."""
    s = s.strip()
    if len(s) < norb:
        raise RuntimeError(f"string '{s}' shorter than norb={norb}")
    v = 0
    for i in range(norb):
        if s[len(s) - 1 - i] == "1":
            v |= 1 << i
    return v


def format_string(v: int, norb: int) -> str:
    return "".join("1" if (int(v) >> i) & 1 else "0" for i in range(norb - 1, -1, -1))


def sort_unique(masks) -> np.ndarray:
    """Numeric order (most-significant word first == integer order for
    single-word strings), duplicates removed."""
    return np.unique(np.asarray(masks, dtype=np.uint64))


def read_strings(path, norb):
    if norb > MAX_NORB:
        raise RuntimeError(f"norb={norb} > {MAX_NORB} is not supported")
    vals = []
    with _open_text(path) as f:
        for line in f:
            if line.strip():
                vals.append(parse_string(line, norb))
    return sort_unique(np.array(vals, dtype=np.uint64))


def write_strings(path, masks, norb):
    with open(path, "w") as f:
        for m in masks:
            f.write(format_string(int(m), norb) + "\n")


def masks_to_words(masks, norb, bit_length=20):
    """Reference word layout: orbital i -> bit i%bl of word i//bl."""
    nw = (norb + bit_length - 1) // bit_length
    masks = np.asarray(masks, dtype=np.uint64)
    out = np.zeros((masks.size, nw), dtype=np.uint64)
    for w in range(nw):
        lo = w * bit_length
        nbits = min(bit_length, norb - lo)
        out[:, w] = (masks >> np.uint64(lo)) & np.uint64((1 << nbits) - 1)
    return out


def words_to_masks(words, norb, bit_length=20):
    words = np.asarray(words, dtype=np.uint64)
    if words.ndim == 1:
        words = words[:, None]
    out = np.zeros(words.shape[0], dtype=np.uint64)
    for w in range(words.shape[1]):
        out |= words[:, w] << np.uint64(w * bit_length)
    return out


def _g16(v: float) -> str:
    # std::ostream with precision(16) and the default float field prints as
    # printf("%.16g")
    return f"{v:.16g}"


def write_rdm_text(directory, norb, one, two):
    """1pRDM.txt and 2pRDM.txt in the reference driver's text format
    (apps/chemistry_tpb_selected_basis_diagonalization/main.cc:275-303):
    "i j value" for i, j (outer to inner) with value one[0][i+L*j] +
    one[1][i+L*j], and "i j k l value" for i, j, k, l with the four spin
    blocks of two[.][i+L*j+L^2*(k+L*l)] summed in block order; 16
    significant digits.  one: 2*L^2 values, two: 4*L^4 (the layouts of
    TPBOperator.correlation)."""
    L = int(norb)
    o = np.asarray(one, dtype=np.float64).reshape(2, L * L)
    t = np.asarray(two, dtype=np.float64).reshape(4, L ** 4)
    g1 = o[0] + o[1]
    g2 = ((t[0] + t[1]) + t[2]) + t[3]
    with open(os.path.join(directory, "1pRDM.txt"), "w") as f:
        f.write("".join(f"{i} {j} {_g16(g1[i + L * j])}\n" for i in range(L) for j in range(L)))
    with open(os.path.join(directory, "2pRDM.txt"), "w") as f:
        for i in range(L):
            f.write("".join(f"{i} {j} {k} {l} {_g16(g2[i + L * j + L * L * (k + L * l)])}\n"
                            for j in range(L) for k in range(L) for l in range(L)))


def read_rdm_text(directory, norb):
    """The spin-summed 1- and 2-RDM from 1pRDM.txt / 2pRDM.txt, as flat
    arrays indexed i+L*j and i+L*j+L^2*(k+L*l)."""
    L = int(norb)
    g1 = np.zeros(L * L)
    for line in open(os.path.join(directory, "1pRDM.txt")):
        i, j, v = line.split()
        g1[int(i) + L * int(j)] = float(v)
    g2 = np.zeros(L ** 4)
    for line in open(os.path.join(directory, "2pRDM.txt")):
        i, j, k, l, v = line.split()
        g2[int(i) + L * int(j) + L * L * (int(k) + L * int(l))] = float(v)
    return g1, g2


def energy_decomposition(I0, h, eri, one, two):
    """Zero-, one- and two-body energies from the RDMs, as the reference
    driver prints them (main.cc:279-281, 296-299): I0; sum_ij h_ij (one[0] +
    one[1])[i+L*j]; 1/2 sum_ijkl (ik|jl) sum_s two[s][i+L*j+L^2*(k+L*l)].
    h: L x L, eri: L^4 array indexed [p, q, r, s] = (pq|rs)."""
    h = np.asarray(h, dtype=np.float64)
    L = h.shape[0]
    o = np.asarray(one, dtype=np.float64).reshape(2, L, L)       # [s, j, i]
    t = np.asarray(two, dtype=np.float64).reshape(4, L, L, L, L)  # [s, l, k, j, i]
    g1 = (o[0] + o[1]).T                                          # [i, j]
    g2 = t.sum(axis=0).transpose(3, 2, 1, 0)                      # [i, j, k, l]
    e1 = float(np.einsum("ij,ij->", h, g1))
    e2 = 0.5 * float(np.einsum("ikjl,ijkl->", np.asarray(eri).reshape(L, L, L, L), g2))
    return dict(zero=float(I0), one=e1, two=e2, total=float(I0) + e1 + e2)


def _wavefunction_path(file):
    # one-rank file name: the basename plus the six-digit rank 0
    return f"{file}000000"


def save_wavefunction(file, norb, adet, bdet, W, bit_length=20):
    """The reference's wavefunction file for one rank
    (chemistry/tpb/restart.h:20-67, file name + "000000"): size_t number of
    alpha strings, of beta strings and words per string; the alpha strings'
    words, the beta strings' words (bit_length bits per word, orbital i in
    word i/bl, bit i%bl); then W (double, alpha-major)."""
    aw = masks_to_words(adet, norb, bit_length)
    bw = masks_to_words(bdet, norb, bit_length)
    W = np.ascontiguousarray(W, dtype=np.float64)
    if W.size != aw.shape[0] * bw.shape[0]:
        raise ValueError("W does not match the string sets")
    with open(_wavefunction_path(file), "wb") as f:
        np.array([aw.shape[0], bw.shape[0], aw.shape[1]], dtype=np.uint64).tofile(f)
        aw.tofile(f)
        bw.tofile(f)
        W.tofile(f)


def load_wavefunction(file, norb, adet, bdet, bit_length=20):
    """W over the given string sets from a wavefunction file
    (chemistry/tpb/restart.h:414-660 for one rank): zero, then for every
    stored (alpha, beta) pair whose strings both occur in the sets the stored
    coefficient at their position (restart from the intersection of the old
    and new determinant sets)."""
    with open(_wavefunction_path(file), "rb") as f:
        na, nb, nw = (int(v) for v in np.fromfile(f, dtype=np.uint64, count=3))
        aw = np.fromfile(f, dtype=np.uint64, count=na * nw).reshape(na, nw)
        bw = np.fromfile(f, dtype=np.uint64, count=nb * nw).reshape(nb, nw)
        Wl = np.fromfile(f, dtype=np.float64, count=na * nb)
    if Wl.size != na * nb:
        raise ValueError(f"{_wavefunction_path(file)}: truncated")
    cur_a = masks_to_words(adet, norb, bit_length)
    cur_b = masks_to_words(bdet, norb, bit_length)
    if cur_a.shape[1] != nw:
        raise ValueError("word count of the file does not match norb / bit_length")
    ia = {row.tobytes(): k for k, row in enumerate(cur_a)}
    ib = {row.tobytes(): k for k, row in enumerate(cur_b)}
    ja = np.array([ia.get(row.tobytes(), -1) for row in aw], dtype=np.int64)
    jb = np.array([ib.get(row.tobytes(), -1) for row in bw], dtype=np.int64)
    W = np.zeros(cur_a.shape[0] * cur_b.shape[0])
    sa, sb = np.nonzero(ja >= 0)[0], np.nonzero(jb >= 0)[0]
    if sa.size and sb.size:
        W.reshape(cur_a.shape[0], cur_b.shape[0])[np.ix_(ja[sa], jb[sb])] += \
            Wl.reshape(na, nb)[np.ix_(sa, sb)]
    return W
