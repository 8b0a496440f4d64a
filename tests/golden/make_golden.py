"""Writes golden vectors produced by the reference implementation itself.

Runs oracle/_ref/ref_driver (the reference CPU path built from its own
headers by oracle/build_ref.sh) on each case and keeps a small digest of its
outputs in tests/golden/ref_<case>.npz:
  tables   entry counts and SHA-256 of (lengths, kets, cran) per list
  hii      sampled elements, sum, sum of squares
  sigma    H*v for a seeded unit vector v: sampled elements, norm, <v|Hv>
  davidson per-step (E0, |r|) printed by the reference, final <W|HW>
  W        |W| samples, sum W^2, sum |W| (sign-free)
  rdm      full 1-RDM; 2-RDM sampled elements and per-block sums/norms
Only these derived numbers are committed, never reference source.

usage: python tests/golden/make_golden.py [case ...]
"""
import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sbd-pub_amd"))
from sbdx import instances, io  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
DATA = os.path.join(HERE, "data")

CASES = {
    "h2o_1em3": dict(fcidump=os.path.join(DATA, "h2o_fcidump.txt.gz"),
                     strings=os.path.join(DATA, "h2o-1em3-alpha.txt"), nb=10),
    "n2_1em3": dict(fcidump=os.path.join(DATA, "n2_fcidump.txt.gz"),
                    strings=os.path.join(DATA, "n2-1em3-alpha.txt"), nb=10),
    "n2_3em4": dict(fcidump=os.path.join(DATA, "n2_fcidump.txt.gz"),
                    strings=os.path.join(DATA, "n2-3em4-alpha.txt"), nb=10, rdm=0),
    "syn10": dict(workload=dict(norb=10, nelec_half=3, nstr=80, nchol=10, seed=77), nb=12),
    "tiny16": dict(workload="tiny16", nb=12, max_it=3),
}

SEED = 7


def sha(arr):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(arr).tobytes()).digest(), dtype=np.uint8)


def read_table(path, w):
    b = np.fromfile(path, dtype=np.int64)
    n = int(b[1])
    ln = b[2:2 + n]
    tot = int(ln.sum())
    ket = b[2 + n:2 + n + tot]
    cr = b[2 + n + tot:2 + n + tot + w * tot]
    return ln, ket, cr


def run_case(name, spec, workdir):
    if "workload" in spec:
        wl = spec["workload"]
        inst = instances.workload(wl) if isinstance(wl, str) else instances.make_instance(**wl)
        fcf = os.path.join(workdir, "fcidump.txt")
        sf = os.path.join(workdir, "alpha.txt")
        instances.write_instance(inst, fcf, sf)
        norb, nstr = inst["norb"], inst["astr"].size
    else:
        fcf = os.path.join(workdir, "fcidump.txt")
        with open(fcf, "w") as f:
            import gzip
            with gzip.open(spec["fcidump"], "rt") as g:
                f.write(g.read())
        sf = spec["strings"]
        norb = io.read_fcidump(fcf)["norb"]
        nstr = io.read_strings(sf, norb).size
    ndet = nstr * nstr
    v = np.random.default_rng(SEED).standard_normal(ndet)
    v /= np.linalg.norm(v)
    vf = os.path.join(workdir, "vec.bin")
    v.tofile(vf)
    env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count()))
    cmd = [DRIVER, fcf, sf, "20", workdir, vf, str(spec.get("max_it", 100)), str(spec["nb"]),
           "1e-8", str(spec.get("rdm", 1))]
    out = subprocess.run(cmd, env=env, check=True, capture_output=True, text=True).stdout
    traj = [(float(m.group(2).split()[0]), float(m.group(1)))
            for m in re.finditer(r"Davidson iteration \S+ \(tol=(\S+)\):(.*)", out)]
    energy = float(re.search(r"energy (\S+)", out).group(1))
    rng = np.random.default_rng(1)
    d = dict(norb=norb, nstr=nstr, ndet=ndet, seed=SEED, energy=energy,
             trajectory=np.array(traj))
    for t, lists in ((0, ("sa", "da")), (1, ("sb", "db"))):
        for nm in lists:
            w = 2 if nm[0] == "s" else 4
            ln, ket, cr = read_table(os.path.join(workdir, f"table_t{t}_{nm}.bin"), w)
            d[f"table_{nm}_count"] = int(ln.sum())
            d[f"table_{nm}_sha"] = sha(np.concatenate([ln, ket, cr]).astype(np.int64))
    hii = np.fromfile(os.path.join(workdir, "hii.bin"))
    idx = np.sort(rng.choice(ndet, size=min(512, ndet), replace=False))
    d.update(hii_idx=idx, hii_val=hii[idx], hii_sum=hii.sum(), hii_sq=(hii * hii).sum())
    sg = np.fromfile(os.path.join(workdir, "sigma.bin"))
    d.update(sigma_idx=idx, sigma_val=sg[idx], sigma_norm=np.linalg.norm(sg), sigma_vhv=v @ sg)
    W = np.fromfile(os.path.join(workdir, "W.bin"))
    d.update(W_idx=idx, W_abs=np.abs(W[idx]), W_sq=(W * W).sum(), W_l1=np.abs(W).sum())
    if spec.get("rdm", 1):
        o1 = np.fromfile(os.path.join(workdir, "rdm1.bin"))
        o2 = np.fromfile(os.path.join(workdir, "rdm2.bin")).reshape(4, -1)
        L4 = norb ** 4
        i2 = np.sort(rng.choice(L4, size=min(4096, L4), replace=False))
        d.update(rdm1=o1, rdm2_idx=i2, rdm2_val=o2[:, i2], rdm2_sum=o2.sum(axis=1),
                 rdm2_norm=np.linalg.norm(o2, axis=1), rdm2_abs=np.abs(o2).sum(axis=1))
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **d)
    print(f"{name}: ndet={ndet} E={energy:.12f} steps={len(traj)}", flush=True)


def main(names):
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} missing: run oracle/build_ref.sh first")
    for name in names or list(CASES):
        with tempfile.TemporaryDirectory() as wd:
            run_case(name, CASES[name], wd)


if __name__ == "__main__":
    main(sys.argv[1:])
