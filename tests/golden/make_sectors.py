"""Lowest energy of each spin-flip sector of a synthetic workload, by the
CPU oracle's sigma and scipy's Lanczos (eigsh), written to
tests/golden/sectors_<workload>.json.

With equal alpha and beta string sets H commutes with the exchange P of the
two strings; in the reference's interleaved convention
(P W)[a,b] = (-1)^(n_a n_b + popcount(a & b)) W[b,a], n_a = n_b electrons
per string, the sign that makes the Hartree-Fock vector even (P HF = HF).  The lowest eigenvalue of each
sector is the lowest of the projected operator Pi H Pi, Pi = (1 +- P)/2 (the
complement contributes zeros, above these negative energies).  The oracle's
sigma is pinned against the reference's own sigma on this workload
(tests/golden/ref_tiny16.npz), so these energies follow from the reference's
Hamiltonian; the reference's Davidson from the Hartree-Fock vector cannot
reach an odd ground state (the HF vector has no odd part).

usage: python tests/golden/make_sectors.py [workload]   (default tiny16)
"""
import json
import os
import sys
import time

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "sbd-pub_amd")]
from oracle import oracle  # noqa: E402
from sbdx import instances  # noqa: E402


def exchange_sign(strings):
    s = np.asarray(strings, dtype=np.uint64)
    ne = bin(int(s[0])).count("1")
    pc = np.array([[bin(int(x) & int(y)).count("1") for y in s] for x in s]) + ne * ne
    return np.where(pc % 2 == 0, 1.0, -1.0)


def main(name="tiny16"):
    inst = instances.workload(name)
    orc = oracle.Oracle.from_instance(inst)
    hii = orc.diag()
    n = inst["astr"].size
    sgn = exchange_sign(inst["astr"])

    def exch(w):
        return (sgn * w.reshape(n, n).T).ravel()

    out = dict(workload=name, ndet=n * n, method="scipy eigsh (k=2, SA, tol 1e-12) on the oracle sigma "
               "projected onto each sector")
    for key, par in (("even", 1.0), ("odd", -1.0)):
        def mv(x, par=par):
            x = np.asarray(x).ravel()
            x = 0.5 * (x + par * exch(x))
            y = orc.sigma(x, hii)
            return 0.5 * (y + par * exch(y))
        op = LinearOperator((n * n, n * n), matvec=mv, dtype=np.float64)
        t = time.time()
        w, v = eigsh(op, k=2, which="SA", tol=1e-12, v0=np.random.default_rng(5).standard_normal(n * n))
        x = v[:, 0]
        r = np.linalg.norm(mv(x) - w[0] * x)
        out[key] = dict(E=[float(e) for e in w], residual=float(r), hf_weight=float(x[0] ** 2))
        print(f"{key}: E = {w} |r| = {r:.2e} HF weight {x[0] ** 2:.3e} ({time.time() - t:.0f} s)", flush=True)
    out["ground"] = min(out["even"]["E"][0], out["odd"]["E"][0])
    out["ground_sector"] = "even" if out["even"]["E"][0] <= out["odd"]["E"][0] else "odd"
    with open(os.path.join(HERE, f"sectors_{name}.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:])
