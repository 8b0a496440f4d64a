"""Times one sigma of the reference CPU path (oracle/_ref/ref_driver, built
from the reference's own sources by oracle/build_ref.sh) against the CPU
oracle's sigma on the same instances and vector with the same thread
count, and checks that the two results agree.  The ratio calibrates
bench.py's CPU baseline, which times the oracle because the reference
cannot run where the GPU is.

TEST INFRASTRUCTURE ONLY.  usage: OMP_NUM_THREADS=8 python oracle/calibrate.py
"""
import gzip
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# the repository root, not this directory, resolves `oracle` (a namespace
# package holding oracle.py)
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path[:0] = [ROOT, os.path.join(ROOT, "sbd-pub_amd"), os.path.join(ROOT, "tests")]
from conftest import DATA, molecule  # noqa: E402
from oracle import oracle  # noqa: E402
from sbdx import instances  # noqa: E402

DRIVER = os.path.join(HERE, "_ref", "ref_driver")
THREADS = int(os.environ.get("OMP_NUM_THREADS", "8"))


def run(name, fcidump, strings, inst, reps):
    n = inst["astr"].size
    v = np.random.default_rng(7).standard_normal(n * n)
    v /= np.linalg.norm(v)
    with tempfile.TemporaryDirectory() as wd:
        vf = os.path.join(wd, "v.bin")
        v.tofile(vf)
        t_ref = []
        for _ in range(reps):
            out = subprocess.run([DRIVER, fcidump, strings, "20", wd, vf, "0", "12", "1e-8", "0"],
                                 env=dict(os.environ, OMP_NUM_THREADS=str(THREADS)), check=True,
                                 capture_output=True, text=True).stdout
            t_ref.append(float(re.search(r"sigma (\S+) s", out).group(1)))
        s_ref = np.fromfile(os.path.join(wd, "sigma.bin"))
    orc = oracle.Oracle.from_instance(inst, THREADS)
    hii = orc.diag()
    t_orc = []
    for _ in range(reps):
        t = time.perf_counter()
        s = orc.sigma(v, hii)
        t_orc.append(time.perf_counter() - t)
    orc.close()
    err = np.max(np.abs(s - s_ref)) / np.max(np.abs(s_ref))
    print(f"{name}: ndet={n * n} reference sigma {min(t_ref):.3f} s, oracle sigma {min(t_orc):.3f} s, "
          f"ratio reference/oracle {min(t_ref) / min(t_orc):.2f}, max rel diff {err:.1e} "
          f"({THREADS} threads, best of {reps})", flush=True)


def main():
    with tempfile.TemporaryDirectory() as wd:
        inst = instances.workload("tiny16")
        f, s = os.path.join(wd, "tiny16.fcidump"), os.path.join(wd, "tiny16.txt")
        instances.write_instance(inst, f, s)
        run("tiny16 (norb 16, 5+5 e, 500 strings)", f, s, inst, 3)
        inst = instances.make_instance(norb=24, nelec_half=6, nstr=1500, nchol=60, seed=11)
        f, s = os.path.join(wd, "syn24.fcidump"), os.path.join(wd, "syn24.txt")
        instances.write_instance(inst, f, s)
        run("syn24 (norb 24, 6+6 e, 1500 strings)", f, s, inst, 2)
        f = os.path.join(wd, "h2o.fcidump")
        with open(f, "w") as out:
            out.write(gzip.open(os.path.join(DATA, "h2o_fcidump.txt.gz"), "rt").read())
        run("h2o-1em4 (norb 24, 5+5 e, 1543 strings)", f, os.path.join(DATA, "h2o-1em4-alpha.txt"),
            molecule("h2o", "1em4"), 2)


if __name__ == "__main__":
    main()
