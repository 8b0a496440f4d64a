"""Golden vectors for the carry-over selection: runs the reference app built
from its own sources (oracle/_ref/diag, by oracle/build_ref.sh) on the
h2o-1em3 data set with each carry-over option and keeps, in
tests/golden/carryover_h2o.npz, the strings it writes
(--carryover_adetfile / --carryover_bdetfile, as integer masks, in the
file's order) and the weights it prints.
Test infrastructure; run here, where the reference tree and its build are:
    python tests/golden/make_carryover.py
"""
import gzip
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sbd-pub_amd"))
from sbdx import io  # noqa: E402

DIAG = os.path.join(ROOT, "oracle", "_ref", "diag")
DATA = os.path.join(HERE, "data")
CASES = {"t1_thr": ("1", "--carryover_threshold", "1e-3"), "t1_ratio": ("1", "--carryover_ratio", "0.2"),
         "t2_thr": ("2", "--carryover_threshold", "1e-2"), "t3_thr": ("3", "--carryover_threshold", "1e-6")}


def main():
    out = {}
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "fcidump.txt"), "w") as f:
            f.write(gzip.open(os.path.join(DATA, "h2o_fcidump.txt.gz"), "rt").read())
        for key, (t, opt, val) in CASES.items():
            args = [DIAG, "--fcidump", "fcidump.txt", "--adetfile", os.path.join(DATA, "h2o-1em3-alpha.txt"),
                    "--method", "0", "--block", "12", "--iteration", "100", "--tolerance", "1.0e-10",
                    "--adet_comm_size", "1", "--bdet_comm_size", "1", "--task_comm_size", "1", "--init", "0",
                    "--shuffle", "0", "--carryover_type", t, opt, val, "--carryover_adetfile", "co_a.txt",
                    "--carryover_bdetfile", "co_b.txt", "--rdm", "0"]
            log = subprocess.run(args, cwd=wd, env=dict(os.environ, OMP_NUM_THREADS="8"), check=True,
                                 capture_output=True, text=True).stdout
            for s in ("a", "b"):
                lines = [ln for ln in open(os.path.join(wd, f"co_{s}.txt")).read().splitlines() if ln.strip()]
                out[f"{key}_{s}"] = np.array([io.parse_string(ln, 24) for ln in lines], dtype=np.uint64)
            nums = [float(x) for x in re.findall(r"truncated weight in carry-over for \w+-det = (\S+)", log)]
            nums += [float(x) / 100.0 for x in re.findall(r"percentage of wf used in augmentation: (\S+)", log)]
            out[f"{key}_w"] = np.array(nums)
            print(key, out[f"{key}_a"].size, out[f"{key}_b"].size, nums)
    np.savez_compressed(os.path.join(HERE, "carryover_h2o.npz"), **out)


if __name__ == "__main__":
    main()
