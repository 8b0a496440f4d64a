"""Golden vectors for the RDM text files of the reference driver: runs the
reference app built from its own sources (oracle/_ref/diag, by
oracle/build_ref.sh) on the h2o-1em3 data set with --rdm 1 and keeps, in
tests/golden/rdm_text_h2o.npz:
  sha1p, sha2p   SHA-256 of the reference's 1pRDM.txt and 2pRDM.txt;
  lines1         1pRDM.txt as written (576 lines);
  idx2, lines2   a sample of 2pRDM.txt's lines (the first 1000, then every
                 97th) with their line numbers;
  nlines2        2pRDM.txt's line count;
  energies       the driver's printed total, zero-, one- and two-body
                 energies.
Test infrastructure; run here, where the reference tree and its build are:
    python tests/golden/make_rdm_text.py
"""
import gzip
import hashlib
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DIAG = os.path.join(ROOT, "oracle", "_ref", "diag")
DATA = os.path.join(HERE, "data")


def main():
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "fcidump.txt"), "w") as f:
            f.write(gzip.open(os.path.join(DATA, "h2o_fcidump.txt.gz"), "rt").read())
        args = [DIAG, "--fcidump", "fcidump.txt", "--adetfile", os.path.join(DATA, "h2o-1em3-alpha.txt"),
                "--method", "0", "--block", "12", "--iteration", "100", "--tolerance", "1.0e-10",
                "--adet_comm_size", "1", "--bdet_comm_size", "1", "--task_comm_size", "1", "--init", "0",
                "--shuffle", "0", "--carryover_type", "0", "--rdm", "1"]
        out = subprocess.run(args, cwd=wd, env=dict(os.environ, OMP_NUM_THREADS="8"), check=True,
                             capture_output=True, text=True).stdout
        b1 = open(os.path.join(wd, "1pRDM.txt"), "rb").read()
        b2 = open(os.path.join(wd, "2pRDM.txt"), "rb").read()
    lines2 = b2.split(b"\n")[:-1]
    idx = np.array(sorted(set(range(1000)) | set(range(0, len(lines2), 97))), dtype=np.int64)
    num = r"(-?[0-9.]+(?:e[-+][0-9]+)?)"
    E = [float(re.search(r"Sample-based diagonalization: Energy = " + num, out).group(1))]
    for tag in ("Zero-Body energy", "One-Body energy", "Two-Body energy"):
        E.append(float(re.search(tag + r" = " + num, out).group(1)))
    np.savez_compressed(os.path.join(HERE, "rdm_text_h2o.npz"),
                        sha1p=np.frombuffer(hashlib.sha256(b1).hexdigest().encode(), dtype=np.uint8),
                        sha2p=np.frombuffer(hashlib.sha256(b2).hexdigest().encode(), dtype=np.uint8),
                        lines1=np.frombuffer(b1, dtype=np.uint8), idx2=idx,
                        lines2=np.frombuffer(b"\n".join(lines2[i] for i in idx) + b"\n", dtype=np.uint8),
                        nlines2=np.array([len(lines2)]), energies=np.array(E))
    print(f"1pRDM {len(b1)} bytes, 2pRDM {len(lines2)} lines, {len(idx)} kept; energies {E}")


if __name__ == "__main__":
    main()
