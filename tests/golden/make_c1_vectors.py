"""Writes tests/golden/c1_par2.json: the MD5 of every .par2 file the oracle's
creator restatement writes for the C1 configuration, of the input file, and
the erasure positions used by the repair check.  Run from the repo root:
    python tests/golden/make_c1_vectors.py
The digests pin the oracle against regressions; they come from the
restatement (the reference itself cannot be run here)."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import c1  # noqa: E402
import oracle  # noqa: E402


def main():
    data = c1.c1_input()
    vols = oracle.par2_create([(c1.C1_NAME, data)], block_size=c1.C1_BLOCK, redundancy=c1.C1_REDUNDANCY,
                              scheme=oracle.SCHEME_EXPONENTIAL, stem=c1.C1_STEM)
    out = {
        "config": {"name": c1.C1_NAME, "bytes": c1.C1_BYTES, "block_size": c1.C1_BLOCK,
                   "redundancy_percent": c1.C1_REDUNDANCY, "scheme": "exponential", "seed": c1.C1_SEED},
        "input_md5": hashlib.md5(data).hexdigest(),
        "files": {name: {"bytes": len(b), "md5": hashlib.md5(b).hexdigest()} for name, b in vols},
        "erasures": c1.c1_erasures(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_par2.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(path)


if __name__ == "__main__":
    main()
