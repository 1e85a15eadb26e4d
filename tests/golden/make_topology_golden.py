"""Writes tests/golden/topology.json: structure/determinant counts, first raw structures and a
SHA-256 over the full expansion, produced by the oracle's independent restatement of the reference
generator/expander (oracle/vb_oracle.cpp) and checked there against the reference's own count
fixtures. The engine's topology must reproduce these bit for bit."""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

CASES = {"cov10": ("COV", 10, 10, 1), "cov12": ("COV", 12, 12, 1), "cov14": ("COV", 14, 14, 1),
         "full4": ("FULL", 4, 4, 1), "full6": ("FULL", 6, 6, 1), "h2_full": ("FULL", 2, 2, 1),
         "doublet5": ("COV", 5, 5, 2)}
FIELDS = ("alpha_masks", "beta_masks", "det_alpha", "det_beta", "term_offsets", "term_structure", "term_coefficient")


def main():
    out = {}
    for key, case in CASES.items():
        o = oracle.topology(*case)
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(o["raw_structures"], np.int32).tobytes())
        for f in FIELDS:
            h.update(np.ascontiguousarray(o[f]).tobytes())
        out[key] = dict(case=list(case), n_structures=int(o["n_structures"]), n_determinants=len(o["det_alpha"]),
                        first_structures=o["raw_structures"][:3].tolist(), sha256=h.hexdigest())
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "topology.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
