"""Writes tests/golden/synthetic.json: SHA-256 of the seeded synthetic generator's arrays
(mt19937_64 + Marsaglia polar normals) so that later changes to the generator are caught."""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "xmvb-cpp_amd"))
import xmvb_amd as vb  # noqa: E402

CASES = {"medium": dict(seed=0, n_bf=120, n_act=10, n_aux=360), "small": dict(seed=20260427, n_bf=24, n_act=6, n_aux=48)}


def main():
    out = {}
    for key, c in CASES.items():
        sp = vb.synthetic_problem(c["seed"], c["n_bf"], c["n_act"], c["n_aux"])
        rec = dict(c)
        for f in ("S", "h", "B", "C"):
            arr = np.ascontiguousarray(getattr(sp, f).ravel(order="F") if f in "Sh" else getattr(sp, f).ravel())
            rec[f] = hashlib.sha256(arr.tobytes()).hexdigest()
        out[key] = rec
    json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "synthetic.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
