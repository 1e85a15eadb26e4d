"""Writes tests/golden/stretched_stop.npz: the orbitals at which the oracle's P-LBFGS (the CPU
restatement of the reference optimizer, oracle/plbfgs.py, default options) stops after its
500-iteration cap on the stretched medium scan point (seed 0, n_bf=120, n_aux=360, n_act=10,
42 covalent structures, scan r=2.5), with its energy trace.

At r=2.5 both this optimizer and the device's run towards a point where two active orbitals
coalesce, the structure metric becomes singular and the energy surface is too steep for Armijo
steps above ~1e-6; tests/test_stretched_surface.py checks those properties at the stored point.
Runs on the CPU in about two minutes; run from the repository root after building the oracle."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "xmvb-cpp_amd"))
sys.path.insert(0, ROOT)
import xmvb_amd as vb  # noqa: E402
from oracle.plbfgs import plbfgs  # noqa: E402

CASE = dict(seed=0, n_bf=120, n_act=10, n_aux=360, scan_r=2.5, max_iterations=500)


def main():
    sp = vb.synthetic_problem(CASE["seed"], CASE["n_bf"], CASE["n_act"], CASE["n_aux"], CASE["scan_r"])
    topo = vb.Topology.build("COV", CASE["n_act"], CASE["n_act"], 1)
    res = plbfgs(CASE["n_bf"], CASE["n_aux"], topo, sp.S, sp.h, sp.B, sp.C, max_iterations=CASE["max_iterations"])
    print(f"{res['iterations']} iterations, {res['termination']}, E = {res['energy']:.9f}, "
          f"|g| = {res['gradient_norm']:.3f}", flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "stretched_stop.npz"), C=res["C"],
                        trace=np.asarray(res["trace"]), energy=res["energy"], iterations=res["iterations"],
                        **{k: v for k, v in CASE.items()})


if __name__ == "__main__":
    main()
