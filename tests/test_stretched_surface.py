"""Why P-LBFGS stalls on the stretched medium scan point (seed 0, n_bf=120, n_aux=360, n_act=10,
scan r=2.5): a property of the surface, checked with the oracle at the point where the oracle's own
P-LBFGS (oracle/plbfgs.py, the reference optimizer's defaults) stops after 500 iterations
(tests/golden/stretched_stop.npz, written by tests/golden/make_stretched_stop.py). The device
optimizer runs into the same region (500 iterations, energy within 4e-4 Eh of this point).

There two active orbitals have coalesced (normalised overlap 1 - 2e-6), the structure metric is
singular to working precision (condition number ~1e11, structure coefficients ~1e3) and the energy
keeps falling towards that boundary along a valley so steep that a step of 1e-5 along the
steepest-descent direction already raises the energy, while the gradient itself is right (central
differences agree). Armijo steps therefore shrink to ~1e-6 and gain ~1e-6 Eh per iteration; with
rounding noise of ~1e-7 Eh in the energy at this conditioning, the line search can also fail
outright, which is how the earlier runs ended."""
import os

import numpy as np
import pytest

import oracle
import xmvb_amd as vb

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "stretched_stop.npz")


@pytest.fixture(scope="module")
def stop():
    d = np.load(FIXTURE)
    sp = vb.synthetic_problem(int(d["seed"]), int(d["n_bf"]), int(d["n_act"]), int(d["n_aux"]), float(d["scan_r"]))
    topo = vb.Topology.build("COV", int(d["n_act"]), int(d["n_act"]), 1)
    return sp, topo, d["C"]


def _energy(sp, topo, C, gradient=False):
    return oracle.evaluate(sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, C, gradient=gradient)


def _max_overlap(S_act):
    d = np.sqrt(np.diag(S_act))
    return float(np.abs(S_act / np.outer(d, d) - np.eye(len(d))).max())


def _cond(S):
    w = np.linalg.eigvalsh(0.5 * (S + S.T))
    return float(w[-1] / w[0])


def test_orbitals_coalesce_and_metric_becomes_singular(stop):
    sp, topo, C = stop
    start = _energy(sp, topo, sp.C)
    end = _energy(sp, topo, C)
    assert _max_overlap(start["S_act"]) < 0.5 and _cond(start["S"]) < 1e4
    assert _max_overlap(end["S_act"]) > 1.0 - 1e-5
    assert _cond(end["S"]) > 1e10
    assert np.abs(end["coef"]).max() > 100.0
    assert end["energy"] < start["energy"] - 80.0  # the valley is deep: -25.7 -> -110.8 Eh


def test_gradient_is_right_but_the_valley_admits_only_tiny_steps(stop):
    sp, topo, C = stop
    ref = _energy(sp, topo, C, gradient=True)
    g = ref["grad"]
    gn = float(np.linalg.norm(g))
    assert gn > 1.0  # far from stationary
    u = g / gn
    ep = _energy(sp, topo, C + 1e-5 * u)["energy"]
    em = _energy(sp, topo, C - 1e-5 * u)["energy"]
    assert abs((ep - em) / 2e-5 - gn) <= 0.01 * gn  # central differences confirm the gradient
    # a step of 1e-5 along steepest descent raises the energy (curvature ~1e7 Eh per unit^2) ...
    assert _energy(sp, topo, C - 1e-5 * u)["energy"] > ref["energy"] + 1e-5
    # ... while a step of 2e-7 lowers it: the direction is a descent direction
    assert _energy(sp, topo, C - 2e-7 * u)["energy"] < ref["energy"] - 1e-7
