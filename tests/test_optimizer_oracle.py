"""The CPU restatement of the projected L-BFGS optimisation (oracle/plbfgs.py), pinned by what the
mathematics fixes rather than by its own formulas: it descends monotonically (Armijo), stops at a
stationary point of the oracle energy (projected gradient below tolerance, <g_i, c_i> = 0), that
point is a local minimum (no lower energy along random tangent directions), and where the surface
has a single basin it lands where an independent optimiser (scipy BFGS) lands."""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from oracle.plbfgs import plbfgs
import xmvb_amd as vb


def _case(kw, n_act, n_el, n_bf, n_aux, seed, split_support):
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_el, 1)
    support = None
    if split_support:  # each orbital on its own block of basis functions, as atom-localised orbitals
        support = np.zeros((n_act, n_bf), np.uint8)
        w = n_bf // n_act
        for i in range(n_act):
            support[i, i * w:(i + 1) * w] = 1
    return sp, topo, support


def _tangent_perturbations(C, S, mask, rng, count):
    for _ in range(count):
        d = np.where(mask, rng.standard_normal(C.shape), 0.0)
        Sc = np.where(mask, C @ S, 0.0)
        d -= C * np.einsum("ik,ik->i", Sc, d)[:, None]  # S-orthogonal to each orbital
        yield d / np.abs(d).max()


@pytest.mark.parametrize("kw,n_act,n_el,n_bf,n_aux,seed,split", [
    ("FULL", 2, 2, 10, 30, 1, True),   # two orbitals on disjoint supports, 3 structures
    ("COV", 4, 4, 16, 48, 2, False),
])
def test_plbfgs_reaches_a_local_minimum(kw, n_act, n_el, n_bf, n_aux, seed, split):
    sp, topo, support = _case(kw, n_act, n_el, n_bf, n_aux, seed, split)
    r = plbfgs(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, support=support, gradient_tolerance=1e-6,
               energy_tolerance=1e-12)
    assert r["converged"] and r["termination"] == "plbfgs_projected_gradient_tolerance"
    assert np.all(np.diff(r["trace"]) <= 1e-12)
    ref = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, r["C"], support=support)
    assert abs(ref["energy"] - r["energy"]) <= 1e-12 * abs(r["energy"])
    g = ref["grad"]
    assert np.abs(np.einsum("ik,ik->i", g, r["C"])).max() <= 1e-9
    mask = np.ones((n_act, n_bf), bool) if support is None else support.astype(bool)
    Sinv = [np.zeros_like(sp.S) for _ in range(n_act)]
    for i in range(n_act):
        idx = np.flatnonzero(mask[i])
        Sinv[i][np.ix_(idx, idx)] = np.linalg.inv(sp.S[np.ix_(idx, idx)])
    nu = max(np.sqrt(g[i] @ Sinv[i] @ g[i]) for i in range(n_act))
    assert nu <= 1e-6
    rng = np.random.default_rng(7)
    for d in _tangent_perturbations(r["C"], sp.S, mask, rng, 8):
        e = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, r["C"] + 1e-3 * d, support=support,
                            gradient=False)["energy"]
        assert e >= r["energy"] - 1e-12


def test_plbfgs_agrees_with_scipy_bfgs_on_split_supports():
    n_act, n_bf, n_aux = 2, 10, 30
    sp, topo, support = _case("FULL", n_act, 2, n_bf, n_aux, 1, True)
    r = plbfgs(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, support=support, gradient_tolerance=1e-7,
               energy_tolerance=1e-13)
    mask = support.astype(bool)

    def f(x):
        Cm = np.zeros((n_act, n_bf))
        Cm[mask] = x
        o = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, Cm, support=support)
        return o["energy"], o["grad"][mask]

    res = minimize(f, np.where(mask, sp.C, 0.0)[mask], jac=True, method="BFGS", options=dict(gtol=1e-9))
    assert abs(res.fun - r["energy"]) <= 1e-10


def test_plbfgs_options_and_limits():
    sp, topo, support = _case("COV", 4, 4, 16, 48, 2, False)
    r = plbfgs(16, 48, topo, sp.S, sp.h, sp.B, sp.C, max_iterations=3)
    assert r["iterations"] == 3 and r["termination"] == "max_iterations" and not r["converged"]
    r0 = plbfgs(16, 48, topo, sp.S, sp.h, sp.B, sp.C, gradient_tolerance=1e9)
    assert r0["termination"] == "plbfgs_initial_tolerance" and r0["iterations"] == 0
    r1 = plbfgs(16, 48, topo, sp.S, sp.h, sp.B, sp.C, max_iterations=12, history_size=0)
    assert r1["iterations"] == 12  # steepest descent in the S metric still descends
    assert np.all(np.diff(r1["trace"]) < 0)
