"""Davidson structure eigensolver (EIGENSOLVER=DAVIDSON) against the FP64 oracle (dense GEP) and
against the engine's dense path, through the C ABI.

With the reference's default accuracy (energy 1e-7, relative residual 1e-3, CI stationarity
||(H - E S) c|| / ||c|| <= 1e-7) the energy error is second order in the residual, so |dE| <= 1e-8
Eh holds; the eigenvector, and with it the gradient, is only first-order accurate (~1e-7 / gap).
With tightened tolerances the gradient meets the dense path's 1e-7 bar."""
import numpy as np
import pytest

import oracle
import xmvb_amd as vb

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-12, residual_tolerance=1e-10, stationarity_tolerance=1e-11)


@pytest.mark.parametrize("kw,n_act,n_el,n_bf,n_aux,seed", [
    ("FULL", 2, 2, 10, 30, 1),     # 3 structures
    ("COV", 4, 4, 16, 48, 2),      # 2 structures
    ("COV", 6, 6, 30, 90, 4),      # 5 structures
    ("COV", 8, 8, 40, 120, 6),     # 14 structures
    ("FULL", 6, 6, 24, 60, 5),     # 175 structures: restarts of the 32-vector basis
])
@pytest.mark.parametrize("tight", [False, True])
def test_davidson_matches_oracle(kw, n_act, n_el, n_bf, n_aux, seed, tight):
    print(f"case davidson {kw} n_act={n_act} n_bf={n_bf} seed={seed} tight={tight}", flush=True)
    ctx = vb.Context(0)
    ctx.set_eigensolver(1, **(TIGHT if tight else {}))
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_el, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    e, g = prob.evaluate(sp.C)
    ref = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C)
    de, dg = abs(e - ref["energy"]), np.abs(g - ref["grad"]).max()
    print(f"n_str={topo.n_structures} dE={de:.2e} dg={dg:.2e}", flush=True)
    assert de <= (1e-10 if tight else 1e-8)
    assert dg <= (1e-7 if tight else 1e-4)
    # S-normalised eigenvector, same ray as the dense solution
    mid = prob.intermediates()
    c, cref = mid["coef"], ref["coef"]
    assert abs(c @ mid["S"] @ c - 1.0) <= 1e-10
    if tight:
        assert np.abs(np.sign(c @ cref) * c - cref).max() <= 1e-7
    # a second evaluation starts from the recycled root and agrees
    e2, g2 = prob.evaluate(sp.C)
    assert abs(e2 - e) <= (1e-11 if tight else 1e-8)
    prob.close()
    ctx.close()


def test_davidson_budget_exhaustion_is_reported():
    """132 covalent structures with an ill-conditioned metric (cond S ~ 8e4): the Jacobi-preconditioned
    iteration does not reach the 1e-7 Eh stationarity bound within the reference's 2 n iterations,
    which the reference reports as an error (davidson_core.hpp:722-729); so does the engine."""
    ctx = vb.Context(0)
    ctx.set_eigensolver(1)
    sp = vb.synthetic_problem(13, 48, 12, 96)
    topo = vb.Topology.build("COV", 12, 12, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    with pytest.raises(vb.VbeError) as err:
        prob.evaluate(sp.C)
    assert err.value.status == 2 and "Davidson" in str(err.value)
    # the problem stays usable: the dense solver evaluates it
    prob.eigensolver = 0
    e, _ = prob.evaluate(sp.C)
    ref = oracle.evaluate(48, 96, topo, sp.S, sp.h, sp.B, sp.C, gradient=False)
    assert abs(e - ref["energy"]) <= 1e-8
    prob.close()
    ctx.close()


def test_davidson_per_call_selection_and_dense_agree():
    ctx = vb.Context(0)
    sp = vb.synthetic_problem(5, 24, 6, 60)
    topo = vb.Topology.build("FULL", 6, 6, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    e_dense, g_dense = prob.evaluate(sp.C)
    ctx.set_eigensolver(1, **TIGHT)
    prob.eigensolver = 1
    e_dav, g_dav = prob.evaluate(sp.C)
    assert abs(e_dav - e_dense) <= 1e-10
    assert np.abs(g_dav - g_dense).max() <= 1e-7
    prob.close()
    ctx.close()


def test_davidson_pool_matches_single():
    ctx = vb.Context(0)
    ctx.set_eigensolver(1, **TIGHT)
    topo = vb.Topology.build("COV", 8, 8, 1)
    pool = vb.Pool(ctx, topo, capacity=3)
    sps = [vb.synthetic_problem(s, 40, 8, 120) for s in (6, 7, 8)]
    for k, sp in enumerate(sps):
        pool.load_synthetic(k, sp)
    e, g = pool.evaluate([0, 1, 2], [sp.C for sp in sps])
    for k, sp in enumerate(sps):
        ref = oracle.evaluate(40, 120, topo, sp.S, sp.h, sp.B, sp.C)
        assert abs(e[k] - ref["energy"]) <= 1e-10
        assert np.abs(g[k] - ref["grad"]).max() <= 1e-7
    pool.close()
    ctx.close()
