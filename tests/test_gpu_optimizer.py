"""Device P-LBFGS (vbe_optimize) against the CPU restatement oracle/plbfgs.py on the same seeded
problems and options: the same accepted iterates (energy after every iteration), the same stopping
reason and point, and at the medium size the energy of the returned orbitals against the oracle."""
import numpy as np
import pytest

import oracle
from oracle.plbfgs import plbfgs
import xmvb_amd as vb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vb.Context(0)
    yield c
    c.close()


def _case(kw, n_act, n_el, n_bf, n_aux, seed, split_support):
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_el, 1)
    support = None
    if split_support:
        support = np.zeros((n_act, n_bf), np.uint8)
        w = n_bf // n_act
        for i in range(n_act):
            support[i, i * w:(i + 1) * w] = 1
    return sp, topo, support


@pytest.mark.parametrize("kw,n_act,n_el,n_bf,n_aux,seed,split,opts", [
    ("FULL", 2, 2, 10, 30, 1, True, dict(gradient_tolerance=1e-6, energy_tolerance=1e-12)),
    ("COV", 4, 4, 16, 48, 2, False, dict(gradient_tolerance=1e-6, energy_tolerance=1e-12, max_iterations=40)),
    ("COV", 6, 6, 30, 90, 4, False, dict(max_iterations=25)),
    ("COV", 4, 4, 16, 48, 3, False, dict(history_size=0, max_iterations=15)),
])
def test_optimizer_follows_oracle_iterates(ctx, kw, n_act, n_el, n_bf, n_aux, seed, split, opts):
    print(f"case optimize {kw} n_act={n_act} n_bf={n_bf} seed={seed} {opts}", flush=True)
    sp, topo, support = _case(kw, n_act, n_el, n_bf, n_aux, seed, split)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    r = prob.optimize(sp.C, support=support, **opts)
    ref = plbfgs(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, support=support, **opts)
    print(f"device {r['iterations']} it E={r['energy']:.12f} {r['termination']}; "
          f"oracle {ref['iterations']} it E={ref['energy']:.12f} {ref['termination']}", flush=True)
    assert r["termination"] == ref["termination"]
    assert r["iterations"] == ref["iterations"]
    assert r["energy_evaluations"] == ref["energy_evaluations"]
    np.testing.assert_allclose(r["trace"], ref["trace"], rtol=0, atol=1e-9)
    assert abs(r["energy"] - ref["energy"]) <= 1e-9
    assert np.abs(r["C"] - ref["C"]).max() <= 1e-6
    assert abs(r["gradient_norm"] - ref["gradient_norm"]) <= 1e-6 + 1e-6 * ref["gradient_norm"]
    prob.close()


def test_optimizer_medium_config(ctx):
    """n_bf=120, n_aux=360, n_act=10: the first accepted iterates against the restatement, then a
    longer run whose returned point the oracle re-evaluates."""
    print("case optimize medium n_act=10 n_bf=120", flush=True)
    sp, topo, _ = _case("COV", 10, 10, 120, 360, 0, False)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    r = prob.optimize(sp.C, max_iterations=6)
    ref = plbfgs(120, 360, topo, sp.S, sp.h, sp.B, sp.C, max_iterations=6)
    np.testing.assert_allclose(r["trace"], ref["trace"], rtol=0, atol=1e-9)
    long = prob.optimize(sp.C, max_iterations=60)
    print(f"medium: {long['iterations']} it E={long['energy']:.10f} |r|={long['gradient_norm']:.2e} "
          f"{long['termination']} evals={long['energy_evaluations']}/{long['gradient_evaluations']}", flush=True)
    assert np.all(np.diff(long["trace"]) <= 1e-12)
    chk = oracle.evaluate(120, 360, topo, sp.S, sp.h, sp.B, long["C"])
    assert abs(chk["energy"] - long["energy"]) <= 1e-8
    assert np.abs(chk["grad"] - long["grad"]).max() <= 1e-7
    prob.close()


def test_optimizer_rejects_bad_options(ctx):
    sp, topo, _ = _case("COV", 4, 4, 16, 48, 2, False)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    with pytest.raises(vb.VbeError):
        prob.optimize(sp.C, armijo_constant=1.5)
    with pytest.raises(vb.VbeError):
        prob.optimize(sp.C, history_size=-1)
    r = prob.optimize(sp.C, gradient_tolerance=1e9)
    assert r["termination"] == "plbfgs_initial_tolerance" and r["iterations"] == 0
    prob.close()


def test_pool_optimize_follows_each_problem_alone(ctx):
    """Several slots (different seeds and basis sizes) converged together in lockstep reach the
    point each reaches alone, with the same iteration count and stopping reason."""
    print("case pool optimize COV n_act=4 slots seeds 2,3,5", flush=True)
    topo = vb.Topology.build("COV", 4, 4, 1)
    cases = [(2, 16, 48), (3, 20, 40), (5, 16, 32)]
    pool = vb.Pool(ctx, topo, capacity=4)
    sps = []
    for slot, (seed, n_bf, n_aux) in enumerate(cases):
        sp = vb.synthetic_problem(seed, n_bf, 4, n_aux)
        pool.load_synthetic(slot, sp)
        sps.append(sp)
    opts = dict(gradient_tolerance=1e-6, energy_tolerance=1e-12, max_iterations=30)
    res = pool.optimize([0, 1, 2], [sp.C for sp in sps], **opts)
    for slot, (sp, r) in enumerate(zip(sps, res)):
        prob = vb.Problem.from_synthetic(ctx, sp, topo)
        alone = prob.optimize(sp.C, **opts)
        prob.close()
        print(f"slot {slot}: pool {r['iterations']} it E={r['energy']:.12f} {r['termination']}; alone "
              f"{alone['iterations']} it E={alone['energy']:.12f}", flush=True)
        assert r["termination"] == alone["termination"]
        assert r["iterations"] == alone["iterations"]
        assert r["energy_evaluations"] == alone["energy_evaluations"]
        assert abs(r["energy"] - alone["energy"]) <= 1e-9
        assert np.abs(r["C"] - alone["C"]).max() <= 1e-6
        assert np.abs(r["grad"] - alone["grad"]).max() <= 1e-6
    pool.close()


def test_pool_optimize_medium_batch(ctx):
    """Eight medium problems (n_bf=120, n_act=10) converging together, checked against the oracle's
    energy and gradient at each returned point."""
    print("case pool optimize medium x8", flush=True)
    topo = vb.Topology.build("COV", 10, 10, 1)
    pool = vb.Pool(ctx, topo, capacity=8)
    sps = [vb.synthetic_problem(s, 120, 10, 360) for s in range(8)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    res = pool.optimize(list(range(8)), [sp.C for sp in sps], max_iterations=25)
    for s in (0, 5):
        chk = oracle.evaluate(120, 360, topo, sps[s].S, sps[s].h, sps[s].B, res[s]["C"])
        assert res[s]["iterations"] == 25
        assert abs(chk["energy"] - res[s]["energy"]) <= 1e-8
        assert np.abs(chk["grad"] - res[s]["grad"]).max() <= 1e-7
    pool.close()


def test_pool_optimize_single_slot(ctx):
    """A pool batch of one slot evaluates its trials with gradients (no retained single-problem
    forward to complete) and still follows the iterates of the problem alone."""
    topo = vb.Topology.build("COV", 4, 4, 1)
    sp = vb.synthetic_problem(2, 16, 4, 48)
    pool = vb.Pool(ctx, topo, capacity=2)
    pool.load_synthetic(1, sp)
    r = pool.optimize([1], [sp.C], max_iterations=20)[0]
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    alone = prob.optimize(sp.C, max_iterations=20)
    prob.close()
    assert r["iterations"] == alone["iterations"] == 20
    assert abs(r["energy"] - alone["energy"]) <= 1e-9
    pool.close()
