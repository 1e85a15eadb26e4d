"""State-averaged evaluations (NSTATE > 1, workflow/evaluator.cpp:22-32,65-111) through the C ABI:
selected structure roots with normalised weights, energy E = sum_s w_s lambda_s + E_nuc and the
gradient of that average, against the FP64 oracle (whose state-averaged gradient is pinned by
central differences on the CPU): |dE| <= 1e-8 Eh, gradient max error <= 1e-7, per-state energies
1e-8. Cases cover the lane kernels (one pair per lane), the lane-group kernel (every pair
deferred), rank-deficient pairs, the one-workgroup and the many-workgroup eigensolve, a single
non-lowest state, and batched pools."""
import numpy as np
import pytest

import oracle
import xmvb_amd as vb

pytestmark = pytest.mark.gpu

STATES, WEIGHTS = [0, 1, 2], [0.5, 0.3, 0.2]


def _check(e, g, ref, states_e=None):
    assert abs(e - ref["energy"]) <= 1e-8, (e, ref["energy"])
    assert float(np.abs(g - ref["grad"]).max()) <= 1e-7
    if states_e is not None:
        assert np.abs(states_e - ref["state_energies"]).max() <= 1e-8


@pytest.mark.parametrize("kw,n_act,n_bf,n_aux,seed,limit", [
    ("COV", 6, 30, 90, 4, None),     # 5 structures: one-workgroup eigensolve replaced by the global path
    ("COV", 8, 40, 120, 6, None),    # 14 structures
    ("COV", 10, 60, 120, 3, None),   # 42 structures, lane kernels
    ("COV", 6, 30, 90, 8, 0.0),      # every pair on the lane-group kernel
    ("FULL", 6, 24, 60, 5, None),    # 175 structures
])
def test_state_average_matches_oracle(kw, n_act, n_bf, n_aux, seed, limit):
    print(f"case states {kw} n_act={n_act} n_bf={n_bf} limit={limit}", flush=True)
    ctx = vb.Context(0)
    if limit is not None:
        ctx.set_pair_condition_limit(limit)
    ctx.set_states(STATES, WEIGHTS)
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_act, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo, e_nuc=0.75)
    e, g = prob.evaluate(sp.C)
    ref = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, e_nuc=0.75, states=STATES, weights=WEIGHTS)
    _check(e, g, ref, prob.state_energies())
    prob.close()
    ctx.close()


def test_single_excited_state_and_rank_deficient_pairs():
    from problems import fragment_problem

    n_bf, S, h, B, Cm, sup = fragment_problem(3, 12, 5, 5, 40)
    topo = vb.Topology.build("COV", 10, 10, 1)
    ctx = vb.Context(0)
    for states, weights in (([1], [1.0]), ([2, 0], [1.0, 3.0])):
        print(f"case states {states} fragments (singular pairs)", flush=True)
        ctx.set_states(states, weights)
        prob = vb.Problem(ctx, n_bf, 40, topo, S, h, B)
        e, g = prob.evaluate(Cm, support=sup)
        assert prob.singular_pairs() > 0
        ref = oracle.evaluate(n_bf, 40, topo, S, h, B, Cm, support=sup, states=states, weights=weights)
        _check(e, g, ref, prob.state_energies())
        prob.close()
    ctx.close()


def test_state_average_in_a_pool_and_per_step_eigensolve():
    """132 structures (many-workgroup eigensolve for few problems) and a batched pool of three."""
    ctx = vb.Context(0)
    ctx.set_states(STATES, WEIGHTS)
    topo = vb.Topology.build("COV", 12, 12, 1)
    sps = [vb.synthetic_problem(300 + s, 36, 12, 72) for s in range(3)]
    pool = vb.Pool(ctx, topo, capacity=3)
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print("case states pool of 3, 132 structures", flush=True)
    e, g = pool.evaluate([0, 1, 2], [sp.C for sp in sps])
    for s in (0, 2):
        ref = oracle.evaluate(36, 72, topo, sps[s].S, sps[s].h, sps[s].B, sps[s].C, states=STATES, weights=WEIGHTS)
        _check(e[s], g[s], ref, pool.state_energies(s))
    e1, g1 = pool.evaluate([1], [sps[1].C])
    assert abs(e1[0] - e[1]) <= 1e-10 and float(np.abs(g1[0] - g[1]).max()) <= 1e-9
    pool.close()
    ctx.close()


def test_state_average_optimisation_and_davidson_rejection():
    ctx = vb.Context(0)
    ctx.set_states([0, 1], [1.0, 1.0])
    topo = vb.Topology.build("COV", 6, 6, 1)
    sp = vb.synthetic_problem(21, 24, 6, 48)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    print("case states P-LBFGS", flush=True)
    res = prob.optimize(sp.C, max_iterations=30)
    tr = np.asarray(res["trace"])
    assert res["iterations"] > 0 and np.all(np.diff(tr) <= 1e-12)
    e, _ = prob.evaluate(res["C"])
    assert abs(e - res["energy"]) <= 1e-10
    ref = oracle.evaluate(24, 48, topo, sp.S, sp.h, sp.B, res["C"], states=[0, 1], weights=[1.0, 1.0],
                          gradient=False)
    assert abs(e - ref["energy"]) <= 1e-8
    prob.eigensolver = 1
    with pytest.raises(vb.VbeError) as err:
        prob.evaluate(sp.C)
    assert err.value.status == 5
    prob.close()
    ctx.close()
