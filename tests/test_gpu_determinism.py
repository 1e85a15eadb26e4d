"""Bitwise reproducibility of energies and gradients (needs GPU 0).

The backward kernels accumulate adjoints in a fixed order (per-wavefront copies summed in wavefront
order, wavefronts taking turns on shared accumulators, one wavefront per block where atomics stay
shared), and the deferred and rank-deficient pair lists are built in pair order. Repeating an
evaluation with the same launch therefore gives the same bits, and a pinned launch (single
problems, the optimizer) gives every slot the same bits whatever it is batched with, so a
full-convergence run follows the same iterates alone and in a batch.
"""
import numpy as np
import pytest

import xmvb_amd as vb

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("kw,n_act,n_bf,n_aux,seed,limit", [
    ("COV", 10, 120, 360, 0, None),   # medium: lane kernels, private wavefront accumulators
    ("COV", 10, 60, 120, 1, 1.0e3),   # many pairs deferred to the lane-group kernel
    ("COV", 6, 30, 90, 2, 0.0),       # every pair on the lane-group kernel
    ("COV", 12, 48, 96, 3, None),     # 6x6 blocks: shared W taken in turns, tiles flushed in turns
    ("FULL", 6, 24, 60, 4, None),
])
def test_repeated_evaluations_are_bitwise_equal(kw, n_act, n_bf, n_aux, seed, limit):
    print(f"case repeat {kw} n_act={n_act} n_bf={n_bf} limit={limit}", flush=True)
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_act, 1)
    ctx = vb.Context(0)
    if limit is not None:
        ctx.set_pair_condition_limit(limit)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    runs = [prob.evaluate(sp.C) for _ in range(3)]
    for e, g in runs[1:]:
        assert e == runs[0][0]
        assert _same(g, runs[0][1])
    prob.close()
    ctx.close()


def test_singular_pairs_are_bitwise_reproducible():
    from problems import fragment_problem

    print("case repeat singular fragments", flush=True)
    n_bf, S, h, B, Cm, sup = fragment_problem(3, 12, 5, 5, 40)
    topo = vb.Topology.build("COV", 10, 10, 1)
    ctx = vb.Context(0)
    prob = vb.Problem(ctx, n_bf, 40, topo, S, h, B)
    runs = [prob.evaluate(Cm, support=sup) for _ in range(3)]
    assert prob.singular_pairs() > 0
    for e, g in runs[1:]:
        assert e == runs[0][0]
        assert _same(g, runs[0][1])
    prob.close()
    ctx.close()


def test_batched_launch_is_bitwise_reproducible():
    topo = vb.Topology.build("COV", 10, 10, 1)
    K = 8
    ctx = vb.Context(0)
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(50 + s, 120, 10, 360) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print("case repeat batched medium x8", flush=True)
    e0, g0 = pool.evaluate(list(range(K)), [sp.C for sp in sps])
    e1, g1 = pool.evaluate(list(range(K)), [sp.C for sp in sps])
    assert _same(e0, e1)
    for a, b in zip(g0, g1):
        assert _same(a, b)
    pool.close()
    ctx.close()


@pytest.mark.parametrize("n_act,n_bf,n_aux,iterations", [(10, 120, 360, 160), (12, 48, 96, 150)])
def test_optimizer_iterates_alone_equal_batched(n_act, n_bf, n_aux, iterations):
    """P-LBFGS from the same start, alone (twice) and in lockstep with three other problems: the
    same energy after every accepted iteration, bit for bit, and the same counts."""
    topo = vb.Topology.build("COV", n_act, n_act, 1)
    ctx = vb.Context(0)
    sps = [vb.synthetic_problem(900 + k, n_bf, n_act, n_aux, scan_r=1.0 + 0.5 * k) for k in range(4)]
    print(f"case optimizer alone vs batched n_act={n_act} n_bf={n_bf} {iterations} iterations", flush=True)
    prob = vb.Problem.from_synthetic(ctx, sps[0], topo)
    alone = [prob.optimize(sps[0].C, max_iterations=iterations) for _ in range(2)]
    prob.close()
    assert alone[0]["iterations"] == alone[1]["iterations"]
    assert _same(alone[0]["trace"], alone[1]["trace"])
    pool = vb.Pool(ctx, topo, capacity=4)
    for k, sp in enumerate(sps):
        pool.load_synthetic(k, sp)
    batched = pool.optimize([0, 1, 2, 3], [sp.C for sp in sps], max_iterations=iterations)
    pool.close()
    b = batched[0]
    assert b["iterations"] == alone[0]["iterations"] >= min(iterations, 100)
    assert b["energy_evaluations"] == alone[0]["energy_evaluations"]
    assert b["energy"] == alone[0]["energy"]
    assert _same(b["C"], alone[0]["C"])
    ctx.close()


def test_seven_electron_blocks_reproduce_to_rounding():
    """n_act = 14 with 7-electron blocks: the packed-kernel adjoint has no room in LDS for per-
    wavefront copies, so it keeps shared atomics; the energy is still the same bits (the forward
    has no accumulation) and the gradient agrees to rounding."""
    print("case repeat 7x7 blocks n_act=14", flush=True)
    sp = vb.synthetic_problem(5, 28, 14, 56)
    topo = vb.Topology.build("COV", 14, 14, 1)
    ctx = vb.Context(0)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    runs = [prob.evaluate(sp.C) for _ in range(2)]
    assert runs[1][0] == runs[0][0]
    assert float(np.abs(runs[1][1] - runs[0][1]).max()) <= 1e-12 * max(1.0, float(np.abs(runs[0][1]).max()))
    prob.close()
    ctx.close()
