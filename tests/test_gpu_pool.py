"""Batched pools (many resident problems per launch) and submit/wait serving through the C ABI.

Each slot's energy and gradient must equal the oracle's (|dE| <= 1e-8, grad 1e-7) and must be
the same evaluated alone and in a mixed batch (tolerance 1e-12: the per-slot arithmetic is the
same, only the block partition of the adjoint sums can change)."""
import numpy as np
import pytest

import oracle
import xmvb_amd as vb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vb.Context(0)
    yield c
    c.close()


def test_heterogeneous_slots_match_oracle(ctx):
    topo = vb.Topology.build("COV", 6, 6, 1)
    dims = [(20, 40), (31, 62), (45, 100), (24, 24)]
    pool = vb.Pool(ctx, topo, capacity=len(dims))
    sps = []
    for s, (n_bf, n_aux) in enumerate(dims):
        sp = vb.synthetic_problem(100 + s, n_bf, 6, n_aux)
        pool.load_synthetic(s, sp, e_nuc=0.25 * s)
        sps.append(sp)
    print("case heterogeneous pool of 4 slots", flush=True)
    e, g = pool.evaluate([0, 1, 2, 3], [sp.C for sp in sps])
    for s, sp in enumerate(sps):
        ref = oracle.evaluate(sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, sp.C, e_nuc=0.25 * s)
        assert abs(e[s] - ref["energy"]) <= 1e-8
        assert np.abs(g[s] - ref["grad"]).max() <= 1e-7
    pool.close()


def test_tiled_and_flat_factor_layouts_in_one_launch(ctx):
    """Slots with n_bf up to 224 keep their RI factors aux-tiled (one-read transform), larger ones in
    the flat layout; a launch over both runs both transform kernels."""
    topo = vb.Topology.build("COV", 4, 4, 1)
    dims = [(40, 96), (240, 72), (33, 41)]
    pool = vb.Pool(ctx, topo, capacity=len(dims))
    sps = []
    for s, (n_bf, n_aux) in enumerate(dims):
        sp = vb.synthetic_problem(300 + s, n_bf, 4, n_aux)
        pool.load_synthetic(s, sp)
        sps.append(sp)
    print("case tiled + flat RI layouts in one pool launch", flush=True)
    e, g = pool.evaluate([0, 1, 2], [sp.C for sp in sps])
    for s, sp in enumerate(sps):
        ref = oracle.evaluate(sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, sp.C)
        assert abs(e[s] - ref["energy"]) <= 1e-8
        assert np.abs(g[s] - ref["grad"]).max() <= 1e-7
    pool.close()


def test_alone_and_in_the_mix_agree(ctx):
    topo = vb.Topology.build("COV", 8, 8, 1)
    K = 6
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(200 + s, 40, 8, 120) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print("case alone vs mix, 6 slots", flush=True)
    alone = [pool.evaluate([s], [sps[s].C]) for s in range(K)]
    order = [3, 0, 5, 1, 4, 2]
    e_mix, g_mix = pool.evaluate(order, [sps[s].C for s in order])
    for i, s in enumerate(order):
        assert abs(e_mix[i] - alone[s][0][0]) <= 1e-12
        assert np.abs(g_mix[i] - alone[s][1][0]).max() <= 1e-12
    # two submissions in flight on disjoint slots
    t1 = pool.submit([0, 1, 2], [sps[s].C for s in (0, 1, 2)])
    t2 = pool.submit([3, 4, 5], [sps[s].C for s in (3, 4, 5)])
    e2, g2 = pool.wait(t2)
    e1, g1 = pool.wait(t1)
    for i, s in enumerate((0, 1, 2)):
        assert abs(e1[i] - alone[s][0][0]) <= 1e-12
    for i, s in enumerate((3, 4, 5)):
        assert abs(e2[i] - alone[s][0][0]) <= 1e-12
        assert np.abs(g2[i] - alone[s][1][0]).max() <= 1e-12
    ref = oracle.evaluate(40, 120, topo, sps[4].S, sps[4].h, sps[4].B, sps[4].C)
    assert abs(e2[1] - ref["energy"]) <= 1e-8
    pool.close()


def test_slot_conflict_is_rejected(ctx):
    topo = vb.Topology.build("COV", 4, 4, 1)
    pool = vb.Pool(ctx, topo, capacity=2)
    sp = vb.synthetic_problem(7, 16, 4, 32)
    pool.load_synthetic(0, sp)
    pool.load_synthetic(1, sp)
    t = pool.submit([0], [sp.C])
    with pytest.raises(vb.VbeError):
        pool.submit([0, 1], [sp.C, sp.C])  # slot 0 still pending
    # a synchronous batched evaluation of a slot that the pending submission holds is rejected too
    import torch

    c_dev = torch.from_numpy(np.concatenate([sp.C.ravel(), sp.C.ravel()])).to("cuda:0")
    g_dev = torch.zeros_like(c_dev)
    torch.cuda.synchronize()
    with pytest.raises(vb.VbeError) as err:
        pool.evaluate_device([0], c_dev.data_ptr(), g_dev.data_ptr())
    assert err.value.status == 1 and "pending submission" in str(err.value)
    with pytest.raises(vb.VbeError):
        pool.load_synthetic(0, sp)
    e_free = pool.evaluate_device([1], c_dev.data_ptr(), g_dev.data_ptr())  # the other slot is free
    assert np.isfinite(e_free[0])
    pool.wait(t)
    with pytest.raises(vb.VbeError):
        pool.evaluate([0, 0], [sp.C, sp.C])  # duplicate slot in one launch
    e, _ = pool.evaluate([1], [sp.C])
    assert np.isfinite(e[0])
    pool.close()


def test_medium_batch_matches_single_problem(ctx):
    topo = vb.Topology.build("COV", 10, 10, 1)
    K = 8
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(s, 120, 10, 360) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print("case medium batch of 8", flush=True)
    e, g = pool.evaluate(list(range(K)), [sp.C for sp in sps])
    prob = vb.Problem.from_synthetic(ctx, sps[0], topo)
    e0, g0 = prob.evaluate(sps[0].C)
    assert abs(e[0] - e0) <= 1e-12 and np.abs(g[0] - g0).max() <= 1e-12
    ref = oracle.evaluate(120, 360, topo, sps[5].S, sps[5].h, sps[5].B, sps[5].C)
    assert abs(e[5] - ref["energy"]) <= 1e-8 and np.abs(g[5] - ref["grad"]).max() <= 1e-7
    prob.close()
    pool.close()


def test_autotune_keeps_results(ctx):
    topo = vb.Topology.build("COV", 8, 8, 1)
    K = 8
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(300 + s, 40, 8, 120) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    slots = list(range(K))
    e0, g0 = pool.evaluate(slots, [sp.C for sp in sps])
    print("case autotune 8 slots", flush=True)
    tuned = pool.autotune(slots, [sp.C for sp in sps], repeats=1)
    assert tuned["fwd_target"] in tuned["table"] and tuned["bwd_target"] in tuned["table"]
    assert all(v["fwd_ms"] > 0.0 and v["bwd_ms"] > 0.0 for v in tuned["table"].values())
    e1, g1 = pool.evaluate(slots, [sp.C for sp in sps])
    for s in range(K):
        assert abs(e1[s] - e0[s]) <= 1e-10
        assert float(np.abs(g1[s] - g0[s]).max()) <= 1e-10
    pool.close()


def test_large_structure_space_batched_matches_alone(ctx):
    """132 structures: the eigensolve runs from global memory, per step over many workgroups when
    a launch holds few problems and one workgroup per problem in larger batches. Both forms must
    give the same energies and gradients (the tridiagonalisation's sums are ordered differently,
    hence the 1e-9 / 1e-8 tolerances)."""
    topo = vb.Topology.build("COV", 12, 12, 1)
    K = 10
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(300 + s, 48, 12, 96) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print("case 132 structures, 10 slots alone vs batched", flush=True)
    e_all, g_all = pool.evaluate(list(range(K)), [sp.C for sp in sps])
    for s in (0, 4, 9):
        e1, g1 = pool.evaluate([s], [sps[s].C])
        assert abs(e_all[s] - e1[0]) <= 1e-9
        assert np.abs(g_all[s] - g1[0]).max() <= 1e-8
    pool.close()


@pytest.mark.parametrize("n_act,n_bf", [(10, 120), (12, 60), (14, 40)])
def test_batched_tail_kernels_match_single(ctx, n_act, n_bf):
    """More than 16 problems per launch switch the Gram product of the active factors and the
    orbital-gradient contraction to their batched (LDS-tiled) kernels; their results must equal
    the one-problem kernels' up to summation order (n_act 12 and 14 take the q-chunked gradient
    staging)."""
    topo = vb.Topology.build("COV", n_act, n_act, 1) if n_act <= 12 else vb.Topology.build("COV", n_act, 2, 1)
    K = 20
    pool = vb.Pool(ctx, topo, capacity=K)
    sps = [vb.synthetic_problem(400 + s, n_bf + (s % 3), n_act, 3 * n_bf) for s in range(K)]
    for s, sp in enumerate(sps):
        pool.load_synthetic(s, sp)
    print(f"case batched tail n_act={n_act} K={K}", flush=True)
    e_all, g_all = pool.evaluate(list(range(K)), [sp.C for sp in sps])
    for s in (0, 7, 19):
        e1, g1 = pool.evaluate([s], [sps[s].C])
        assert abs(e_all[s] - e1[0]) <= 1e-10 * max(1.0, abs(e1[0]))
        assert np.abs(g_all[s] - g1[0]).max() <= 1e-9
    pool.close()
