"""The batching workload's shape: independent synthetic problems with uneven sizes, n_act drawn from
{4, 6, 8, 10, 12} (n_electrons = n_act, singlet, all covalent Rumer structures) and n_bf uniform in
[40, 200] with n_aux = 3 n_bf, bucketed by active space into pools (one topology per pool, any n_bf
per slot) and evaluated one launch sequence per bucket. Every problem is checked against the FP64
oracle. A sample of 20 problems (4 per bucket) keeps the oracle's share of the run to seconds."""
import numpy as np
import pytest

import oracle
import xmvb_amd as vb

pytestmark = pytest.mark.gpu


def test_mixed_size_buckets_match_oracle():
    rng = np.random.default_rng(20260427)
    ctx = vb.Context(0)
    worst_e, worst_g = 0.0, 0.0
    for n_act in (4, 6, 8, 10, 12):
        topo = vb.Topology.build("COV", n_act, n_act, 1)
        sizes = [int(x) for x in rng.integers(40, 201, size=4)]
        print(f"case mixed bucket n_act={n_act} n_bf={sizes}", flush=True)
        pool = vb.Pool(ctx, topo, capacity=len(sizes))
        sps = []
        for slot, n_bf in enumerate(sizes):
            sp = vb.synthetic_problem(1000 * n_act + slot, n_bf, n_act, 3 * n_bf)
            pool.load_synthetic(slot, sp)
            sps.append(sp)
        e, g = pool.evaluate(list(range(len(sizes))), [sp.C for sp in sps])
        for slot, sp in enumerate(sps):
            ref = oracle.evaluate(sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, sp.C)
            de, dg = abs(e[slot] - ref["energy"]), np.abs(g[slot] - ref["grad"]).max()
            worst_e, worst_g = max(worst_e, de), max(worst_g, dg)
            assert de <= 1e-8, (n_act, sp.n_bf, de)
            assert dg <= 1e-7, (n_act, sp.n_bf, dg)
        pool.close()
    print(f"worst |dE| {worst_e:.2e}, worst gradient error {worst_g:.2e}", flush=True)
    ctx.close()
