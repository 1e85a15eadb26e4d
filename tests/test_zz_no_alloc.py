"""No-allocation entry point vbe_eval_device_noalloc against vbe_eval_device (needs GPU 0).

Both entry points run the same kernels on the same resident problem; the no-allocation one carves
its workspace from the library's static device scratch. Its energy and gradient must equal
vbe_eval_device's (|dE| <= 1e-12 Eh, gradient 1e-12) and the oracle's (1e-8 Eh, 1e-7) on two calls
in a row, the second with a smaller problem, and the calls must leave the device's free memory
unchanged.
"""
import numpy as np
import pytest

import oracle
import xmvb_amd as vb

pytestmark = pytest.mark.gpu


def _eval_both(ctx, kw, n_act, n_el, n_bf, n_aux, seed):
    import torch

    print(f"case {kw} n_act={n_act} n_bf={n_bf} n_aux={n_aux} seed={seed}", flush=True)
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_el, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    c = torch.from_numpy(sp.C.copy()).to("cuda:0")
    g_ref = torch.zeros_like(c)
    g_na = torch.full_like(c, float("nan"))
    torch.cuda.synchronize()
    e_ref = prob.evaluate_device(c.data_ptr(), g_ref.data_ptr())
    free0 = torch.cuda.mem_get_info(0)[0]
    e_na = prob.evaluate_device_noalloc(c.data_ptr(), g_na.data_ptr())
    free1 = torch.cuda.mem_get_info(0)[0]
    out = dict(sp=sp, topo=topo, e_ref=e_ref, e_na=e_na, g_ref=g_ref.cpu().numpy(), g_na=g_na.cpu().numpy(),
               free=(free0, free1))
    prob.close()
    return out


def _check(r, n_bf, n_aux):
    assert abs(r["e_na"] - r["e_ref"]) <= 1e-12
    assert np.all(np.isfinite(r["g_na"]))
    assert float(np.abs(r["g_na"] - r["g_ref"]).max()) <= 1e-12
    sp = r["sp"]
    ref = oracle.evaluate(n_bf, n_aux, r["topo"], sp.S, sp.h, sp.B, sp.C)
    assert abs(r["e_na"] - ref["energy"]) <= 1e-8
    assert float(np.abs(r["g_na"] - ref["grad"]).max()) <= 1e-7


def test_noalloc_matches_eval_device_two_calls():
    ctx = vb.Context(0)
    # first call: the medium size; second call: a smaller problem on the same scratch
    first = _eval_both(ctx, "COV", 10, 10, 120, 360, 0)
    _check(first, 120, 360)
    second = _eval_both(ctx, "COV", 6, 6, 48, 144, 3)
    _check(second, 48, 144)
    # after the first call, no further call changes the device's free memory
    assert second["free"][0] == second["free"][1]
    ctx.close()


def test_noalloc_rejects_workspace_above_the_scratch_maximum():
    import torch

    ctx = vb.Context(0)
    # 14 orbitals, 429 structures, 3432 determinants: the determinant H/S alone need 188 MB
    n_bf, n_aux = 28, 56
    sp = vb.synthetic_problem(5, n_bf, 14, n_aux)
    topo = vb.Topology.build("COV", 14, 14, 1)
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    c = torch.from_numpy(sp.C.copy()).to("cuda:0")
    g = torch.zeros_like(c)
    torch.cuda.synchronize()
    with pytest.raises(vb.VbeError) as err:
        prob.evaluate_device_noalloc(c.data_ptr(), g.data_ptr())
    assert err.value.status == 6 and "scratch maximum" in str(err.value)
    prob.close()
    ctx.close()
