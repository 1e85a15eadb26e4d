"""The oracle's state-averaged energy and gradient (selected structure roots with normalised
weights, workflow/evaluator.cpp:22-32,65-111, mcscf/states/weights.cpp:18-33), pinned by what the
mathematics fixes: the energy equals the weighted sum of the generalised eigenvalues of the oracle's
own H and S (scipy), the gradient equals central finite differences of that energy, a one-state
selection reproduces the single-root evaluation, and a zero weight drops its state."""
import numpy as np
import pytest
from scipy.linalg import eigh

import oracle
import xmvb_amd as vb


@pytest.fixture(scope="module")
def case():
    sp = vb.synthetic_problem(7, 14, 4, 28)
    topo = vb.Topology.build("FULL", 4, 4, 1)  # 20 structures
    return sp, topo


def _eval(case, **kw):
    sp, topo = case
    return oracle.evaluate(sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, kw.pop("C", sp.C), e_nuc=0.5, **kw)


def test_state_average_energy_is_the_weighted_eigenvalue_sum(case):
    st, wt = [0, 1, 3], [5.0, 3.0, 2.0]  # normalised to 0.5, 0.3, 0.2
    r = _eval(case, states=st, weights=wt, gradient=False)
    ev = eigh(r["H"], r["S"], eigvals_only=True)
    expect = sum(w / 10.0 * ev[s] for s, w in zip(st, wt)) + 0.5
    assert abs(r["energy"] - expect) <= 1e-11 * max(1.0, abs(expect))
    assert np.allclose(r["state_energies"], ev[st] + 0.5, rtol=0, atol=1e-11)


def test_single_state_selection_reproduces_the_lowest_root(case):
    r0 = _eval(case)
    r1 = _eval(case, states=[0], weights=[2.0])
    assert abs(r1["energy"] - r0["energy"]) <= 1e-13
    assert np.abs(r1["grad"] - r0["grad"]).max() <= 1e-12
    r2 = _eval(case, states=[0, 2], weights=[1.0, 0.0])
    assert abs(r2["energy"] - r0["energy"]) <= 1e-13
    assert np.abs(r2["grad"] - r0["grad"]).max() <= 1e-12


@pytest.mark.parametrize("st,wt", [([0, 1, 3], [0.5, 0.3, 0.2]), ([2], [1.0])])
def test_state_average_gradient_matches_central_differences(case, st, wt):
    sp, _ = case
    g = _eval(case, states=st, weights=wt)["grad"]
    h = 1e-5
    fd = np.zeros_like(sp.C)
    for i in range(sp.C.shape[0]):
        for j in range(sp.C.shape[1]):
            Cp, Cm = sp.C.copy(), sp.C.copy()
            Cp[i, j] += h
            Cm[i, j] -= h
            ep = _eval(case, C=Cp, states=st, weights=wt, gradient=False)["energy"]
            em = _eval(case, C=Cm, states=st, weights=wt, gradient=False)["energy"]
            fd[i, j] = (ep - em) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-7 * max(1.0, np.abs(g).max())


def test_invalid_selections_are_rejected(case):
    with pytest.raises(RuntimeError):
        _eval(case, states=[0, 1], weights=[1.0, -1.0])
    with pytest.raises(RuntimeError):
        _eval(case, states=[25], weights=[1.0])
