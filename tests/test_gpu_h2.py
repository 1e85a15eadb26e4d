"""The H2 deck end to end on the device: R = 0.74 angstrom, 6-31G** (10 functions), two VB orbitals
each on its own atom's functions, STR=FULL (1 covalent + 2 ionic structures), exact two-electron
integrals as pivoted-Cholesky factors, P-LBFGS from the atomic core guess. The device energy is
compared with the FP64 oracle and its P-LBFGS restatement on the same integrals (1e-10 Eh)."""
import numpy as np
import pytest

import oracle
from oracle.plbfgs import plbfgs
import xmvb_amd as vb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deck():
    return vb.h2_deck()


def _problem(ctx, d):
    return vb.Problem(ctx, d["n_bf"], d["n_aux"], d["topo"], d["S"], d["h"], d["B"], e_nuc=d["e_nuc"])


def test_h2_single_evaluation(deck):
    print("case h2 evaluation", flush=True)
    ctx = vb.Context(0)
    prob = _problem(ctx, deck)
    e, g = prob.evaluate(deck["C0"], support=deck["support"])
    ref = oracle.evaluate(deck["n_bf"], deck["n_aux"], deck["topo"], deck["S"], deck["h"], deck["B"], deck["C0"],
                          support=deck["support"], e_nuc=deck["e_nuc"])
    assert abs(e - ref["energy"]) <= 1e-12
    assert np.abs(g - ref["grad"]).max() <= 1e-11
    assert np.all(g[deck["support"] == 0] == 0.0)
    prob.close()
    ctx.close()


@pytest.mark.parametrize("opts", [dict(), dict(gradient_tolerance=1e-7, energy_tolerance=1e-13)])
def test_h2_plbfgs_matches_oracle(deck, opts):
    print(f"case h2 plbfgs {opts}", flush=True)
    ctx = vb.Context(0)
    prob = _problem(ctx, deck)
    r = prob.optimize(deck["C0"], support=deck["support"], **opts)
    ref = plbfgs(deck["n_bf"], deck["n_aux"], deck["topo"], deck["S"], deck["h"], deck["B"], deck["C0"],
                 support=deck["support"], **opts)
    e_ref = ref["energy"] + deck["e_nuc"]
    print(f"device E={r['energy']:.12f} ({r['iterations']} it, {r['termination']}); oracle E={e_ref:.12f} "
          f"({ref['iterations']} it)", flush=True)
    assert r["converged"] and r["termination"] == ref["termination"]
    assert r["iterations"] == ref["iterations"]
    assert abs(r["energy"] - e_ref) <= 1e-10
    assert np.abs(r["C"] - ref["C"]).max() <= 1e-8
    prob.close()
    ctx.close()
