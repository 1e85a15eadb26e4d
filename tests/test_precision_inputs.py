"""The precision fixture (tests/golden/precision.npz) holds oracle outputs for seeded synthetic
inputs; this checks that the generator still produces exactly those inputs (SHA-256 of S, h, B, C)."""
import hashlib
import os

import numpy as np

import xmvb_amd as vb

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "precision.npz")


def test_precision_fixture_inputs_reproduce():
    gold = np.load(GOLDEN)
    sp = vb.synthetic_problem(int(gold["seed"]), int(gold["n_bf"]), int(gold["n_act"]), int(gold["n_aux"]))
    for f, want in zip(("S", "h", "B", "C"), gold["input_sha256"]):
        arr = getattr(sp, f).ravel(order="F") if f in "Sh" else getattr(sp, f).ravel()
        assert hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest() == str(want), f


def test_precision_fixture_gradient_is_scale_invariant():
    """E is invariant under rescaling any one orbital, so <dE/dc_i, c_i> = 0 for every orbital;
    a dropped normalisation term in the oracle's pullback would break this."""
    gold = np.load(GOLDEN)
    sp = vb.synthetic_problem(int(gold["seed"]), int(gold["n_bf"]), int(gold["n_act"]), int(gold["n_aux"]))
    g = gold["grad"]
    proj = np.einsum("ib,ib->i", g, sp.C)
    assert np.abs(proj).max() <= 1e-9 * max(1.0, np.abs(g).max() * np.abs(sp.C).max() * g.shape[1])
