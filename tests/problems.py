"""Test problem constructors shared by the oracle and GPU parity tests."""
import numpy as np


def fragment_problem(seed, n_frag_bf, nact_a, nact_b, n_aux):
    """Two AO fragments with zero cross overlap and fragment-localised orbitals: many determinant
    overlap blocks are exactly singular (nullity 1..3)."""
    rng = np.random.default_rng(seed)
    n_bf = 2 * n_frag_bf
    S = np.zeros((n_bf, n_bf))
    for f in range(2):
        a = rng.normal(size=(n_frag_bf, n_frag_bf))
        S[f * n_frag_bf:(f + 1) * n_frag_bf, f * n_frag_bf:(f + 1) * n_frag_bf] = np.eye(n_frag_bf) + 0.1 * a @ a.T / n_frag_bf
    x = rng.normal(scale=np.sqrt(0.5), size=(n_bf, n_bf))
    h = 0.5 * (x + x.T) - 2 * np.eye(n_bf)
    B = np.zeros((n_aux, n_bf * (n_bf + 1) // 2))
    for P in range(n_aux):
        g = rng.normal(scale=np.sqrt(1 / n_bf), size=(n_bf, n_bf))
        g = 0.5 * (g + g.T)
        for hi in range(n_bf):
            for lo in range(hi + 1):
                B[P, hi * (hi + 1) // 2 + lo] = g[hi, lo]
    n_act = nact_a + nact_b
    Cm = np.zeros((n_act, n_bf))
    sup = np.zeros((n_act, n_bf), np.uint8)
    for k in range(n_act):
        f = 0 if k < nact_a else 1
        sup[k, f * n_frag_bf:(f + 1) * n_frag_bf] = 1
        Cm[k, f * n_frag_bf:(f + 1) * n_frag_bf] = rng.normal(size=n_frag_bf) / np.sqrt(n_frag_bf)
    return n_bf, S, h, B.flatten(order="F"), Cm, sup
