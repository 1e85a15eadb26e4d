"""Pins the FP64 oracle against checks that do not restate its formulas: brute-force minors
(numpy determinants), an orthonormal-basis FCI projection, scipy's generalised eigensolver, dense
einsum integrals, central finite differences (including exactly singular determinant blocks) and
per-orbital scale invariance."""
import itertools

import numpy as np
import pytest
import scipy.linalg

import fci_reference
from problems import fragment_problem
import oracle
import xmvb_amd as vb


def _minor(x, rows, cols):
    keep_r = [i for i in range(x.shape[0]) if i not in rows]
    keep_c = [j for j in range(x.shape[1]) if j not in cols]
    if not keep_r:
        return 1.0
    return np.linalg.det(x[np.ix_(keep_r, keep_c)])


def _brute_cofactors(x):
    n = x.shape[0]
    c1 = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            c1[i, j] = (-1) ** (i + j) * _minor(x, [i], [j])
    c2 = np.zeros((n, n, n, n))
    for r1, r2, l1, l2 in itertools.product(range(n), repeat=4):
        if r1 == r2 or l1 == l2:
            continue
        # d^2 det / dX[r1,l1] dX[r2,l2]
        sgn = (-1) ** (r1 + r2 + l1 + l2)
        sgn *= (1 if (r1 < r2) == (l1 < l2) else -1)
        c2[r1, r2, l1, l2] = sgn * _minor(x, [r1, r2], [l1, l2])
    return c1, c2


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7])
def test_regular_cofactors_vs_minors(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, n)) + 0.5 * np.eye(n)
    det, nul, c1, c2 = oracle.resolve(x, want_cof2=True)
    b1, b2 = _brute_cofactors(x)
    assert nul == 0
    assert det == pytest.approx(np.linalg.det(x), rel=1e-12)
    assert np.abs(c1 - b1).max() <= 1e-12 * max(1, np.abs(b1).max())
    if n >= 2:
        assert np.abs(c2 - b2).max() <= 1e-12 * max(1, np.abs(b2).max())


@pytest.mark.parametrize("n,k", [(2, 1), (3, 1), (3, 2), (4, 1), (4, 2), (5, 1), (5, 2), (6, 3), (7, 2)])
def test_singular_cofactors_vs_minors(n, k):
    """Exactly rank-(n-k) blocks: nullity classification and SVD-form cofactors (contractions.cpp:172-236,901-983)."""
    rng = np.random.default_rng(10 * n + k)
    x = rng.normal(size=(n, n - k)) @ rng.normal(size=(n - k, n))
    det, nul, c1, c2 = oracle.resolve(x, want_cof2=True)
    b1, b2 = _brute_cofactors(x)
    assert nul == k
    assert det == 0.0
    if k == 1:
        assert np.abs(c1 - b1).max() <= 1e-10 * max(1, np.abs(b1).max())
    else:
        assert np.abs(c1).max() <= 1e-10
    if n >= 2 and k <= 2:
        assert np.abs(c2 - b2).max() <= 1e-10 * max(1, np.abs(b2).max())


def test_zero_block_full_nullity():
    det, nul, c1, _ = oracle.resolve(np.zeros((4, 4)))
    assert nul == 4 and det == 0.0 and not c1.any()


def _problem(kw, n_act, n_el, n_bf, n_aux, seed, mult=1):
    sp = vb.synthetic_problem(seed, n_bf, n_act, n_aux)
    topo = vb.Topology.build(kw, n_act, n_el, mult)
    return sp, topo


@pytest.mark.parametrize("kw,n_act,n_el,n_bf,n_aux", [("COV", 4, 4, 10, 20), ("FULL", 2, 2, 8, 12),
                                                     ("FULL", 4, 4, 10, 20), ("COV", 6, 6, 14, 28),
                                                     ("FULL", 3, 2, 9, 12), ("COV", 5, 5, 12, 20, )])
def test_structure_matrices_vs_orthonormal_fci(kw, n_act, n_el, n_bf, n_aux):
    mult = 2 if n_el % 2 else 1
    sp, topo = _problem(kw, n_act, n_el, n_bf, n_aux, 11, mult)
    r = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, gradient=False)
    H, S = fci_reference.structure_matrices(r["S_act"], r["h_act"], r["g_packed"], topo)
    # reference parity tolerances for cofactor vs orthogonal forms: |dH| < 2e-9, |dS| < 2e-11
    # (tests/vbscf/diagnostics/action_parity.cpp:14-86)
    assert np.abs(H - r["H"]).max() < 2e-9
    assert np.abs(S - r["S"]).max() < 2e-11


def test_active_integrals_vs_dense_einsum():
    n_bf, n_act, n_aux = 11, 4, 9
    sp, topo = _problem("COV", n_act, n_act, n_bf, n_aux, 3)
    r = oracle.evaluate(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C, gradient=False)
    Cn = sp.C / np.sqrt(np.einsum("ka,ab,kb->k", sp.C, sp.S, sp.C))[:, None]
    np.testing.assert_allclose(r["S_act"], Cn @ sp.S @ Cn.T, atol=1e-13)
    np.testing.assert_allclose(r["h_act"], Cn @ sp.h @ Cn.T, atol=1e-12)
    Bm = sp.B.reshape((n_aux, -1), order="F")
    Bu = np.zeros((n_aux, n_bf, n_bf))
    for hi in range(n_bf):
        for lo in range(hi + 1):
            Bu[:, hi, lo] = Bu[:, lo, hi] = Bm[:, hi * (hi + 1) // 2 + lo]
    L = np.einsum("kp,Apq,lq->Akl", Cn, Bu, Cn)
    eri = np.einsum("Apq,Ars->pqrs", L, L)
    g = r["g_packed"]
    for p, q, rr, s in itertools.product(range(n_act), repeat=4):
        assert g[fci_reference.pair(fci_reference.pair(p, q), fci_reference.pair(rr, s))] == pytest.approx(
            eri[p, q, rr, s], abs=1e-12)


def test_gep_vs_scipy():
    sp, topo = _problem("COV", 8, 8, 30, 60, 4)
    r = oracle.evaluate(30, 60, topo, sp.S, sp.h, sp.B, sp.C, gradient=False)
    w_ref, x_ref = scipy.linalg.eigh(r["H"], r["S"])
    w, x = oracle.gep(r["H"], r["S"])
    np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-10)
    assert r["energy"] == pytest.approx(w_ref[0], abs=1e-10)
    c = r["coef"]
    assert c @ r["S"] @ c == pytest.approx(1.0, abs=1e-12)
    assert min(np.abs(c - x_ref[:, 0]).max(), np.abs(c + x_ref[:, 0]).max()) < 1e-8


def _fd_check(n_bf, n_aux, topo, S, h, B, Cm, support=None, entries=8, eps=1e-5, seed=0):
    r = oracle.evaluate(n_bf, n_aux, topo, S, h, B, Cm, support=support)
    g = r["grad"]
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(entries):
        k = int(rng.integers(Cm.shape[0]))
        cols = np.arange(n_bf) if support is None else np.nonzero(support[k])[0]
        mu = int(rng.choice(cols))
        cp, cm = Cm.copy(), Cm.copy()
        cp[k, mu] += eps
        cm[k, mu] -= eps
        ep = oracle.evaluate(n_bf, n_aux, topo, S, h, B, cp, support=support, gradient=False)["energy"]
        em = oracle.evaluate(n_bf, n_aux, topo, S, h, B, cm, support=support, gradient=False)["energy"]
        worst = max(worst, abs((ep - em) / (2 * eps) - g[k, mu]))
    return r, worst


@pytest.mark.parametrize("kw,n_act,n_el,n_bf,n_aux", [("COV", 4, 4, 12, 20), ("FULL", 2, 2, 10, 15),
                                                     ("COV", 6, 6, 16, 30), ("FULL", 4, 4, 14, 25),
                                                     ("COV", 8, 8, 24, 40), ("COV", 5, 5, 12, 20)])
def test_gradient_vs_finite_differences(kw, n_act, n_el, n_bf, n_aux):
    mult = 2 if n_el % 2 else 1
    sp, topo = _problem(kw, n_act, n_el, n_bf, n_aux, 7, mult)
    r, worst = _fd_check(n_bf, n_aux, topo, sp.S, sp.h, sp.B, sp.C)
    assert worst < 2e-8
    # E is invariant under rescaling each orbital: <dE/dc_k, c_k> = 0
    assert np.abs(np.einsum("ka,ka->k", r["grad"], sp.C)).max() < 1e-12


@pytest.mark.parametrize("kw,n_frag_bf,a,b,n_aux", [("COV", 5, 2, 2, 12), ("FULL", 5, 2, 2, 12), ("COV", 6, 3, 3, 14)])
def test_singular_pairs_gradient_vs_finite_differences(kw, n_frag_bf, a, b, n_aux):
    n_bf, S, h, B, Cm, sup = fragment_problem(3, n_frag_bf, a, b, n_aux)
    topo = vb.Topology.build(kw, a + b, a + b, 1)
    r, worst = _fd_check(n_bf, n_aux, topo, S, h, B, Cm, support=sup, entries=10)
    sact = r["S_act"]
    singular = sum(
        abs(np.linalg.det(sact[np.ix_(fci_reference.occ(int(r2)), fci_reference.occ(int(l2)))])) < 1e-12
        for l2 in topo.alpha_masks for r2 in topo.alpha_masks)
    assert singular > 0
    assert worst < 2e-8
    assert np.all(r["grad"][sup == 0] == 0.0)
    H, Sm = fci_reference.structure_matrices(r["S_act"], r["h_act"], r["g_packed"], topo)
    assert np.abs(H - r["H"]).max() < 2e-9 and np.abs(Sm - r["S"]).max() < 2e-11
