"""Structure generation and determinant expansion: product vs oracle bit-exact, pinned by the
reference's own count fixtures and by closed forms / spin algebra."""
import hashlib
import json
import os
from math import comb

import numpy as np
import pytest

import oracle
import xmvb_amd as vb

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "topology.json")
FIELDS = ("alpha_masks", "beta_masks", "det_alpha", "det_beta", "term_offsets", "term_structure", "term_coefficient")

CASES = [("COV", 4, 4, 1), ("COV", 6, 6, 1), ("COV", 8, 8, 1), ("COV", 10, 10, 1), ("COV", 12, 12, 1),
         ("FULL", 2, 2, 1), ("FULL", 3, 2, 1), ("FULL", 4, 4, 1), ("FULL", 5, 4, 1), ("FULL", 6, 6, 1),
         ("ION", 6, 6, 1), ("COV ION(1)", 6, 6, 1), ("COV", 5, 5, 2), ("FULL", 6, 5, 2), ("COV", 7, 5, 2), ("COV", 6, 4, 3),
         ("FULL", 4, 3, 2)]


def digest(topo) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(topo.raw_structures, np.int32).tobytes())
    for f in FIELDS:
        h.update(np.ascontiguousarray(getattr(topo, f)).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("kw,n,ne,mult", CASES)
def test_product_matches_oracle_bit_exact(kw, n, ne, mult):
    t = vb.Topology.build(kw, n, ne, mult)
    o = oracle.topology(kw, n, ne, mult)
    assert np.array_equal(t.raw_structures, o["raw_structures"])
    for f in FIELDS:
        assert np.array_equal(getattr(t, f), o[f]), f


def test_reference_count_fixtures():
    # tests/input/test_generated_structures.cpp:84-141
    assert vb.count_structures("FULL", 18, 18, 1) == 449141836
    assert vb.count_structures("FULL", 20, 20, 1) == 5924217936
    assert vb.count_structures("FULL", 3, 2, 1) == 6
    assert vb.Topology.build("FULL", 3, 2, 1).n_structures == 6
    for n in range(1, 8):
        for ne in range(1, n + 1):
            assert vb.count_structures("FULL", n, ne, ne + 1) == comb(n, ne)
        assert vb.count_structures("FULL", n, 2 * n, 1) == 1
    assert oracle.count_structures("FULL", 18, 18, 1) == 449141836


@pytest.mark.parametrize("n", [4, 6, 8, 10, 12, 14])
def test_rumer_closed_forms(n):
    # Catalan structure counts, 2^(n/2) determinants per covalent structure, unique alpha = C(n, n/2)
    # (tests/vbscf/unit/test_structure_tiles.cpp:426-487: n=4 -> 2, n=6 -> 5)
    t = vb.Topology.build("COV", n, n, 1)
    catalan = comb(n, n // 2) - comb(n, n // 2 - 1)
    assert t.n_structures == catalan
    assert len(t.term_structure) == catalan * 2 ** (n // 2)
    assert len(t.alpha_masks) == comb(n, n // 2) == t.n_determinants
    # beta string of every determinant is the complement of its alpha string
    full = (1 << n) - 1
    assert np.all(t.beta_masks[t.det_beta] == (full ^ t.alpha_masks[t.det_alpha].astype(np.uint64)))
    assert np.all(np.abs(t.term_coefficient) == 1.0)


def test_full_all_alpha_subsets():
    # all-alpha FULL spaces enumerate every occupied subset exactly once
    # (tests/input/test_generated_structures.cpp:33-65)
    for n in range(1, 7):
        for ne in range(1, n + 1):
            t = vb.Topology.build("FULL", n, ne, ne + 1)
            rows = {tuple(r) for r in t.raw_structures}
            assert len(rows) == comb(n, ne) == t.n_structures
            assert all(list(r) == sorted(r) for r in rows)


def _apply_raise(vec, n):
    """S_+ = sum_p a+_{p alpha} a_{p beta} on {(alpha_mask, beta_mask): coef}, alpha-before-beta order."""
    out = {}
    for (am, bm), c in vec.items():
        for p in range(n):
            if not (bm >> p & 1) or (am >> p & 1):
                continue
            # annihilate beta p: passes all alpha electrons and the beta electrons below p
            sign = (-1) ** (bin(am).count("1") + bin(bm & ((1 << p) - 1)).count("1"))
            bm2 = bm & ~(1 << p)
            sign *= (-1) ** bin(am & ((1 << p) - 1)).count("1")
            key = (am | (1 << p), bm2)
            out[key] = out.get(key, 0.0) + sign * c
    return {k: v for k, v in out.items() if v != 0.0}


@pytest.mark.parametrize("kw,n,ne,mult", [("COV", 4, 4, 1), ("COV", 6, 6, 1), ("FULL", 4, 4, 1), ("FULL", 3, 2, 1),
                                          ("COV", 8, 8, 1), ("FULL", 6, 6, 1), ("COV", 5, 5, 2), ("FULL", 6, 5, 2)])
def test_structures_are_spin_eigenfunctions(kw, n, ne, mult):
    # S_+ Psi = 0 and S_z = (mult-1)/2  =>  S^2 = S(S+1) with S = (mult-1)/2
    # (tests/vbscf/unit/test_structure_spin.cpp:44-57 raising-operator test)
    t = vb.Topology.build(kw, n, ne, mult)
    vecs = [dict() for _ in range(t.n_structures)]
    for d in range(t.n_determinants):
        key = (int(t.alpha_masks[t.det_alpha[d]]), int(t.beta_masks[t.det_beta[d]]))
        for k in range(t.term_offsets[d], t.term_offsets[d + 1]):
            vecs[t.term_structure[k]][key] = vecs[t.term_structure[k]].get(key, 0.0) + t.term_coefficient[k]
    for v in vecs:
        assert v and _apply_raise(v, n) == {}


def test_contaminated_vector_fails_raise():
    # a lone determinant of a singlet pair is not spin adapted
    assert _apply_raise({(0b01, 0b10): 1.0}, 2) != {}


def test_golden_digests():
    with open(GOLDEN) as f:
        gold = json.load(f)
    for key, rec in gold.items():
        kw, n, ne, mult = rec["case"]
        t = vb.Topology.build(kw, n, ne, mult)
        assert t.n_structures == rec["n_structures"] and t.n_determinants == rec["n_determinants"]
        assert t.raw_structures[:3].tolist() == rec["first_structures"]
        assert digest(t) == rec["sha256"], key


def test_invalid_class_rejected():
    with pytest.raises(vb.VbeError):
        vb.Topology.build("BOGUS", 4, 4, 1)
    with pytest.raises(vb.VbeError):
        vb.count_structures("ION(5)", 4, 4, 1)


def _flip_pairs_brute(topo):
    """Unordered determinant pairs {d, e} with one representative per spin-flip couple
    {d, e} / {flip d, flip e} (the smaller packed index), by enumeration."""
    nd = len(topo.det_alpha)
    key = {(int(topo.alpha_masks[topo.det_alpha[d]]), int(topo.beta_masks[topo.det_beta[d]])): d for d in range(nd)}
    flip = [key.get((int(topo.beta_masks[topo.det_beta[d]]), int(topo.alpha_masks[topo.det_alpha[d]])))
            for d in range(nd)]
    if topo.n_alpha != topo.n_beta or any(f is None for f in flip):
        return None, nd * (nd + 1) // 2
    kept = 0
    for e in range(nd):
        for d in range(e + 1):
            lo, hi = min(flip[d], flip[e]), max(flip[d], flip[e])
            kept += e * (e + 1) // 2 + d <= hi * (hi + 1) // 2 + lo
    return flip, kept


@pytest.mark.parametrize("kw,n,ne,mult", [("COV", 6, 6, 1), ("COV", 8, 8, 1), ("FULL", 2, 2, 1), ("FULL", 4, 4, 1),
                                          ("FULL", 8, 14, 1), ("FULL", 6, 6, 1), ("COV", 6, 4, 3)])
def test_spin_flip_pairing(kw, n, ne, mult):
    """The spin-flip partner map is an involution that exchanges the strings, every couple of
    pairs {d, e} / {flip d, flip e} has exactly one representative, and the closed-form count the
    benchmark uses for the evaluated pairs equals the enumeration (the lane kernels' work list
    follows the same rule)."""
    import bench

    topo = vb.Topology.build(kw, n, ne, mult)
    flip, kept = _flip_pairs_brute(topo)
    assert bench.lane_pairs_evaluated(topo) == kept
    if flip is None:
        assert topo.n_alpha != topo.n_beta
        return
    nd = len(flip)
    for d in range(nd):
        assert flip[flip[d]] == d
        assert topo.alpha_masks[topo.det_alpha[flip[d]]] == topo.beta_masks[topo.det_beta[d]]
    # couples: each unordered pair maps to a representative, each representative covers 1 or 2 pairs
    covered = {}
    for e in range(nd):
        for d in range(e + 1):
            lo, hi = min(flip[d], flip[e]), max(flip[d], flip[e])
            rep = min(e * (e + 1) // 2 + d, hi * (hi + 1) // 2 + lo)
            covered[rep] = covered.get(rep, 0) + 1
    assert len(covered) == kept and set(covered.values()) <= {1, 2}
    assert sum(covered.values()) == nd * (nd + 1) // 2
