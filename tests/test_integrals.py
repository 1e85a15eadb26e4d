"""AO integrals of the embedded 6-31G** hydrogen basis (McMurchie-Davidson), pinned by closed forms
and invariances rather than by the recurrences themselves:
  * single-primitive s functions: the textbook two-centre overlap, kinetic, nuclear-attraction and
    (ss|ss) formulas (Szabo & Ostlund, Modern Quantum Chemistry, appendix A);
  * single-primitive p functions on two centres along z: <px|px> = exp(-a R^2 / 2) and
    <pz|pz> = exp(-a R^2 / 2) (1 - a R^2) for equal exponents; p kinetic diagonal 5a/2;
  * the two-electron matrix is symmetric positive semidefinite and its pivoted-Cholesky factors
    reproduce it; RHF and FCI energies are invariant under rotating and translating the molecule;
  * the VBSCF energy of the H2 deck lies between FCI and RHF in the same basis."""
import math

import numpy as np
import pytest
import scipy.linalg as sl
from scipy.special import erf

import xmvb_amd as vb
from oracle.plbfgs import plbfgs

R = 0.74 * vb.BOHR_PER_ANGSTROM
A1 = 0.1612778   # second s shell of H (one primitive): functions 1 (atom 0) and 6 (atom 1)
AP = 1.1         # p shell of H in 6-31G**: functions 2,3,4 (atom 0) and 7,8,9 (atom 1)


def boys0(t):
    return 1.0 if t < 1e-15 else 0.5 * math.sqrt(math.pi / t) * erf(math.sqrt(t))


def s_overlap(a, b, r):
    return (4 * a * b / (a + b) ** 2) ** 0.75 * math.exp(-a * b / (a + b) * r * r)


def s_kinetic(a, b, r):
    mu = a * b / (a + b)
    return mu * (3 - 2 * mu * r * r) * s_overlap(a, b, r)


def s_nuclear(a, b, ra, rb, rc):
    p = a + b
    rp = (a * ra + b * rb) / p
    norm = (2 * a / math.pi) ** 0.75 * (2 * b / math.pi) ** 0.75
    return -norm * 2 * math.pi / p * math.exp(-a * b / p * (ra - rb) ** 2) * boys0(p * (rp - rc) ** 2)


def s_eri(a, b, c, d, ra, rb, rc, rd):
    p, q = a + b, c + d
    rp, rq = (a * ra + b * rb) / p, (c * rc + d * rd) / q
    norm = 1.0
    for e in (a, b, c, d):
        norm *= (2 * e / math.pi) ** 0.75
    return (norm * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(-a * b / p * (ra - rb) ** 2 -
            c * d / q * (rc - rd) ** 2) * boys0(p * q / (p + q) * (rp - rq) ** 2))


def full_eri(eri, n):
    idx = lambda a, b: max(a, b) * (max(a, b) + 1) // 2 + min(a, b)
    g = np.zeros((n, n, n, n))
    for a in range(n):
        for b in range(n):
            for c in range(n):
                for d in range(n):
                    g[a, b, c, d] = eri[idx(a, b), idx(c, d)]
    return g


def rhf_energy(ints):
    S, h, n = ints["S"], ints["h"], ints["n_bf"]
    g = full_eri(ints["eri"], n)
    _, C = sl.eigh(h, S)
    for _ in range(200):
        D = 2 * np.outer(C[:, 0], C[:, 0])
        F = h + np.einsum("abcd,cd->ab", g, D) - 0.5 * np.einsum("acbd,cd->ab", g, D)
        _, C = sl.eigh(F, S)
    return 0.5 * np.sum(D * (h + F)) + ints["e_nuc"]


def fci_energy(ints):
    """Two-electron singlet FCI in the Lowdin-orthonormalised basis (symmetric spatial products)."""
    S, h, n = ints["S"], ints["h"], ints["n_bf"]
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    ho = X.T @ h @ X
    g = np.einsum("abcd,ai,bj,ck,dl->ijkl", full_eri(ints["eri"], n), X, X, X, X)
    pairs = [(i, j) for i in range(n) for j in range(i, n)]
    Hm = np.zeros((len(pairs), len(pairs)))
    for P, (i, j) in enumerate(pairs):
        for Q, (k, l) in enumerate(pairs):
            # spatial part (|ij> + |ji>)/sqrt(2 (1 + delta_ij)) for both
            def amp(a, b, c, d):  # <ab| H |cd> for a product of two spatial orbitals
                one = (ho[a, c] if b == d else 0.0) + (ho[b, d] if a == c else 0.0)
                return one + g[a, c, b, d]
            v = amp(i, j, k, l) + amp(i, j, l, k) + amp(j, i, k, l) + amp(j, i, l, k)
            Hm[P, Q] = v / (2 * math.sqrt((1 + (i == j)) * (1 + (k == l))))
    return np.linalg.eigvalsh(Hm)[0] + ints["e_nuc"]


@pytest.fixture(scope="module")
def ints():
    return vb.ao_integrals("6-31G**", [1, 1], [[0, 0, 0], [0, 0, R]])


def test_s_functions_match_closed_forms(ints):
    assert abs(ints["S"][1, 6] - s_overlap(A1, A1, R)) < 1e-14
    assert abs(ints["T"][1, 6] - s_kinetic(A1, A1, R)) < 1e-14
    assert abs(ints["T"][1, 1] - 1.5 * A1) < 1e-14
    v = s_nuclear(A1, A1, 0.0, R, 0.0) + s_nuclear(A1, A1, 0.0, R, R)
    assert abs(ints["V"][1, 6] - v) < 1e-13
    idx = lambda a, b: max(a, b) * (max(a, b) + 1) // 2 + min(a, b)
    for (a, b, c, d) in [(1, 1, 6, 6), (1, 6, 1, 6), (1, 1, 1, 1), (1, 6, 6, 6)]:
        centres = [0.0 if k < 5 else R for k in (a, b, c, d)]
        ref = s_eri(A1, A1, A1, A1, *centres)
        assert abs(ints["eri"][idx(a, b), idx(c, d)] - ref) < 1e-13, (a, b, c, d)


def test_p_functions_match_closed_forms(ints):
    e = math.exp(-AP * R * R / 2)
    assert abs(ints["S"][2, 7] - e) < 1e-14        # px(A) px(B)
    assert abs(ints["S"][3, 8] - e) < 1e-14        # py(A) py(B)
    assert abs(ints["S"][4, 9] - e * (1 - AP * R * R)) < 1e-14  # pz(A) pz(B)
    assert abs(ints["S"][2, 9]) < 1e-15 and abs(ints["S"][2, 3]) < 1e-15
    for k in (2, 3, 4):
        assert abs(ints["T"][k, k] - 2.5 * AP) < 1e-13
        assert abs(ints["S"][k, k] - 1.0) < 1e-14


def test_two_electron_matrix_and_factors(ints):
    eri = ints["eri"]
    assert np.abs(eri - eri.T).max() == 0.0
    assert np.linalg.eigvalsh(eri).min() > -1e-13
    B = vb.pivoted_cholesky(eri, 1e-13)
    assert np.abs(B.T @ B - eri).max() < 1e-12


def test_energies_invariant_under_rotation_and_translation(ints):
    e_rhf, e_fci = rhf_energy(ints), fci_energy(ints)
    rot = sl.expm(np.array([[0, -0.3, 0.5], [0.3, 0, -0.7], [-0.5, 0.7, 0]]))
    xyz = np.array([[0, 0, 0], [0, 0, R]]) @ rot.T + np.array([0.4, -1.1, 2.3])
    moved = vb.ao_integrals("6-31G**", [1, 1], xyz)
    assert abs(rhf_energy(moved) - e_rhf) < 1e-11
    assert abs(fci_energy(moved) - e_fci) < 1e-11
    assert abs(moved["e_nuc"] - ints["e_nuc"]) < 1e-14
    assert e_fci < e_rhf


def test_h2_vbscf_between_fci_and_rhf(ints):
    d = vb.h2_deck()
    r = plbfgs(d["n_bf"], d["n_aux"], d["topo"], d["S"], d["h"], d["B"], d["C0"], support=d["support"],
               gradient_tolerance=1e-7, energy_tolerance=1e-13)
    e_vb = r["energy"] + d["e_nuc"]
    assert r["converged"]
    assert fci_energy(ints) < e_vb < rhf_energy(ints)
    # the optimised orbitals are mirror images of each other (reflection z -> R - z exchanges atoms)
    c = r["C"]
    mirror = c[0].copy()
    mirror[4] = -mirror[4]
    assert np.abs(np.abs(c[1][5:]) - np.abs(mirror[:5])).max() < 1e-6
