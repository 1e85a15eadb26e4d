"""Independent structure-space H/S by orthonormal-basis full CI (test helper, numpy only).

Orbitals are orthonormalised with S_act = R^T R; a nonorthogonal determinant of sorted occupied
strings (P_alpha, P_beta) expands over orthonormal determinants (K_alpha, K_beta) with coefficient
det R[K_alpha, P_alpha] * det R[K_beta, P_beta] (exterior power of R). Structure vectors are then
projected onto the orthonormal determinant space, where S = V^T V and
H = sum_pq h_pq <E_pq> + 1/2 sum_pqrs (pq|rs) (<E_pq E_rs> - delta_qr <E_ps>) with spin-summed
excitation operators E_pq built from second-quantised fermion signs. This is the pattern of the
reference's own parity test between its cofactor and orthogonal representations
(tests/vbscf/diagnostics/action_parity.cpp:14-86) and shares no code with the oracle.
"""
from __future__ import annotations

import itertools

import numpy as np


def pair(a, b):
    hi, lo = max(a, b), min(a, b)
    return hi * (hi + 1) // 2 + lo


def unpack_g(g_packed, n):
    g = np.zeros((n, n, n, n))
    for p, q, r, s in itertools.product(range(n), repeat=4):
        g[p, q, r, s] = g_packed[pair(pair(p, q), pair(r, s))]
    return g


def occ(mask):
    return [i for i in range(64) if mask >> i & 1]


def excitation(n_orb, strings):
    """E_pq restricted to one spin on a list of string masks: dict (p,q) -> matrix."""
    index = {m: i for i, m in enumerate(strings)}
    mats = {}
    for p in range(n_orb):
        for q in range(n_orb):
            M = np.zeros((len(strings), len(strings)))
            for j, m in enumerate(strings):
                if not (m >> q & 1):
                    continue
                m1 = m & ~(1 << q)
                sign = (-1) ** bin(m & ((1 << q) - 1)).count("1")
                if m1 >> p & 1:
                    continue
                sign *= (-1) ** bin(m1 & ((1 << p) - 1)).count("1")
                m2 = m1 | (1 << p)
                M[index[m2], j] += sign
            mats[p, q] = M
    return mats


def structure_matrices(sact, hact, g_packed, topo):
    n = sact.shape[0]
    na, nb = topo.n_alpha, topo.n_beta
    R = np.linalg.cholesky(sact).T  # upper, sact = R^T R
    Ri = np.linalg.inv(R)
    hbar = Ri.T @ hact @ Ri
    g = unpack_g(g_packed, n)
    gbar = np.einsum("ap,bq,cr,ds,abcd->pqrs", Ri, Ri, Ri, Ri, g, optimize=True)
    a_str = [sum(1 << i for i in c) for c in itertools.combinations(range(n), na)]
    b_str = [sum(1 << i for i in c) for c in itertools.combinations(range(n), nb)]
    Na, Nb = len(a_str), len(b_str)

    def ext(P, strings, k):
        v = np.zeros(len(strings))
        Pl = occ(P)
        for i, m in enumerate(strings):
            K = occ(m)
            v[i] = np.linalg.det(R[np.ix_(K, Pl)]) if k else 1.0
        return v

    T = topo.incidence()
    V = np.zeros((Na * Nb, topo.n_structures))
    for d in range(topo.n_determinants):
        va = ext(int(topo.alpha_masks[topo.det_alpha[d]]), a_str, na)
        vb = ext(int(topo.beta_masks[topo.det_beta[d]]), b_str, nb)
        V += np.outer(np.kron(va, vb), T[d])
    Ea, Eb = excitation(n, a_str), excitation(n, b_str)
    Ia, Ib = np.eye(Na), np.eye(Nb)
    # alpha-before-beta ordering: beta operators pick up (-1)^{n_alpha} twice per E_pq -> no sign
    U = {(p, q): (np.kron(Ea[p, q], Ib) + np.kron(Ia, Eb[p, q])) @ V for p in range(n) for q in range(n)}
    S = V.T @ V
    H = np.zeros_like(S)
    for p in range(n):
        for q in range(n):
            H += hbar[p, q] * (V.T @ U[p, q])
    for p, q, r, s in itertools.product(range(n), repeat=4):
        c = gbar[p, q, r, s]
        if c == 0.0:
            continue
        term = U[q, p].T @ U[r, s]
        if q == r:
            term = term - V.T @ U[p, s]
        H += 0.5 * c * term
    return H, S
