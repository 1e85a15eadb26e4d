"""CPU restatement of the projected L-BFGS orbital optimisation (test infrastructure only).

Only tests/ may import this module; the engine never does. It follows the reference's P-LBFGS
backend step by step on the oracle's FP64 energy and gradient:
  - outer loop, descent fallbacks, stall handling, convergence tests:
    src/vbscf/optimization/backends/plbfgs.cpp:20-226
  - Armijo backtracking: src/vbscf/optimization/globalization/line_search.cpp:33-95
  - zero-step and no-progress predicates: src/vbscf/optimization/driver/checks.hpp:33-66
  - two-loop recursion, curvature filter, secant transport and append:
    src/vbscf/optimization/preconditioners/transported_lbfgs.cpp:14-176
  - defaults: src/vbscf/optimization/driver/options.hpp:35-72
The chart is the whitened product-sphere chart of each orbital on its support (the energy is
invariant under rescaling an orbital, orbitals/charts/chart.cpp:560-660, whitening.cpp), written in
raw coefficients: reduced gradient <-> r_i = g_i - (S c_i)(c_i^T g_i); reduced step <-> the
S-orthogonal projection of a step; identity inverse <-> S_i^-1 on the support. The initial inverse
is the scaled identity (LbfgsInitialInverse::ScaledIdentity) and the convergence measure is the
largest per-orbital dual norm sqrt(r_i^T S_i^-1 r_i).
"""
from __future__ import annotations

import numpy as np

from . import evaluate

EPS = np.finfo(float).eps
TERMINATIONS = ("max_iterations", "plbfgs_initial_tolerance", "plbfgs_projected_gradient_tolerance",
                "plbfgs_line_search_failed", "plbfgs_line_search_stalled", "plbfgs_non_descent_direction",
                "nonredundant_space_empty")


def _restricted_inverse(S, mask):
    idx = np.flatnonzero(mask)
    out = np.zeros_like(S)
    out[np.ix_(idx, idx)] = np.linalg.inv(S[np.ix_(idx, idx)])
    return out


def plbfgs(n_bf, n_aux, topo, S, h, B, C0, support=None, max_iterations=2000, history_size=100,
           gradient_tolerance=1e-3, energy_tolerance=1e-7, initial_step_size=1.0, minimum_step_size=1e-7,
           armijo_constant=1e-4, states=None, weights=None):
    S = np.asarray(S, dtype=np.float64)
    n_act = C0.shape[0]
    masks = np.ones((n_act, n_bf), bool) if support is None else np.asarray(support).astype(bool)
    Sinv = [_restricted_inverse(S, masks[i]) for i in range(n_act)]
    reduced_size = int(sum(masks[i].sum() - 1 for i in range(n_act)))
    sup = None if support is None else np.ascontiguousarray(support, np.uint8)
    counts = dict(energy_evaluations=0, gradient_evaluations=0)

    def energy(Cm):
        counts["energy_evaluations"] += 1
        try:
            return evaluate(n_bf, n_aux, topo, S, h, B, Cm, support=sup, gradient=False, states=states,
                            weights=weights)["energy"]
        except RuntimeError:
            return float("nan")

    def gradient(Cm):
        counts["gradient_evaluations"] += 1
        return evaluate(n_bf, n_aux, topo, S, h, B, Cm, support=sup, gradient=True, states=states,
                        weights=weights)["grad"]

    def chart(Cm, Gm):
        Sc = np.where(masks, Cm @ S, 0.0)
        rho = np.sqrt(np.einsum("ik,ik->i", Cm, Sc))
        if not np.all(rho > 0):
            raise RuntimeError("orbital has no positive norm on its support")
        Cm, Gm, Sc = Cm / rho[:, None], Gm * rho[:, None], Sc / rho[:, None]
        beta = np.einsum("ik,ik->i", Cm, Gm)
        R = Gm - Sc * beta[:, None]
        nu = np.sqrt(np.maximum([R[i] @ Sinv[i] @ R[i] for i in range(n_act)], 0.0))
        return Cm, Gm, Sc, R, float(nu.max()), float(np.abs(Cm).max())

    def transport(pair, Cm, Sc):
        s, y = pair
        sig = np.einsum("ik,ik->i", Sc, s)
        bet = np.einsum("ik,ik->i", Cm, y)
        sp, yp = s - Cm * sig[:, None], y - Sc * bet[:, None]
        ss = float(sum(sp[i] @ S @ sp[i] for i in range(n_act)))
        yy = float(sum(yp[i] @ Sinv[i] @ yp[i] for i in range(n_act)))
        sy = float(np.sum(sp * yp))
        sn, yn = np.sqrt(max(ss, 0.0)), np.sqrt(max(yy, 0.0))
        ok = sn > 0 and yn > 0 and np.isfinite(sn) and np.isfinite(yn) and np.isfinite(sy) and sy > np.sqrt(EPS) * sn * yn
        return ok, sp, yp, sy, yy

    def direction(Cm, Gm, Sc, R, history):
        kept = []
        for pair in history:  # oldest to newest
            ok, sp, yp, sy, yy = transport(pair, Cm, Sc)
            if ok:
                kept.append((sp, yp, 1.0 / sy, sy / yy))
        q = R.copy()
        alphas = [0.0] * len(kept)
        for k in range(len(kept) - 1, -1, -1):
            sp, yp, rho, _ = kept[k]
            alphas[k] = rho * float(np.sum(sp * q))
            if not np.isfinite(alphas[k]):
                raise RuntimeError("non-finite first loop")
            q = q - alphas[k] * yp
        gamma = kept[-1][3] if kept else 1.0
        z = gamma * np.stack([Sinv[i] @ q[i] for i in range(n_act)])
        for k in range(len(kept)):
            sp, yp, rho, _ = kept[k]
            b = rho * float(np.sum(yp * z))
            if not np.isfinite(b):
                raise RuntimeError("non-finite second loop")
            z = z + sp * (alphas[k] - b)
        t = -z
        return t, float(np.sum(Gm * t)), float(np.abs(t).max())

    zero_step = 128.0 * EPS

    def armijo(Cm, E0, t, dd, tinf, cmax, step0):
        if not np.isfinite(dd) or dd >= 0.0:
            return None
        step = max(minimum_step_size, step0)
        while step >= minimum_step_size:
            if step * tinf <= zero_step * max(1.0, cmax):
                step *= 0.5
                continue
            Ct = Cm + step * t
            Et = energy(Ct)
            if np.isfinite(Et) and Et <= E0 + armijo_constant * step * dd:
                return Ct, Et, gradient(Ct)
            step *= 0.5
        return None

    C = np.where(masks, np.asarray(C0, dtype=np.float64), 0.0)
    E = energy(C)
    if not np.isfinite(E):
        raise RuntimeError("the starting orbitals cannot be evaluated")
    G = gradient(C)
    C, G, Sc, R, gnorm, cmax = chart(C, G)
    trace = [E]
    history = []
    iterations, converged, termination = 0, False, 0
    previous_energy = E
    for it in range(max_iterations):
        if reduced_size == 0:
            termination = 6
            break
        gnorm_before = gnorm
        if it == 0 and gnorm < gradient_tolerance:
            converged, termination = True, 1
            break
        t, dd, tinf = direction(C, G, Sc, R, history)
        if not (np.isfinite(dd) and dd < 0.0):
            history = []
            t, dd, tinf = direction(C, G, Sc, R, history)
        if not (np.isfinite(dd) and dd < 0.0) or tinf <= zero_step * max(1.0, cmax):
            history = []
            t, dd, tinf = direction(C, G, Sc, R, history)
        if not (np.isfinite(dd) and dd < 0.0):
            termination = 5
            break
        reference_energy, cmax_before = E, cmax
        Cp, Gp = C, G
        acc = armijo(C, E, t, dd, tinf, cmax, min(1.0, initial_step_size))
        if acc is None:
            termination = 3
            break
        Ct, E, Gt = acc
        C, G, Sc, R, gnorm, cmax = chart(Ct, Gt)
        step_inf = float(np.abs(C - Cp).max())
        no_progress = (abs(E - reference_energy) <= 4096.0 * EPS * max(1.0, abs(reference_energy))
                       and not gnorm_before < gradient_tolerance and gnorm >= 0.9 * gnorm_before)
        if (step_inf <= zero_step * max(1.0, cmax_before) or no_progress) and gnorm_before >= gradient_tolerance:
            C, G, Sc, R, gnorm, cmax = chart(Cp, Gp)
            history = []
            t, dd, tinf = direction(C, G, Sc, R, history)
            sd_step = max(minimum_step_size, min(min(1.0, initial_step_size), 1.0 / max(1.0, gnorm)))
            acc = armijo(C, reference_energy, t, dd, tinf, cmax, sd_step)
            if acc is None:
                E = reference_energy
                termination = 4
                break
            Ct, E, Gt = acc
            Cp, Gp = C, G
            C, G, Sc, R, gnorm, cmax = chart(Ct, Gt)
        iterations += 1
        trace.append(E)
        energy_change = E - previous_energy
        previous_energy = E
        if abs(energy_change) < energy_tolerance and gnorm < gradient_tolerance:
            converged, termination = True, 2
            break
        if history_size > 0:
            ok = transport((C - Cp, G - Gp), C, Sc)[0]
            if ok:
                history.append((C - Cp, G - Gp))
                if len(history) > history_size:
                    history.pop(0)
    return dict(energy=E, gradient_norm=gnorm, iterations=iterations, converged=converged,
                termination=TERMINATIONS[termination], C=C, grad=G, trace=np.array(trace), **counts)
