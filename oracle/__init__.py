"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the
engine never does. The oracle restates the reference's algorithm in oracle/vb_oracle.cpp.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"oracle not built: {LIB_PATH}")
        lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_count_structures.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
        lib.oracle_topology.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P]
        lib.oracle_topology_fill.argtypes = [P] * 8
        lib.oracle_resolve.argtypes = [C.c_int, P, C.POINTER(C.c_double), C.POINTER(C.c_int), P, P]
        lib.oracle_spin_pair.argtypes = [C.c_int, P, P, P, C.c_uint64, C.c_uint64, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int)]
        lib.oracle_det_pair.argtypes = [C.c_int, P, P, P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.oracle_gep.argtypes = [C.c_int, P, P, P, P]
        lib.oracle_eval.argtypes = ([C.c_int] * 3 + [P] * 5 + [C.c_int, C.c_int, P, C.c_int, P, C.c_int] + [P] * 5 +
                                    [C.c_double, C.c_int, P] + [P] * 9)
        lib.oracle_eval_states.argtypes = ([C.c_int] * 3 + [P] * 5 + [C.c_int, C.c_int, P, C.c_int, P, C.c_int] +
                                           [P] * 5 + [C.c_double, C.c_int, C.c_int, P, P] + [P] * 11)
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _err(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {library().oracle_last_error().decode()}")


def count_structures(kw: str, n_act: int, n_el: int, mult: int = 1) -> int:
    out = C.c_uint64()
    _err(library().oracle_count_structures(kw.encode(), n_act, n_el, mult, C.byref(out)))
    return out.value


def topology(kw: str, n_act: int, n_el: int, mult: int = 1) -> dict:
    lib = library()
    sizes = np.zeros(5, np.int64)
    _err(lib.oracle_topology(kw.encode(), n_act, n_el, mult, _p(sizes)))
    ns, na, nb, nd, nt = (int(x) for x in sizes)
    out = dict(alpha_masks=np.zeros(na, np.uint64), beta_masks=np.zeros(nb, np.uint64),
               det_alpha=np.zeros(nd, np.int32), det_beta=np.zeros(nd, np.int32),
               term_offsets=np.zeros(nd + 1, np.int64), term_structure=np.zeros(nt, np.int32),
               term_coefficient=np.zeros(nt), raw_structures=np.zeros((ns, n_el), np.int32))
    _err(lib.oracle_topology_fill(*(_p(out[k]) for k in ("alpha_masks", "beta_masks", "det_alpha", "det_beta",
                                                          "term_offsets", "term_structure", "term_coefficient",
                                                          "raw_structures"))))
    out["n_structures"] = ns
    return out


def resolve(x: np.ndarray, want_cof2: bool = False):
    """det, nullity, first-order cofactor matrix, dense second-order cofactors of block X (rows = right)."""
    n = x.shape[0]
    xf = np.asfortranarray(x, dtype=np.float64)
    det, nul = C.c_double(), C.c_int()
    cof1 = np.zeros((n, n), order="F")
    cof2 = np.zeros(n ** 4) if want_cof2 else None
    library().oracle_resolve(n, _p(xf), C.byref(det), C.byref(nul), _p(cof1), _p(cof2))
    if want_cof2:
        cof2 = cof2.reshape((n, n, n, n), order="F")  # [r1, r2, l1, l2]
    return det.value, nul.value, cof1, cof2


def spin_pair(sact, hact, g, mL: int, mR: int):
    n = sact.shape[0]
    det, hs, nul = C.c_double(), C.c_double(), C.c_int()
    _err(library().oracle_spin_pair(n, _p(np.asfortranarray(sact)), _p(np.asfortranarray(hact)),
                                    _p(np.ascontiguousarray(g)), mL, mR, C.byref(det), C.byref(hs), C.byref(nul)))
    return det.value, hs.value, nul.value


def det_pair(sact, hact, g, aL: int, aR: int, bL: int, bR: int):
    n = sact.shape[0]
    s, h = C.c_double(), C.c_double()
    _err(library().oracle_det_pair(n, _p(np.asfortranarray(sact)), _p(np.asfortranarray(hact)),
                                   _p(np.ascontiguousarray(g)), aL, aR, bL, bR, C.byref(s), C.byref(h)))
    return s.value, h.value


def gep(H, S):
    n = H.shape[0]
    w = np.zeros(n)
    x = np.zeros((n, n), order="F")
    _err(library().oracle_gep(n, _p(np.asfortranarray(H)), _p(np.asfortranarray(S)), _p(w), _p(x)))
    return w, x


def evaluate(n_bf, n_aux, topo, S, h, B, Cm, support=None, e_nuc=0.0, gradient=True, want_det=False,
             states=None, weights=None):
    """One FP64 E(+grad) evaluation. topo: object/dict with the expansion arrays and n_act.
    states/weights: selected structure roots and their (unnormalised) weights for a state-averaged
    energy and gradient; default the lowest root alone. out["state_energies"]: E_s + E_nuc."""
    get = (lambda k: topo[k]) if isinstance(topo, dict) else (lambda k: getattr(topo, k))
    n_act = int(get("n_act"))
    am, bm = np.ascontiguousarray(get("alpha_masks"), np.uint64), np.ascontiguousarray(get("beta_masks"), np.uint64)
    da, db = np.ascontiguousarray(get("det_alpha"), np.int32), np.ascontiguousarray(get("det_beta"), np.int32)
    off = np.ascontiguousarray(get("term_offsets"), np.int64)
    ts = np.ascontiguousarray(get("term_structure"), np.int32)
    tc = np.ascontiguousarray(get("term_coefficient"), np.float64)
    ns = int(get("n_structures"))
    nd = len(da)
    np_ = n_act * (n_act + 1) // 2
    Sf, hf = np.asfortranarray(S, np.float64), np.asfortranarray(h, np.float64)
    Bc, Cc = np.ascontiguousarray(B, np.float64), np.ascontiguousarray(Cm, np.float64)
    sup = None if support is None else np.ascontiguousarray(support, np.uint8)
    e = np.zeros(1)
    out = dict(grad=np.zeros((n_act, n_bf)) if gradient else None, H=np.zeros((ns, ns), order="F"),
               S=np.zeros((ns, ns), order="F"), coef=np.zeros(ns), S_act=np.zeros((n_act, n_act), order="F"),
               h_act=np.zeros((n_act, n_act), order="F"), g_packed=np.zeros(np_ * (np_ + 1) // 2),
               H_det=np.zeros((nd, nd), order="F") if want_det else None,
               S_det=np.zeros((nd, nd), order="F") if want_det else None)
    if states is None:
        _err(library().oracle_eval(n_bf, n_act, n_aux, _p(Sf), _p(hf), _p(Bc), _p(Cc), _p(sup), ns, len(am), _p(am),
                                   len(bm), _p(bm), nd, _p(da), _p(db), _p(off), _p(ts), _p(tc), e_nuc, int(gradient),
                                   _p(e), _p(out["grad"]), _p(out["H"]), _p(out["S"]), _p(out["coef"]),
                                   _p(out["S_act"]), _p(out["h_act"]), _p(out["g_packed"]), _p(out["H_det"]),
                                   _p(out["S_det"])))
    else:
        si = np.ascontiguousarray(states, np.int32)
        sw = np.ascontiguousarray(np.ones(len(si)) if weights is None else weights, np.float64)
        out["state_energies"] = np.zeros(len(si))
        _err(library().oracle_eval_states(n_bf, n_act, n_aux, _p(Sf), _p(hf), _p(Bc), _p(Cc), _p(sup), ns, len(am),
                                          _p(am), len(bm), _p(bm), nd, _p(da), _p(db), _p(off), _p(ts), _p(tc), e_nuc,
                                          int(gradient), len(si), _p(si), _p(sw), _p(e), _p(out["grad"]),
                                          _p(out["H"]), _p(out["S"]), _p(out["coef"]), _p(out["S_act"]),
                                          _p(out["h_act"]), _p(out["g_packed"]), _p(out["H_det"]), _p(out["S_det"]),
                                          _p(out["state_energies"])))
    out["energy"] = float(e[0])
    return out
