"""Python binding of the MI355X VBSCF energy+gradient engine (ctypes over include/xmvb_amd.h).

The host-side mirror of the reference's evaluation interface:
  * ``Topology`` - Rumer/Young structure generation + determinant expansion
    (src/input/loading/generated_structures.cpp, src/vbscf/structures/expansion/expander.cpp);
  * ``Context`` / ``Problem`` - device-resident problem and the two-phase evaluation contract of
    VbScfObjective (src/vbscf/optimization/objective/function.cpp:89-275):
    ``evaluate_forward`` (trial energy) then ``complete_gradient``; ``evaluate`` does both, like
    OrbitalGradientEvaluator::evaluate (src/vbscf/derivatives/gradient/orbital/evaluator.cpp:198-418);
  * ``parse_ctrl_method`` - the $CTRL method flag with the HIPVBSCF token (src/input/deck/model.cpp:360-381).

Errors from the C ABI are raised as ``VbeError`` carrying the status code; there is no CPU
fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libxmvb_amd.so")

STATUS = {
    0: "VBE_OK",
    1: "VBE_ERR_ARGUMENT",
    2: "VBE_ERR_NUMERIC",
    3: "VBE_ERR_METRIC",
    4: "VBE_ERR_DEVICE",
    5: "VBE_ERR_UNSUPPORTED",
    6: "VBE_ERR_MEMORY",
    7: "VBE_ERR_TOPOLOGY",
}
METHODS = {0: "VBSCF", 1: "CUVBSCF", 2: "CASSCF", 3: "DFT", 4: "HIPVBSCF"}

# every symbol include/xmvb_amd.h declares
EXPORTED_SYMBOLS = (
    "vbe_parse_ctrl_method", "vbe_check_hipvbscf_options", "vbe_count_structures", "vbe_topology_build",
    "vbe_topology_copy", "vbe_topology_view_of", "vbe_topology_free", "vbe_synthetic_problem",
    "vbe_context_create", "vbe_context_destroy", "vbe_last_error", "vbe_global_error", "vbe_problem_upload",
    "vbe_problem_release", "vbe_problem_device_bytes", "vbe_eval", "vbe_eval_device", "vbe_eval_forward",
    "vbe_complete_gradient", "vbe_last_intermediates", "vbe_last_phase_times", "vbe_last_singular_pairs",
    "vbe_pool_create", "vbe_pool_destroy", "vbe_pool_load", "vbe_pool_device_bytes", "vbe_pool_unload",
    "vbe_context_memory", "vbe_pool_autotune", "vbe_pool_eval_device",
    "vbe_pool_submit", "vbe_pool_wait", "vbe_pool_poll", "vbe_pool_phase_times", "vbe_set_pair_condition_limit", "vbe_set_ri_precision",
    "vbe_optimizer_defaults", "vbe_optimize", "vbe_pool_optimize", "vbe_set_eigensolver",
    "vbe_gaussian_basis_size", "vbe_ao_integrals", "vbe_pivoted_cholesky", "vbe_eval_device_noalloc",
    "vbe_server_create", "vbe_server_destroy", "vbe_server_submit", "vbe_server_collect", "vbe_server_stats",
    "vbe_set_states", "vbe_last_state_energies", "vbe_pool_state_energies",
)

TERMINATIONS = {0: "max_iterations", 1: "plbfgs_initial_tolerance", 2: "plbfgs_projected_gradient_tolerance",
                3: "plbfgs_line_search_failed", 4: "plbfgs_line_search_stalled", 5: "plbfgs_non_descent_direction",
                6: "nonredundant_space_empty"}


class VbeError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Dims(C.Structure):
    _fields_ = [("n_bf", C.c_int32), ("n_act", C.c_int32), ("n_alpha", C.c_int32), ("n_beta", C.c_int32),
                ("n_aux", C.c_int32)]


class TopologyView(C.Structure):
    _fields_ = [
        ("n_structures", C.c_int32),
        ("n_alpha_strings", C.c_int32), ("alpha_masks", C.POINTER(C.c_uint64)),
        ("n_beta_strings", C.c_int32), ("beta_masks", C.POINTER(C.c_uint64)),
        ("n_determinants", C.c_int32), ("det_alpha", C.POINTER(C.c_int32)), ("det_beta", C.POINTER(C.c_int32)),
        ("term_offsets", C.POINTER(C.c_int64)), ("term_structure", C.POINTER(C.c_int32)),
        ("term_coefficient", C.POINTER(C.c_double)),
    ]


class Flags(C.Structure):
    _fields_ = [("precision", C.c_int32), ("eigensolver", C.c_int32), ("want_gradient", C.c_int32)]


class PhaseTimes(C.Structure):
    _fields_ = [(name, C.c_float) for name in (
        "orbital_prep_ms", "transform_ms", "ri_one_index_ms", "pair_forward_ms", "structure_ms", "eigensolve_ms",
        "pair_backward_ms", "pair_backward_kernel_ms", "gradient_tail_ms", "total_ms", "pair_forward_kernel_ms",
        "pair_backward_field_ms", "pair_backward_adjoint_ms")]


class OptimizerOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("history_size", C.c_int32), ("gradient_tolerance", C.c_double),
                ("energy_tolerance", C.c_double), ("initial_step_size", C.c_double),
                ("minimum_step_size", C.c_double), ("armijo_constant", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("batch_size", C.c_int32), ("energy", C.c_double),
                ("t_arrival", C.c_double), ("t_submit", C.c_double), ("t_dispatch", C.c_double),
                ("t_complete", C.c_double), ("device_ms", C.c_double)]


class ServerCounters(C.Structure):
    _fields_ = [(name, C.c_int64) for name in ("queued", "in_flight_launches", "launches", "launched_requests",
                                                "completed", "uncollected")]


class OptimizerResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("gradient_norm", C.c_double), ("iterations", C.c_int32),
                ("energy_evaluations", C.c_int32), ("gradient_evaluations", C.c_int32), ("converged", C.c_int32),
                ("termination", C.c_int32)]


_lib = None


def library() -> C.CDLL:
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(LIB_PATH)
        P, I32, I64, U64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        lib.vbe_global_error.restype = C.c_char_p
        lib.vbe_parse_ctrl_method.argtypes = [C.c_char_p, C.POINTER(I32)]
        lib.vbe_check_hipvbscf_options.argtypes = [I32, I32, I32, I32]
        lib.vbe_count_structures.argtypes = [C.c_char_p, I32, I32, I32, C.POINTER(U64)]
        lib.vbe_topology_build.argtypes = [C.c_char_p, I32, I32, I32, C.POINTER(P), P]
        lib.vbe_topology_copy.argtypes = [P] * 9
        lib.vbe_topology_view_of.argtypes = [P, C.POINTER(TopologyView)]
        lib.vbe_topology_free.argtypes = [P]
        lib.vbe_topology_free.restype = None
        lib.vbe_synthetic_problem.argtypes = [U64, I32, I32, I32, D, P, P, P, P]
        lib.vbe_context_create.argtypes = [I32, U64, C.POINTER(P)]
        lib.vbe_context_destroy.argtypes = [P]
        lib.vbe_context_destroy.restype = None
        lib.vbe_last_error.argtypes = [P, C.c_char_p, C.c_size_t]
        lib.vbe_set_pair_condition_limit.argtypes = [P, D]
        lib.vbe_set_ri_precision.argtypes = [P, I32]
        lib.vbe_set_eigensolver.argtypes = [P, I32, D, D, D, I32]
        lib.vbe_gaussian_basis_size.argtypes = [C.c_char_p, I32, P, C.POINTER(I32)]
        lib.vbe_ao_integrals.argtypes = [C.c_char_p, I32, P, P, P, P, P, P, P, C.POINTER(D)]
        lib.vbe_pivoted_cholesky.argtypes = [I32, P, D, I32, C.POINTER(I32), P]
        lib.vbe_problem_upload.argtypes = [P, C.POINTER(Dims), P, P, P, C.POINTER(TopologyView), D, C.POINTER(P)]
        lib.vbe_problem_release.argtypes = [P]
        lib.vbe_problem_release.restype = None
        lib.vbe_problem_device_bytes.argtypes = [P, C.POINTER(U64)]
        lib.vbe_eval.argtypes = [P, P, P, P, Flags, C.POINTER(D), P, P]
        lib.vbe_eval_device.argtypes = [P, P, P, P, Flags, C.POINTER(D), P]
        lib.vbe_eval_device_noalloc.argtypes = [P, P, P, P, Flags, C.POINTER(D), P]
        lib.vbe_eval_forward.argtypes = [P, P, P, P, C.POINTER(D)]
        lib.vbe_complete_gradient.argtypes = [P, P, P]
        lib.vbe_last_intermediates.argtypes = [P, P, P, P, P, P, P, P]
        lib.vbe_last_phase_times.argtypes = [P, C.POINTER(PhaseTimes)]
        lib.vbe_last_singular_pairs.argtypes = [P, C.POINTER(I64)]
        lib.vbe_pool_create.argtypes = [P, C.POINTER(TopologyView), I32, I32, I32, I32, C.POINTER(P)]
        lib.vbe_pool_destroy.argtypes = [P]
        lib.vbe_pool_destroy.restype = None
        lib.vbe_pool_load.argtypes = [P, I32, I32, I32, P, P, P, D]
        lib.vbe_pool_device_bytes.argtypes = [P, C.POINTER(U64)]
        lib.vbe_pool_unload.argtypes = [P, I32]
        lib.vbe_pool_autotune.argtypes = [P, I32, P, P, I32, C.POINTER(I32), C.POINTER(I32), P]
        lib.vbe_context_memory.argtypes = [P, C.POINTER(U64), C.POINTER(U64)]
        lib.vbe_pool_eval_device.argtypes = [P, I32, P, P, I32, P, P]
        lib.vbe_pool_submit.argtypes = [P, I32, P, P, I32, C.POINTER(I32)]
        lib.vbe_pool_wait.argtypes = [P, I32, P, P]
        lib.vbe_pool_poll.argtypes = [P, I32, C.POINTER(I32)]
        lib.vbe_pool_phase_times.argtypes = [P, C.POINTER(PhaseTimes)]
        lib.vbe_optimizer_defaults.argtypes = [C.POINTER(OptimizerOptions)]
        lib.vbe_optimizer_defaults.restype = None
        lib.vbe_optimize.argtypes = [P, P, P, P, C.POINTER(OptimizerOptions), C.POINTER(OptimizerResult), P, P, I32]
        lib.vbe_pool_optimize.argtypes = [P, I32, P, P, P, C.POINTER(OptimizerOptions), P, P]
        lib.vbe_set_states.argtypes = [P, I32, P, P]
        lib.vbe_last_state_energies.argtypes = [P, P, P]
        lib.vbe_pool_state_energies.argtypes = [P, I32, P]
        lib.vbe_server_create.argtypes = [P, I32, I32, C.POINTER(P)]
        lib.vbe_server_destroy.argtypes = [P]
        lib.vbe_server_destroy.restype = None
        lib.vbe_server_submit.argtypes = [P, P, I32, P, I32, P, D, C.POINTER(I64)]
        lib.vbe_server_collect.argtypes = [P, I32, I32, C.POINTER(RequestInfo), C.POINTER(I32)]
        lib.vbe_server_stats.argtypes = [P, C.POINTER(ServerCounters)]
        _lib = lib
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check_global(rc: int):
    if rc != 0:
        raise VbeError(rc, library().vbe_global_error().decode())


def parse_ctrl_method(ctrl_line: str) -> str:
    out = C.c_int32()
    _check_global(library().vbe_parse_ctrl_method(ctrl_line.encode(), C.byref(out)))
    return METHODS[out.value]


def check_hipvbscf_options(eigensolver: int = 0, ri: bool = True, n_states: int = 1, n_inactive_electrons: int = 0):
    _check_global(library().vbe_check_hipvbscf_options(eigensolver, int(ri), n_states, n_inactive_electrons))


def count_structures(structure_class: str, n_act: int, n_electrons: int, multiplicity: int = 1) -> int:
    out = C.c_uint64()
    _check_global(library().vbe_count_structures(structure_class.encode(), n_act, n_electrons, multiplicity,
                                                 C.byref(out)))
    return out.value


@dataclass
class Topology:
    """Structure set expanded into signed determinants (StructureExpansion, expansion/types.hpp:390-423)."""
    n_act: int
    n_electrons: int
    multiplicity: int
    n_alpha: int
    n_beta: int
    raw_structures: np.ndarray  # (n_str, n_electrons) 1-based labels
    alpha_masks: np.ndarray
    beta_masks: np.ndarray
    det_alpha: np.ndarray
    det_beta: np.ndarray
    term_offsets: np.ndarray
    term_structure: np.ndarray
    term_coefficient: np.ndarray

    @property
    def n_structures(self) -> int:
        return self.raw_structures.shape[0]

    @property
    def n_determinants(self) -> int:
        return self.det_alpha.shape[0]

    @classmethod
    def build(cls, structure_class: str, n_act: int, n_electrons: int, multiplicity: int = 1) -> "Topology":
        lib = library()
        handle = C.c_void_p()
        sizes = np.zeros(5, np.int64)
        _check_global(lib.vbe_topology_build(structure_class.encode(), n_act, n_electrons, multiplicity,
                                             C.byref(handle), _ptr(sizes)))
        try:
            ns, na, nb, nd, nt = (int(x) for x in sizes)
            raw = np.zeros((ns, n_electrons), np.int32)
            am, bm = np.zeros(na, np.uint64), np.zeros(nb, np.uint64)
            da, db = np.zeros(nd, np.int32), np.zeros(nd, np.int32)
            off, ts, tc = np.zeros(nd + 1, np.int64), np.zeros(nt, np.int32), np.zeros(nt, np.float64)
            _check_global(lib.vbe_topology_copy(handle, _ptr(am), _ptr(bm), _ptr(da), _ptr(db), _ptr(off), _ptr(ts),
                                                _ptr(tc), _ptr(raw)))
        finally:
            lib.vbe_topology_free(handle)
        n_beta = (n_electrons - multiplicity + 1) // 2
        return cls(n_act, n_electrons, multiplicity, n_electrons - n_beta, n_beta, raw, am, bm, da, db, off, ts, tc)

    def view(self) -> TopologyView:
        v = TopologyView()
        v.n_structures = self.n_structures
        v.n_alpha_strings = len(self.alpha_masks)
        v.alpha_masks = self.alpha_masks.ctypes.data_as(C.POINTER(C.c_uint64))
        v.n_beta_strings = len(self.beta_masks)
        v.beta_masks = self.beta_masks.ctypes.data_as(C.POINTER(C.c_uint64))
        v.n_determinants = self.n_determinants
        v.det_alpha = self.det_alpha.ctypes.data_as(C.POINTER(C.c_int32))
        v.det_beta = self.det_beta.ctypes.data_as(C.POINTER(C.c_int32))
        v.term_offsets = self.term_offsets.ctypes.data_as(C.POINTER(C.c_int64))
        v.term_structure = self.term_structure.ctypes.data_as(C.POINTER(C.c_int32))
        v.term_coefficient = self.term_coefficient.ctypes.data_as(C.POINTER(C.c_double))
        return v

    def incidence(self) -> np.ndarray:
        """Dense (n_det, n_str) signed incidence matrix T."""
        t = np.zeros((self.n_determinants, self.n_structures))
        for d in range(self.n_determinants):
            for k in range(self.term_offsets[d], self.term_offsets[d + 1]):
                t[d, self.term_structure[k]] += self.term_coefficient[k]
        return t


@dataclass
class SyntheticProblem:
    n_bf: int
    n_act: int
    n_aux: int
    S: np.ndarray  # (n_bf, n_bf) Fortran order
    h: np.ndarray
    B: np.ndarray  # (n_aux * n_bf(n_bf+1)/2,) column-major packed
    C: np.ndarray  # (n_act, n_bf) orbital-major


def synthetic_problem(seed: int, n_bf: int, n_act: int, n_aux: int, scan_r: float = -1.0) -> SyntheticProblem:
    S = np.zeros((n_bf, n_bf), order="F")
    h = np.zeros((n_bf, n_bf), order="F")
    B = np.zeros(n_aux * (n_bf * (n_bf + 1) // 2))
    Cm = np.zeros((n_act, n_bf))
    _check_global(library().vbe_synthetic_problem(seed, n_bf, n_act, n_aux, scan_r, _ptr(S), _ptr(h), _ptr(B),
                                                  _ptr(Cm)))
    return SyntheticProblem(n_bf, n_act, n_aux, S, h, B, Cm)


BOHR_PER_ANGSTROM = 1.0 / 0.529177249  # src/input/loading/loader.cpp:42


def ao_integrals(basis: str, Z, xyz_bohr, eri: bool = True) -> dict:
    """AO integrals of a molecule from the embedded basis data (hydrogen, 6-31G family):
    S, T, V, h = T + V, packed-pair two-electron integrals, atom of each function, E_nuc."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
    n = C.c_int32()
    _check_global(library().vbe_gaussian_basis_size(basis.encode(), len(Z), _ptr(Z), C.byref(n)))
    nb = n.value
    npair = nb * (nb + 1) // 2
    out = dict(S=np.zeros((nb, nb), order="F"), T=np.zeros((nb, nb), order="F"), V=np.zeros((nb, nb), order="F"),
               eri=np.zeros((npair, npair), order="F") if eri else None, atom=np.zeros(nb, np.int32))
    e = C.c_double()
    _check_global(library().vbe_ao_integrals(basis.encode(), len(Z), _ptr(Z), _ptr(xyz), _ptr(out["S"]),
                                             _ptr(out["T"]), _ptr(out["V"]), _ptr(out["eri"]), _ptr(out["atom"]),
                                             C.byref(e)))
    out["h"] = out["T"] + out["V"]
    out["e_nuc"] = e.value
    out["n_bf"] = nb
    return out


def pivoted_cholesky(M: np.ndarray, tolerance: float = 1e-13) -> np.ndarray:
    """B (rank x n) with M ~= B^T B (pivoted Cholesky, largest remaining diagonal first)."""
    M = np.asfortranarray(M, dtype=np.float64)
    n = M.shape[0]
    B = np.zeros((n, n), order="F")
    r = C.c_int32()
    _check_global(library().vbe_pivoted_cholesky(n, _ptr(M), tolerance, n, C.byref(r), _ptr(B)))
    return np.asfortranarray(B[: r.value, :])


def h2_deck(distance_angstrom: float = 0.74, basis: str = "6-31G**") -> dict:
    """The H2 deck: two hydrogens on the z axis, two VB orbitals each expanded only on its own atom's
    basis functions (the $FRAG/$ORB assignment of an HAO deck), STR=FULL NAO=2 NAE=2 singlet
    (1 covalent + 2 ionic structures), exact two-electron integrals written as pivoted-Cholesky
    factors (rank <= npair, remaining diagonal <= 1e-13). Starting orbitals: the lowest root of the
    core Hamiltonian restricted to each atom's functions. Returns S, h, B (n_aux x npair), n_aux,
    e_nuc, support, C0 and the topology."""
    R = distance_angstrom * BOHR_PER_ANGSTROM
    ints = ao_integrals(basis, [1, 1], [[0.0, 0.0, 0.0], [0.0, 0.0, R]])
    n = ints["n_bf"]
    B = pivoted_cholesky(ints["eri"], 1e-13)
    support = np.zeros((2, n), np.uint8)
    C0 = np.zeros((2, n))
    for k in range(2):
        idx = np.flatnonzero(ints["atom"] == k)
        support[k, idx] = 1
        Sk = ints["S"][np.ix_(idx, idx)]
        hk = ints["h"][np.ix_(idx, idx)]
        L = np.linalg.cholesky(Sk)
        Li = np.linalg.inv(L)
        w, U = np.linalg.eigh(Li @ hk @ Li.T)
        C0[k, idx] = Li.T @ U[:, 0]
    topo = Topology.build("FULL", 2, 2, 1)
    return dict(S=ints["S"], h=ints["h"], B=np.ascontiguousarray(B.ravel(order="F")), n_aux=B.shape[0], n_bf=n,
                e_nuc=ints["e_nuc"], support=support, C0=C0, topo=topo, eri=ints["eri"], atom=ints["atom"])


class Context:
    """One HIP device + stream (cuda::Context precedent, src/cuda/memory/context.hpp:66-115)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._lib = library()
        self._h = C.c_void_p()
        self._children = weakref.WeakSet()  # problems, pools and schedulers: closed before the context
        _check_global(self._lib.vbe_context_create(device, hbm_budget_bytes, C.byref(self._h)))

    def error(self) -> str:
        buf = C.create_string_buffer(4096)
        self._lib.vbe_last_error(self._h, buf, 4096)
        return buf.value.decode()

    def check(self, rc: int):
        if rc != 0:
            raise VbeError(rc, self.error())

    def memory(self) -> tuple:
        """(bytes of device storage charged to this context, its HBM budget)."""
        used, budget = C.c_uint64(), C.c_uint64()
        self.check(self._lib.vbe_context_memory(self._h, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def set_pair_condition_limit(self, limit: float):
        """Pairs whose spin blocks have max|X| max|X^-1| >= limit (0..1e9) take the
        complete-pivoting path; 0 sends every pair there."""
        self.check(self._lib.vbe_set_pair_condition_limit(self._h, C.c_double(limit)))

    def set_eigensolver(self, eigensolver: int, energy_tolerance: float = 1e-7, residual_tolerance: float = 1e-3,
                        stationarity_tolerance: float = 1.0, reuse_roots: bool = True):
        """0: dense GEP; 1: Davidson in the metric's factored coordinates with the reference's physical
        stopping: Ritz energy stable to min(energy_tolerance, 1e-8), stationarity residual at most
        the square root of that, or stationarity_tolerance when smaller. Applies to problems and
        pools created afterwards; Problem.evaluate's `eigensolver` argument selects it per call."""
        self.check(self._lib.vbe_set_eigensolver(self._h, eigensolver, energy_tolerance, residual_tolerance,
                                                 stationarity_tolerance, int(reuse_roots)))
        self.eigensolver = eigensolver

    def set_states(self, states=(0,), weights=None):
        """Selected structure states (root indices, distinct, at most 8) and their weights (default
        equal; normalised) for problems and pools created afterwards: the energy and gradient are
        those of the weighted average of the states' energies."""
        st = np.ascontiguousarray(states, dtype=np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and len(w) != len(st):
            raise ValueError("one weight per state")
        self.check(self._lib.vbe_set_states(self._h, len(st), _ptr(st), _ptr(w)))
        self.n_states = len(st)

    def set_ri_precision(self, precision: int):
        """0: FP64 RI factors; 1: plain FP32 RI factors (FP64 accumulation; a comparison mode that
        misses the 1e-8 Eh bar); 2: compensated FP32 RI factors (FP32 plus an 8-bit correction,
        decoded exactly to FP64; within the bar). Applies to problems and pools created afterwards."""
        self.check(self._lib.vbe_set_ri_precision(self._h, precision))
        self.ri_precision = precision

    def close(self):
        if self._h:
            # schedulers first (they drive pools), then pools and problems
            for child in sorted(list(self._children), key=lambda c: not isinstance(c, Scheduler)):
                child.close()
            self._lib.vbe_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """Device-resident molecule-static data of one VBSCF problem."""

    def __init__(self, ctx: Context, n_bf: int, n_aux: int, topo: Topology, S: np.ndarray, h: np.ndarray,
                 B: np.ndarray, e_nuc: float = 0.0):
        self.ctx, self.topo = ctx, topo
        self.precision = getattr(ctx, "ri_precision", 0)
        self.eigensolver = getattr(ctx, "eigensolver", 0)
        self.n_states = getattr(ctx, "n_states", 1)
        self.n_bf, self.n_act, self.n_aux = n_bf, topo.n_act, n_aux
        dims = Dims(n_bf, topo.n_act, topo.n_alpha, topo.n_beta, n_aux)
        self._S = np.asfortranarray(S, dtype=np.float64)
        self._hm = np.asfortranarray(h, dtype=np.float64)
        self._B = np.ascontiguousarray(B, dtype=np.float64)
        self._view = topo.view()
        self._h = C.c_void_p()
        ctx._children.add(self)
        ctx.check(ctx._lib.vbe_problem_upload(ctx._h, C.byref(dims), _ptr(self._S), _ptr(self._hm), _ptr(self._B),
                                              C.byref(self._view), e_nuc, C.byref(self._h)))

    @classmethod
    def from_synthetic(cls, ctx: Context, sp: SyntheticProblem, topo: Topology, e_nuc: float = 0.0) -> "Problem":
        return cls(ctx, sp.n_bf, sp.n_aux, topo, sp.S, sp.h, sp.B, e_nuc)

    def _orbitals(self, Cm):
        Cm = np.ascontiguousarray(Cm, dtype=np.float64)
        if Cm.shape != (self.n_act, self.n_bf):
            raise ValueError(f"orbitals must have shape {(self.n_act, self.n_bf)}")
        return Cm

    def evaluate(self, Cm: np.ndarray, support: np.ndarray | None = None, gradient: bool = True):
        """Energy (and gradient wrt raw orbital slots) of orbitals Cm (n_act, n_bf)."""
        Cm = self._orbitals(Cm)
        sup = None if support is None else np.ascontiguousarray(support, dtype=np.uint8)
        e = C.c_double()
        g = np.zeros_like(Cm) if gradient else None
        self.ctx.check(self.ctx._lib.vbe_eval(self.ctx._h, self._h, _ptr(Cm), _ptr(sup), Flags(self.precision, self.eigensolver, int(gradient)),
                                              C.byref(e), None, _ptr(g)))
        return (e.value, g) if gradient else e.value

    def evaluate_forward(self, Cm: np.ndarray, support: np.ndarray | None = None) -> float:
        Cm = self._orbitals(Cm)
        sup = None if support is None else np.ascontiguousarray(support, dtype=np.uint8)
        e = C.c_double()
        self.ctx.check(self.ctx._lib.vbe_eval_forward(self.ctx._h, self._h, _ptr(Cm), _ptr(sup), C.byref(e)))
        return e.value

    def complete_gradient(self) -> np.ndarray:
        g = np.zeros((self.n_act, self.n_bf))
        self.ctx.check(self.ctx._lib.vbe_complete_gradient(self.ctx._h, self._h, _ptr(g)))
        return g

    def evaluate_device(self, C_ptr: int, grad_ptr: int, gradient: bool = True, support_ptr: int | None = None):
        """Evaluation with device-resident orbitals/gradient (raw device pointers)."""
        e = C.c_double()
        self.ctx.check(self.ctx._lib.vbe_eval_device(self.ctx._h, self._h, C.c_void_p(C_ptr),
                                                     None if support_ptr is None else C.c_void_p(support_ptr),
                                                     Flags(self.precision, self.eigensolver, int(gradient)), C.byref(e), C.c_void_p(grad_ptr)))
        return e.value

    def evaluate_device_noalloc(self, C_ptr: int, grad_ptr: int, gradient: bool = True,
                                support_ptr: int | None = None):
        """evaluate_device with the per-evaluation workspace in the library's static device scratch
        (vbe_eval_device_noalloc: no device allocation in the call)."""
        e = C.c_double()
        self.ctx.check(self.ctx._lib.vbe_eval_device_noalloc(self.ctx._h, self._h, C.c_void_p(C_ptr),
                                                             None if support_ptr is None else C.c_void_p(support_ptr),
                                                             Flags(self.precision, self.eigensolver, int(gradient)),
                                                             C.byref(e), C.c_void_p(grad_ptr)))
        return e.value

    def optimize(self, Cm: np.ndarray, support: np.ndarray | None = None, **options) -> dict:
        """Full P-LBFGS convergence from orbitals Cm; options override the VbScfOptimizerOptions
        defaults (max_iterations, history_size, gradient_tolerance, energy_tolerance,
        initial_step_size, minimum_step_size, armijo_constant)."""
        Cm = self._orbitals(Cm).copy()
        sup = None if support is None else np.ascontiguousarray(support, dtype=np.uint8)
        opt = OptimizerOptions()
        self.ctx._lib.vbe_optimizer_defaults(C.byref(opt))
        for k, v in options.items():
            if not hasattr(opt, k):
                raise TypeError(f"unknown optimizer option {k}")
            setattr(opt, k, v)
        res = OptimizerResult()
        g = np.zeros_like(Cm)
        trace = np.zeros(opt.max_iterations + 1)
        self.ctx.check(self.ctx._lib.vbe_optimize(self.ctx._h, self._h, _ptr(Cm), _ptr(sup), C.byref(opt), C.byref(res),
                                                  _ptr(g), _ptr(trace), len(trace)))
        out = {name: getattr(res, name) for name, _ in OptimizerResult._fields_}
        out["termination"] = TERMINATIONS.get(res.termination, res.termination)
        out.update(C=Cm, grad=g, trace=trace[:min(len(trace), res.iterations + 1)].copy())
        return out

    def intermediates(self) -> dict:
        ns, n = self.topo.n_structures, self.n_act
        np_ = n * (n + 1) // 2
        out = dict(H=np.zeros((ns, ns), order="F"), S=np.zeros((ns, ns), order="F"), coef=np.zeros(ns),
                   S_act=np.zeros((n, n), order="F"), h_act=np.zeros((n, n), order="F"),
                   g_packed=np.zeros(np_ * (np_ + 1) // 2))
        self.ctx.check(self.ctx._lib.vbe_last_intermediates(self.ctx._h, self._h, *(_ptr(out[k]) for k in
                                                                                     ("H", "S", "coef", "S_act",
                                                                                      "h_act", "g_packed"))))
        return out

    def state_energies(self) -> np.ndarray:
        """Total energies of the selected states at the last evaluation."""
        out = np.zeros(getattr(self, "n_states", 1))
        self.ctx.check(self.ctx._lib.vbe_last_state_energies(self.ctx._h, self._h, _ptr(out)))
        return out

    def phase_times(self) -> dict:
        t = PhaseTimes()
        self.ctx._lib.vbe_last_phase_times(self._h, C.byref(t))
        return {name: getattr(t, name) for name, _ in PhaseTimes._fields_}

    def singular_pairs(self) -> int:
        c = C.c_int64()
        self.ctx._lib.vbe_last_singular_pairs(self._h, C.byref(c))
        return c.value

    def device_bytes(self) -> int:
        b = C.c_uint64()
        self.ctx._lib.vbe_problem_device_bytes(self._h, C.byref(b))
        return b.value

    def close(self):
        if self._h:
            self.ctx._lib.vbe_problem_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pool:
    """Resident problem slots sharing one structure topology, evaluated in batched launches
    (one kernel per phase for all listed slots) and served through submit/wait tickets."""

    def __init__(self, ctx: Context, topo: Topology, capacity: int):
        self.ctx, self.topo, self.capacity = ctx, topo, capacity
        self.n_act = topo.n_act
        self._view = topo.view()
        self._h = C.c_void_p()
        self.n_bf = [0] * capacity
        self.n_states = getattr(ctx, "n_states", 1)
        ctx._children.add(self)
        ctx.check(ctx._lib.vbe_pool_create(ctx._h, C.byref(self._view), topo.n_act, topo.n_alpha, topo.n_beta,
                                           capacity, C.byref(self._h)))

    def load(self, slot: int, n_bf: int, n_aux: int, S, h, B, e_nuc: float = 0.0):
        S = np.asfortranarray(S, dtype=np.float64)
        h = np.asfortranarray(h, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        rc = self.ctx._lib.vbe_pool_load(self._h, slot, n_bf, n_aux, _ptr(S), _ptr(h), _ptr(B), e_nuc)
        if rc == 6:  # VBE_ERR_MEMORY: the engine leaves the slot unloaded
            self.n_bf[slot] = 0
        self.ctx.check(rc)
        self.n_bf[slot] = n_bf

    def load_synthetic(self, slot: int, sp: "SyntheticProblem", e_nuc: float = 0.0):
        self.load(slot, sp.n_bf, sp.n_aux, sp.S, sp.h, sp.B, e_nuc)

    def _sizes(self, slots):
        return [self.n_act * self.n_bf[s] for s in slots]

    def evaluate_device(self, slots, C_ptr: int, grad_ptr: int, gradient: bool = True) -> np.ndarray:
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        e = np.zeros(len(sl))
        self.ctx.check(self.ctx._lib.vbe_pool_eval_device(self._h, len(sl), _ptr(sl), C.c_void_p(C_ptr),
                                                          int(gradient), _ptr(e),
                                                          C.c_void_p(grad_ptr) if gradient else None))
        return e

    def submit(self, slots, orbitals, gradient: bool = True) -> int:
        """orbitals: list of (n_act, n_bf_slot) arrays in slot order."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(o, dtype=np.float64).ravel()
                                                    for o in orbitals]))
        t = C.c_int32()
        self.ctx.check(self.ctx._lib.vbe_pool_submit(self._h, len(sl), _ptr(sl), _ptr(flat), int(gradient),
                                                     C.byref(t)))
        self._pending = getattr(self, "_pending", {})
        self._pending[t.value] = (list(sl), gradient)
        return t.value

    def poll(self, ticket: int) -> bool:
        """True once the ticket's device work has finished (wait() then does not block)."""
        d = C.c_int32()
        self.ctx.check(self.ctx._lib.vbe_pool_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int):
        slots, gradient = self._pending.pop(ticket)
        e = np.zeros(len(slots))
        sizes = self._sizes(slots)
        g = np.zeros(sum(sizes)) if gradient else None
        self.ctx.check(self.ctx._lib.vbe_pool_wait(self._h, ticket, _ptr(e), _ptr(g)))
        grads = None
        if gradient:
            grads, at = [], 0
            for s, sz in zip(slots, sizes):
                grads.append(g[at:at + sz].reshape(self.n_act, self.n_bf[s]))
                at += sz
        return e, grads

    def evaluate(self, slots, orbitals, gradient: bool = True):
        return self.wait(self.submit(slots, orbitals, gradient))

    def optimize(self, slots, orbitals, supports=None, **options) -> list:
        """P-LBFGS for several slots at once (lockstep, one batched launch per phase); orbitals and
        supports: lists of (n_act, n_bf_slot) arrays in slot order. Returns one dict per slot."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(o, dtype=np.float64).ravel()
                                                    for o in orbitals]))
        sup = None if supports is None else np.ascontiguousarray(
            np.concatenate([np.ascontiguousarray(m, dtype=np.uint8).ravel() for m in supports]))
        opt = OptimizerOptions()
        self.ctx._lib.vbe_optimizer_defaults(C.byref(opt))
        for k, v in options.items():
            if not hasattr(opt, k):
                raise TypeError(f"unknown optimizer option {k}")
            setattr(opt, k, v)
        res = (OptimizerResult * len(sl))()
        g = np.zeros_like(flat)
        self.ctx.check(self.ctx._lib.vbe_pool_optimize(self._h, len(sl), _ptr(sl), _ptr(flat), _ptr(sup), C.byref(opt),
                                                       C.cast(res, C.c_void_p), _ptr(g)))
        out, at = [], 0
        for j, s in enumerate(sl):
            sz = self.n_act * self.n_bf[s]
            d = {name: getattr(res[j], name) for name, _ in OptimizerResult._fields_}
            d["termination"] = TERMINATIONS.get(d["termination"], d["termination"])
            d.update(C=flat[at:at + sz].reshape(self.n_act, self.n_bf[s]).copy(),
                     grad=g[at:at + sz].reshape(self.n_act, self.n_bf[s]).copy())
            out.append(d)
            at += sz
        return out

    def state_energies(self, slot: int) -> np.ndarray:
        """Total energies of the selected states of a slot at its last evaluation (pools created
        with more than one selected state)."""
        out = np.zeros(self.n_states)
        self.ctx.check(self.ctx._lib.vbe_pool_state_energies(self._h, slot, _ptr(out)))
        return out

    def phase_times(self) -> dict:
        t = PhaseTimes()
        self.ctx._lib.vbe_pool_phase_times(self._h, C.byref(t))
        return {name: getattr(t, name) for name, _ in PhaseTimes._fields_}

    def device_bytes(self) -> int:
        b = C.c_uint64()
        self.ctx._lib.vbe_pool_device_bytes(self._h, C.byref(b))
        return b.value

    def autotune(self, slots, orbitals, repeats: int = 2) -> dict:
        """Time the candidate pair-kernel block targets on these slots and keep the fastest for
        later launches; returns the choice and the per-candidate forward/backward pair-phase ms."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(o, dtype=np.float64).ravel() for o in orbitals]))
        f, b = C.c_int32(), C.c_int32()
        table = np.zeros(2 * 6)
        self.ctx.check(self.ctx._lib.vbe_pool_autotune(self._h, len(sl), _ptr(sl), _ptr(flat), int(repeats),
                                                       C.byref(f), C.byref(b), _ptr(table)))
        cands = [256, 512, 1024, 2048, 4096, 8192]
        return dict(fwd_target=f.value, bwd_target=b.value,
                    table={c: dict(fwd_ms=table[2 * k], bwd_ms=table[2 * k + 1]) for k, c in enumerate(cands)})

    def unload(self, slot: int):
        """Release one slot's device storage (its bytes return to the context budget)."""
        self.ctx.check(self.ctx._lib.vbe_pool_unload(self._h, int(slot)))
        self.n_bf[int(slot)] = 0

    def close(self):
        if self._h:
            self.ctx._lib.vbe_pool_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scheduler:
    """Native request scheduler over resident pool slots (vbe_server_*): a dispatcher thread in the
    engine batches queued E(+grad) requests per pool continuously, on the context's shared
    streams. `submit` queues a request for a loaded slot; `collect` returns completed requests as
    dicts with the energy, status, batch size and times (CLOCK_MONOTONIC seconds, comparable with
    time.perf_counter); gradients land in the array passed to submit."""

    def __init__(self, ctx: Context, max_batch: int = 64, max_inflight_per_pool: int = 4):
        self.ctx = ctx
        self._h = C.c_void_p()
        self._grads = {}  # request id -> gradient array kept alive until collected
        ctx._children.add(self)
        ctx.check(ctx._lib.vbe_server_create(ctx._h, max_batch, max_inflight_per_pool, C.byref(self._h)))
        self._buf = (RequestInfo * 256)()

    def submit(self, pool: "Pool", slot: int, Cm: np.ndarray, grad: np.ndarray | None = None,
               arrival: float = 0.0) -> int:
        Cm = np.ascontiguousarray(Cm, dtype=np.float64)
        if Cm.size != pool.n_act * pool.n_bf[slot]:
            raise ValueError("orbitals do not match the slot's n_act x n_bf")
        if grad is not None and (grad.dtype != np.float64 or not grad.flags.c_contiguous or grad.size != Cm.size):
            raise ValueError("gradient buffer must be a C-contiguous float64 array of the orbitals' size")
        rid = C.c_int64()
        self.ctx.check(self.ctx._lib.vbe_server_submit(self._h, pool._h, slot, _ptr(Cm), int(grad is not None),
                                                       _ptr(grad), float(arrival), C.byref(rid)))
        if grad is not None:
            self._grads[rid.value] = grad
        return rid.value

    def collect(self, wait_us: int = 0) -> list:
        n = C.c_int32()
        self.ctx.check(self.ctx._lib.vbe_server_collect(self._h, len(self._buf), int(wait_us), self._buf,
                                                        C.byref(n)))
        out = []
        for i in range(n.value):
            r = self._buf[i]
            out.append(dict(id=r.id, status=r.status, batch_size=r.batch_size, energy=r.energy,
                            t_arrival=r.t_arrival, t_submit=r.t_submit, t_dispatch=r.t_dispatch,
                            t_complete=r.t_complete, device_ms=r.device_ms, grad=self._grads.pop(r.id, None)))
        return out

    def counters(self) -> dict:
        c = ServerCounters()
        self.ctx.check(self.ctx._lib.vbe_server_stats(self._h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in ServerCounters._fields_}

    def close(self):
        if self._h:
            self.ctx._lib.vbe_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
