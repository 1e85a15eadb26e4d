/*
 * xmvb_amd.h — C ABI of the MI355X VBSCF energy+gradient engine.
 *
 * The engine evaluates one VBSCF energy and its analytic orbital gradient ("E+grad") for a
 * structure set of nonorthogonal spin-coupled VB structures on one AMD Instinct MI355X
 * (gfx950): RI AO->orbital integral transform, determinant-pair Lowdin/cofactor matrix
 * elements, structure H/S assembly, the generalised eigenproblem H c = E S c and the
 * cofactor-path backward, all on the device in FP64.
 *
 * Conventions (identical to the reference program Tao-Xia/xmvb-cpp):
 *   - matrices are column-major;
 *   - orbitals are orbital-major: value of orbital k on basis function mu at [k*n_bf + mu]
 *     (VbScfInput::orbital_preparation_input.orbital_value_table with a dense index table,
 *      src/vbscf/orbitals/preparation/input.hpp:13-167);
 *   - RI factors B are n_aux x n_bf(n_bf+1)/2, column-major, AO pair (mu,nu) at
 *     hi(hi+1)/2+lo (RiAoFactorization::metric_whitened_ao_pair_factors,
 *     src/integrals/ao/ri.hpp:15-31; PairIndex, src/core/pair_index.hpp:15-55), so that
 *     (mu nu|la si) = sum_P B[P,(mu nu)] B[P,(la si)];
 *   - spin strings are occupation bitmasks over active orbitals (bit i = orbital i).
 *
 * Every function returns a vbe_status; no C++ exception crosses this boundary
 * (the reference throws std::invalid_argument/runtime_error and maps them to exit code 2,
 * src/cli/main.cpp:17-20). The text of the last error of a context is available through
 * vbe_last_error(); errors raised before a context exists go to vbe_global_error().
 *
 * Threading: the problems and pools of a context launch on the context's four shared HIP streams
 * (completion is tracked per launch with events). Calls on one pool are serialised by the pool's
 * lock, so the request scheduler's dispatcher thread (vbe_server_*) and the caller's loads and
 * unloads may interleave; calls on one problem must not overlap. Contexts are independent.
 */
#ifndef XMVB_AMD_H
#define XMVB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vbe_status {
    VBE_OK = 0,
    VBE_ERR_ARGUMENT = 1,     /* invalid argument / inconsistent sizes        (std::invalid_argument) */
    VBE_ERR_NUMERIC = 2,      /* orbital normalisation or S_act factorisation failed (preparer.cpp:78-83,399-401) */
    VBE_ERR_METRIC = 3,       /* structure metric not positive definite (dsygvd info > 0) */
    VBE_ERR_DEVICE = 4,       /* HIP runtime error */
    VBE_ERR_UNSUPPORTED = 5,  /* option combination this engine rejects (cuvbscf.cpp:38-50 precedent) */
    VBE_ERR_MEMORY = 6,       /* HBM budget exceeded (std::length_error precedent) */
    VBE_ERR_TOPOLOGY = 7      /* structure generation/expansion failed */
} vbe_status;

/* Electronic methods selected by bare $CTRL tokens. Mirrors ElectronicMethod
 * (src/input/deck/metadata.hpp:19) plus HIPVBSCF, the token that selects this engine. */
typedef enum vbe_method {
    VBE_METHOD_VBSCF = 0,
    VBE_METHOD_CUVBSCF = 1,
    VBE_METHOD_CASSCF = 2,
    VBE_METHOD_DFT = 3,
    VBE_METHOD_HIPVBSCF = 4
} vbe_method;

typedef enum vbe_precision {
    VBE_PRECISION_FP64 = 0,     /* everything FP64 */
    VBE_PRECISION_FP32_RI = 1,  /* RI factors B stored in plain FP32 (half the memory and transform
                                   traffic), the transform accumulating in FP64. The FP32 rounding of B
                                   moves energies by up to ~1e-7 Eh: this mode does NOT meet the 1e-8 Eh
                                   bar and is kept only as a comparison point */
    VBE_PRECISION_FP32C_RI = 2  /* compensated FP32 storage of B: hi = fp32(B) plus an 8-bit correction
                                   k of its rounding error, B = hi + k 2^(e(hi) - 158) (e(hi) the biased
                                   exponent of hi), 5 bytes per element and at most 2^-31 relative error;
                                   decoded exactly to FP64, the transform accumulating in FP64. Meets
                                   the 1e-8 Eh bar. Determinant pairs, structure matrices, metric
                                   Cholesky, eigensolve and gradient stay FP64 in every mode */
} vbe_precision;

typedef enum vbe_eigensolver {
    VBE_EIGENSOLVER_DENSE = 0,    /* EIGENSOLVER=DENSE (linalg/eigensolver.cpp:365-387) */
    VBE_EIGENSOLVER_DAVIDSON = 1  /* lowest root by Davidson in the metric's factored coordinates
                                     (evaluator.cpp:339-391, davidson_core.hpp:271-730) */
} vbe_eigensolver;

typedef struct vbe_context vbe_context;
typedef struct vbe_problem vbe_problem;
typedef struct vbe_pool vbe_pool;
typedef struct vbe_topology vbe_topology;

typedef struct vbe_dims {
    int32_t n_bf;        /* AO basis functions                                     */
    int32_t n_act;       /* active VB orbitals (all electrons active, no core)     */
    int32_t n_alpha;     /* active alpha electrons                                 */
    int32_t n_beta;      /* active beta electrons                                  */
    int32_t n_aux;       /* RI factor rows                                         */
} vbe_dims;

/* Signed structure -> determinant incidence (StructureExpansion, src/vbscf/structures/expansion/types.hpp:390-423). */
typedef struct vbe_topology_view {
    int32_t n_structures;
    int32_t n_alpha_strings;
    const uint64_t* alpha_masks;      /* ascending                                       */
    int32_t n_beta_strings;
    const uint64_t* beta_masks;       /* ascending                                       */
    int32_t n_determinants;
    const int32_t* det_alpha;         /* alpha string index per determinant              */
    const int32_t* det_beta;          /* beta string index per determinant               */
    const int64_t* term_offsets;      /* n_determinants + 1 CSR offsets                  */
    const int32_t* term_structure;    /* structure index per term                        */
    const double* term_coefficient;   /* signed coefficient per term                     */
} vbe_topology_view;

typedef struct vbe_flags {
    int32_t precision;      /* vbe_precision   */
    int32_t eigensolver;    /* vbe_eigensolver */
    int32_t want_gradient;  /* 0: energy only (a trial evaluation), 1: energy + gradient */
} vbe_flags;

typedef struct vbe_phase_times {
    /* device time per phase of the last evaluation, milliseconds (HIP events on the evaluation's
     * stream). ri_one_index_ms isolates the one-index kernel inside transform_ms;
     * pair_backward_kernel_ms the determinant-pair kernels inside pair_backward_ms. The last three
     * isolate single pair kernels: the main forward pair kernel, and the two backward pair kernels
     * (F blocks with the S_act/h_act adjoints; the packed-kernel adjoint). */
    float orbital_prep_ms, transform_ms, ri_one_index_ms, pair_forward_ms, structure_ms, eigensolve_ms,
        pair_backward_ms, pair_backward_kernel_ms, gradient_tail_ms, total_ms;
    float pair_forward_kernel_ms, pair_backward_field_ms, pair_backward_adjoint_ms;
} vbe_phase_times;

/* ---- $CTRL method flag ------------------------------------------------------------------ */

/* Parse the bare tokens of a $CTRL line into a method. Mirrors apply_ctrl_flag
 * (src/input/deck/model.cpp:360-381): VBSCF/CUVBSCF/CASSCF/DFT plus HIPVBSCF; conflicting
 * methods fail with VBE_ERR_ARGUMENT; TBVBSCF and VBPDFT are rejected as unsupported. Tokens
 * that are not method tokens are ignored. *method is VBE_METHOD_VBSCF when no token is given. */
int vbe_parse_ctrl_method(const char* ctrl_line, int32_t* method);

/* Reject option combinations HIPVBSCF does not implement, in the manner of run_cuvbscf
 * (src/cli/cuvbscf.cpp:38-50): eigensolver DENSE or DAVIDSON, two-electron integrals given as
 * RI-shaped factors (ri=1: INT=RI factors, or exact integrals factorised by vbe_pivoted_cholesky),
 * 1 to 8 averaged states (more than one with the dense eigensolver), no inactive (core) electrons. Returns VBE_OK or VBE_ERR_UNSUPPORTED and
 * writes the reason to vbe_global_error(). */
int vbe_check_hipvbscf_options(int32_t eigensolver, int32_t ri, int32_t n_states, int32_t n_inactive_electrons);

/* ---- AO integrals for small decks (host) ---------------------------------------------------
 * Replaces the libcint materialised provider (src/integrals/libcint/materialized_provider.cpp) for
 * the basis data embedded in the library: hydrogen in 6-31G / 6-31G* / 6-31G** (basis/6-31gss.gbs,
 * Cartesian functions, per atom: s(3G), s, then p x, y, z for **). McMurchie-Davidson integrals. */
int vbe_gaussian_basis_size(const char* basis, int32_t n_atoms, const int32_t* Z, int32_t* n_bf);
/* S, T, V (n_bf x n_bf column-major, any may be NULL), two-electron integrals over packed pairs
 * eri_packed[PQ + npair RS] (npair = n_bf(n_bf+1)/2, pair = hi(hi+1)/2+lo; NULL to skip), the atom
 * of each function and the nuclear repulsion energy. Coordinates in bohr (the deck loader converts
 * angstrom with 1/0.529177249, src/input/loading/loader.cpp:42). */
int vbe_ao_integrals(const char* basis, int32_t n_atoms, const int32_t* Z, const double* xyz_bohr, double* S,
                     double* T, double* V, double* eri_packed, int32_t* atom_of_function, double* e_nuc);
/* Pivoted Cholesky M ~= B^T B of a symmetric positive semidefinite n x n matrix (B: max_rank x n,
 * column-major; *rank rows used), stopping when the largest remaining diagonal is <= tolerance:
 * turns exact two-electron integrals into the engine's RI-shaped factors. */
int vbe_pivoted_cholesky(int32_t n, const double* M, double tolerance, int32_t max_rank, int32_t* rank, double* B);

/* ---- topology (host) ---------------------------------------------------------------------- */

/* Closed-form structure count for a structure-class keyword ("COV", "FULL", "ION", "COV ION(1)"),
 * count_generated_structures (src/input/loading/generated_structures.cpp:621-658). */
int vbe_count_structures(const char* structure_class, int32_t n_act, int32_t n_electrons, int32_t multiplicity,
                         uint64_t* count);

/* Generate Rumer/Young structures (generated_structures.cpp:673-716) and expand them into signed
 * determinants (StructureExpander::expand, src/vbscf/structures/expansion/expander.cpp:287-558).
 * sizes[0..4] = n_structures, n_alpha_strings, n_beta_strings, n_determinants, n_terms. */
int vbe_topology_build(const char* structure_class, int32_t n_act, int32_t n_electrons, int32_t multiplicity,
                       vbe_topology** out, int64_t sizes[5]);
/* Copy the arrays of a built topology into caller buffers sized from sizes[]; raw_structures
 * (optional, n_structures x n_electrons labels, 1-based) may be NULL. */
int vbe_topology_copy(const vbe_topology* topo, uint64_t* alpha_masks, uint64_t* beta_masks, int32_t* det_alpha,
                      int32_t* det_beta, int64_t* term_offsets, int32_t* term_structure, double* term_coefficient,
                      int32_t* raw_structures);
int vbe_topology_view_of(const vbe_topology* topo, vbe_topology_view* view);
void vbe_topology_free(vbe_topology* topo);

/* ---- synthetic problems (host) -------------------------------------------------------------- */

/* Seeded synthetic RI problem: mt19937_64(seed) with a Marsaglia polar normal sampler; draws, in
 * this order and column-major: A ~ N(0,1) (n_bf^2) -> S = I + 0.1 A A^T / n_bf; X ~ N(0, var 0.5)
 * -> h = (X + X^T)/2 - 2 I; for each aux row P: G_P ~ N(0, var 1/n_bf) -> B_P = (G_P + G_P^T)/2
 * packed; C ~ N(0,1)/sqrt(n_bf) (orbital-major n_act x n_bf). scan_r >= 0 scales the blocks of h
 * and B that couple the first and second half of the basis by exp(-0.5 scan_r) (bond-stretch scan
 * geometry parameter); pass a negative scan_r for none. Output buffers: S, h (n_bf^2), B
 * (n_aux * n_bf(n_bf+1)/2), C (n_act * n_bf). */
int vbe_synthetic_problem(uint64_t seed, int32_t n_bf, int32_t n_act, int32_t n_aux, double scan_r, double* S,
                          double* h, double* B, double* C);

/* ---- device engine -------------------------------------------------------------------------- */

/* Create a context on HIP device `device` with an HBM budget in bytes (0 = 90% of free memory,
 * the cuda::Context precedent, src/cuda/memory/context.hpp:66-115); every device allocation of
 * the context's problems and pools is charged to the budget (VBE_ERR_MEMORY when exceeded). */
int vbe_context_create(int32_t device, uint64_t hbm_budget_bytes, vbe_context** out);
void vbe_context_destroy(vbe_context* ctx);
/* Copy the last error of a context into buf (NUL-terminated, truncated to len). */
int vbe_last_error(const vbe_context* ctx, char* buf, size_t len);
/* Determinant pairs are evaluated one per lane with row-pivoted inversion when both spin blocks
 * satisfy max|X| max|X^-1| < limit; the other pairs go through the complete-pivoting path that
 * applies DeterminantOverlapResolver's FullPivLU rank rule (src/vbscf/determinants/algebra/
 * overlap.cpp:50-110) directly. Any limit in [0, 1e9] keeps results within rounding of the
 * reference (below 1e9 a row-pivoted block is provably full rank under that rule); 0 sends every
 * pair through the complete-pivoting path. Default 1e5: worse-conditioned blocks then follow the
 * reference's pivot order, which keeps their rounding close to it. Applies to later evaluations.
 * No counterpart in the reference, which has a single (CPU) pair path. */
int vbe_set_pair_condition_limit(vbe_context* ctx, double limit);
/* Storage precision of the RI factors for problems and pools created after the call
 * (vbe_precision; default VBE_PRECISION_FP64). Evaluations must then pass the same precision in
 * vbe_flags. No counterpart in the reference, whose RI factors are FP64 only. */
int vbe_set_ri_precision(vbe_context* ctx, int32_t precision);

/* Structure eigensolver of problems and pools created afterwards (vbe_eval / vbe_eval_device select
 * it per call through vbe_flags.eigensolver, which then stays with the problem) and the Davidson
 * accuracy. Davidson runs in the metric's factored coordinates y = L^T c, S = L L^T, as the
 * reference's solve_factored_structure_roots (src/vbscf/derivatives/gradient/active_space/
 * evaluator.cpp:339-391) with its physical stopping (StructureSolveAccuracy::
 * configure_physical_stopping, src/vbscf/core/contracts/eigensolver.hpp:51-56): Ritz energy stable
 * to min(energy_tolerance, 1e-8) Eh and stationarity residual ||A y - E y|| / ||y|| at most the
 * square root of that. stationarity_tolerance can only tighten the latter (pass 1.0 for the
 * reference's bound); residual_tolerance (the reference's gradient_tolerance, 1e-3) is validated
 * and kept, the physical stopping does not use it. reuse_roots (1) starts each slot from its
 * previous eigenvector (evaluator.cpp:233-246). */
int vbe_set_eigensolver(vbe_context* ctx, int32_t eigensolver, double energy_tolerance, double residual_tolerance,
                        double stationarity_tolerance, int32_t reuse_roots);
const char* vbe_global_error(void);

/* Selected structure states of problems and pools created afterwards: n_states (1..8) distinct
 * root indices (ascending eigenvalue; NULL: 0..n_states-1) and weights (NULL: equal), normalised
 * as mcscf::normalize_weights (src/mcscf/states/weights.cpp:7-33: non-negative, positive sum).
 * The evaluation's energy is then E11 + sum_s w_s lambda_s + E_nuc and its gradient that of this
 * average (workflow/evaluator.cpp:22-32,65-111; the deck's NSTATE / WSTATE(1), equal weights,
 * src/input/deck/model.cpp:167-209). Default: the lowest root alone. More than one state, or a
 * state other than the lowest, takes the dense eigensolver (the Davidson path rejects it). */
int vbe_set_states(vbe_context* ctx, int32_t n_states, const int32_t* state_indices, const double* weights);

/* Upload the molecule-static data of one problem: AO overlap S and core Hamiltonian h
 * (n_bf x n_bf), RI factors B, the structure topology and E_nuc. Host buffers are copied; the
 * problem stays resident in HBM until vbe_problem_release. Replaces the one-time upload of the
 * CUDA precedent cuda::solve_scf (src/cuda/vb/scf.hpp:41-46). */
int vbe_problem_upload(vbe_context* ctx, const vbe_dims* dims, const double* S_ao, const double* h_ao,
                       const double* B_packed, const vbe_topology_view* topo, double e_nuc, vbe_problem** out);
void vbe_problem_release(vbe_problem* prob);
/* Device bytes held by the problem (static data + evaluation workspace). */
int vbe_problem_device_bytes(const vbe_problem* prob, uint64_t* bytes);

/* One evaluation from host orbitals: C_raw (n_act x n_bf, orbital-major), optional support mask
 * (same layout, nonzero = coefficient slot present; NULL = dense). Writes E_total and, when
 * flags.want_gradient, grad_raw (n_act x n_bf; zero on absent slots). E_states (optional) receives
 * the lowest n_str eigenvalues + E_nuc. Host-side replacement of OrbitalGradientEvaluator::evaluate
 * (src/vbscf/derivatives/gradient/orbital/evaluator.cpp:198-418) with EIGENSOLVER=DENSE. */
int vbe_eval(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support, vbe_flags flags,
             double* E_total, double* E_states, double* grad_raw);

/* Same with device-resident orbitals and gradient (device pointers, n_act*n_bf doubles each;
 * d_support may be NULL). Work is enqueued on the context stream and the call returns after the
 * energy is known on the host. */
int vbe_eval_device(vbe_context* ctx, vbe_problem* prob, const double* d_C_raw, const uint8_t* d_support,
                    vbe_flags flags, double* E_total, double* d_grad_raw);

/* No reference counterpart: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same inputs, outputs and results
 * as vbe_eval_device, but the whole per-evaluation workspace (orbital intermediates, one-index
 * factors, packed kernel, determinant and structure matrices, eigensolver work, backward partial
 * rows, pair lists) lives in a static device array of VBE_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, and the call makes no device allocation. Problems whose workspace exceeds that
 * maximum are rejected with VBE_ERR_MEMORY (the medium and batching-form sizes fit, n_act = 14 with
 * 429 structures does not). Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU race. Davidson does not start from the problem's previous root here, and the call discards a
 * retained forward state (vbe_complete_gradient then needs a new vbe_eval_forward). */
#define VBE_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int vbe_eval_device_noalloc(vbe_context* ctx, vbe_problem* prob, const double* d_C_raw, const uint8_t* d_support,
                            vbe_flags flags, double* E_total, double* d_grad_raw);

/* Two-phase contract of VbScfObjective (src/vbscf/optimization/objective/function.cpp:89-275):
 * vbe_eval_forward evaluates the energy of trial orbitals and keeps the forward state on the
 * device; vbe_complete_gradient turns the retained forward state into the gradient. A new
 * forward discards the previous one. */
int vbe_eval_forward(vbe_context* ctx, vbe_problem* prob, const double* C_raw, const uint8_t* support,
                     double* E_total);
int vbe_complete_gradient(vbe_context* ctx, vbe_problem* prob, double* grad_raw);

/* Structure matrices, eigenvector and active-space intermediates of the last evaluation
 * (any pointer may be NULL): H, S (n_str^2), coef (n_str), S_act, h_act (n_act^2), packed
 * active two-electron kernel (np(np+1)/2, np = n_act(n_act+1)/2). */
int vbe_last_intermediates(vbe_context* ctx, vbe_problem* prob, double* H, double* S, double* coef, double* S_act,
                           double* h_act, double* g_packed);

/* Device time per phase of the last evaluation. */
int vbe_last_phase_times(const vbe_problem* prob, vbe_phase_times* times);
/* Total energies E11 + lambda_s + E_nuc of the selected states at the problem's last evaluation
 * (VbScfResult::selected_state_total_energies, workflow/evaluator.cpp:98-106); E_states[n_states]. */
int vbe_last_state_energies(vbe_context* ctx, vbe_problem* prob, double* E_states);
/* The same for a pool slot after its last evaluation (pools with more than one selected state). */
int vbe_pool_state_energies(vbe_pool* pool, int32_t slot, double* E_states);

/* Count of determinant pairs whose overlap block was rank deficient in the last evaluation. */
int vbe_last_singular_pairs(const vbe_problem* prob, int64_t* count);

/* ---- batched and served evaluation: pools of resident problems ------------------------------
 * No reference counterpart: the reference evaluates one problem per process. A pool holds
 * `capacity` resident problem slots sharing one structure topology (same active space and
 * structure set, e.g. the points of a bond-stretch scan or the problems of one size bucket);
 * each slot has its own n_bf, n_aux, S, h, B and E_nuc. One evaluation launch processes any set
 * of slots with one kernel per phase, and two streams let a second submission overlap the first
 * (continuous batching). Orbital/gradient arrays are the per-slot n_act x n_bf blocks of the
 * listed slots, concatenated in list order. */
int vbe_pool_create(vbe_context* ctx, const vbe_topology_view* topo, int32_t n_act, int32_t n_alpha, int32_t n_beta,
                    int32_t capacity, vbe_pool** out);
void vbe_pool_destroy(vbe_pool* pool);
/* Upload (or replace) the static data of one slot; allocates the slot's workspace. */
int vbe_pool_load(vbe_pool* pool, int32_t slot, int32_t n_bf, int32_t n_aux, const double* S_ao, const double* h_ao,
                  const double* B_packed, double e_nuc);
int vbe_pool_device_bytes(const vbe_pool* pool, uint64_t* bytes);
/* Release the device storage of one slot (its budget returns to the context); the slot can be
 * loaded again later. Rejected (VBE_ERR_ARGUMENT) while a pending submission holds the slot.
 * With vbe_context_memory this is what an admission policy needs to keep the resident problems
 * under the context's HBM budget. */
int vbe_pool_unload(vbe_pool* pool, int32_t slot);
/* Bytes of device storage currently charged to the context, and its budget. */
int vbe_context_memory(const vbe_context* ctx, uint64_t* in_use, uint64_t* budget);
/* Synchronous batched evaluation with device-resident orbitals / gradients; energies[count] on the host. */
int vbe_pool_eval_device(vbe_pool* pool, int32_t count, const int32_t* slots, const double* d_C,
                         int32_t want_gradient, double* energies, double* d_grad);
/* Asynchronous batched evaluation from host orbitals: returns a ticket at once; at most two
 * submissions are in flight and a slot may belong to only one of them. */
int vbe_pool_submit(vbe_pool* pool, int32_t count, const int32_t* slots, const double* C, int32_t want_gradient,
                    int32_t* ticket);
/* Non-blocking completion test of a ticket (the vbe_poll of a serving loop): *done = 1 once its
 * device work has finished, after which vbe_pool_wait returns without blocking. */
int vbe_pool_poll(vbe_pool* pool, int32_t ticket, int32_t* done);
/* Wait for a ticket; energies[count] and (if requested at submit) the concatenated gradients. */
int vbe_pool_wait(vbe_pool* pool, int32_t ticket, double* energies, double* grads);
/* Device time per phase of the pool's most recently completed launch. */
int vbe_pool_phase_times(const vbe_pool* pool, vbe_phase_times* times);
/* Launch-shape autotuner (no reference counterpart): evaluates the listed slots (host orbitals C,
 * as for vbe_pool_submit; E+grad) at each candidate pair-kernel block target per launch
 * (256 ... 8192, VBE_AUTOTUNE_CANDIDATES of them), `repeats` times each, and keeps for the pool the
 * target with the shortest forward pair phase and, separately, the shortest backward pair phase
 * (device times of the pool's phase events). The choice applies to later launches of the pool;
 * results differ between shapes only by the summation order of the adjoint partial sums.
 * table (optional, VBE_AUTOTUNE_CANDIDATES x 2): best forward and backward ms per candidate. */
#define VBE_AUTOTUNE_CANDIDATES 6
int vbe_pool_autotune(vbe_pool* pool, int32_t count, const int32_t* slots, const double* C, int32_t repeats,
                      int32_t* fwd_target, int32_t* bwd_target, double* table);

/* ---- full convergence: projected L-BFGS on device-resident state -----------------------------
 * Options of VbScfOptimizerOptions (src/vbscf/optimization/driver/options.hpp:15-108) that the
 * P-LBFGS backend reads; vbe_optimizer_defaults fills the reference defaults. */
typedef struct vbe_optimizer_options {
    int32_t max_iterations;     /* accepted iterations (2000)                                */
    int32_t history_size;       /* secant pairs kept (100)                                   */
    double gradient_tolerance;  /* projected-gradient norm (1e-3)                            */
    double energy_tolerance;    /* |E_k - E_k-1| (1e-7)                                      */
    double initial_step_size;   /* first Armijo trial step, capped at 1 (1.0)                */
    double minimum_step_size;   /* Armijo gives up below this step (1e-7)                    */
    double armijo_constant;     /* sufficient-decrease constant (1e-4)                       */
} vbe_optimizer_options;

/* Termination reasons of run_plbfgs_backend (src/vbscf/optimization/backends/plbfgs.cpp:20-226). */
typedef enum vbe_termination {
    VBE_TERM_MAX_ITERATIONS = 0,       /* iteration limit reached                              */
    VBE_TERM_INITIAL_TOLERANCE = 1,    /* "plbfgs_initial_tolerance"                           */
    VBE_TERM_CONVERGED = 2,            /* "plbfgs_projected_gradient_tolerance"                */
    VBE_TERM_LINE_SEARCH_FAILED = 3,   /* "plbfgs_line_search_failed"                          */
    VBE_TERM_LINE_SEARCH_STALLED = 4,  /* "plbfgs_line_search_stalled"                         */
    VBE_TERM_NON_DESCENT = 5,          /* "plbfgs_non_descent_direction"                       */
    VBE_TERM_EMPTY_SPACE = 6           /* "nonredundant_space_empty"                           */
} vbe_termination;

typedef struct vbe_optimizer_result {
    double energy;                  /* E_total at the returned orbitals                        */
    double gradient_norm;           /* projected-gradient norm there (the convergence measure) */
    int32_t iterations;             /* accepted iterations                                     */
    int32_t energy_evaluations;     /* forward evaluations (Armijo trials included)            */
    int32_t gradient_evaluations;   /* backward evaluations                                    */
    int32_t converged;
    int32_t termination;            /* vbe_termination                                         */
} vbe_optimizer_result;

void vbe_optimizer_defaults(vbe_optimizer_options* options);

/* Converge the orbitals of one problem: run_plbfgs_backend (plbfgs.cpp:20-226) with the Armijo
 * backtracking of try_armijo_backtracking_nonredundant_direction (globalization/line_search.cpp:
 * 33-95) and transported secant pairs (preconditioners/transported_lbfgs.cpp). Orbitals, gradients
 * and the secant history stay on the device; the host only reads scalars (energies, directional
 * derivatives, norms) to take the line-search decisions. C_raw (n_act x n_bf, in/out): start point
 * in, converged orbitals (each normalised to c^T S c = 1) out; support as in vbe_eval; grad_raw
 * (optional) receives the gradient at the returned orbitals; energy_trace (optional, trace_capacity
 * entries) the energy after each accepted iteration, starting with the initial point. */
int vbe_optimize(vbe_context* ctx, vbe_problem* prob, double* C_raw, const uint8_t* support,
                 const vbe_optimizer_options* options, vbe_optimizer_result* result, double* grad_raw,
                 double* energy_trace, int32_t trace_capacity);

/* The same optimisation for `count` resident slots of a pool at once (no reference counterpart:
 * the reference converges one problem per process). The slots advance in lockstep, each phase of
 * every slot's iteration (chart, search direction, Armijo trial, secant append) being one batched
 * launch, and every round of trials one launch sequence of the pool; trials are evaluated with
 * their gradient. Each slot follows exactly the iterates it would follow alone. C_raw, support
 * (or NULL) and grad_raw (or NULL): concatenated n_act x n_bf blocks of the listed slots in list
 * order; results[count]. The pool must have no submission in flight. */
int vbe_pool_optimize(vbe_pool* pool, int32_t count, const int32_t* slots, double* C_raw, const uint8_t* support,
                      const vbe_optimizer_options* options, vbe_optimizer_result* results, double* grad_raw);

/* ---- request scheduler (no reference counterpart: the reference evaluates one problem per
 * process, src/cli/run.cpp; SURVEY's asynchronous vbe_submit / vbe_poll for serving) ----------
 * A native dispatcher thread batches E(+grad) requests for loaded pool slots continuously: when a
 * pool has a free launch lane, every queued request of that pool whose slot is not in flight (up to
 * max_batch, one gradient flag per launch) goes into one launch. Pools' lanes run on their context's
 * shared streams (at most four per context). Times are CLOCK_MONOTONIC seconds (Python's
 * time.perf_counter on Linux). The caller keeps each request's gradient buffer valid until the
 * request is collected, and must not unload a slot that has requests queued or in flight. */
typedef struct vbe_server vbe_server;

typedef struct vbe_request_info {
    int64_t id;
    int32_t status;      /* VBE_OK or the launch's status for this slot */
    int32_t batch_size;  /* requests in the launch that served it */
    double energy;       /* E_total */
    double t_arrival;    /* as passed to vbe_server_submit (its submit time when <= 0) */
    double t_submit;     /* accepted into the queue */
    double t_dispatch;   /* launched */
    double t_complete;   /* results copied out */
    double device_ms;    /* device time of its launch (forward and backward, HIP events) */
} vbe_request_info;

typedef struct vbe_server_counters {
    int64_t queued, in_flight_launches, launches, launched_requests, completed, uncollected;
} vbe_server_counters;

/* max_inflight_per_pool: launches of one pool in flight at once, 1..4 (its launch lanes). */
int vbe_server_create(vbe_context* ctx, int32_t max_batch, int32_t max_inflight_per_pool, vbe_server** out);
/* Completes every queued and in-flight request, stops the dispatcher and frees the server. */
void vbe_server_destroy(vbe_server* server);
/* Queue one request for slot `slot` of `pool`: C (n_act x n_bf of the slot, host, copied at the
 * call), want_gradient, grad_out (n_act x n_bf host buffer, written on completion; NULL without
 * gradient), arrival (the request's arrival time for latency accounting, <= 0: now). */
int vbe_server_submit(vbe_server* server, vbe_pool* pool, int32_t slot, const double* C, int32_t want_gradient,
                      double* grad_out, double arrival, int64_t* id);
/* Completed requests, oldest first, up to max_count; waits up to wait_us microseconds for one when
 * none is ready. *count may be 0. */
int vbe_server_collect(vbe_server* server, int32_t max_count, int32_t wait_us, vbe_request_info* out,
                       int32_t* count);
int vbe_server_stats(vbe_server* server, vbe_server_counters* counters);

#ifdef __cplusplus
}
#endif

#endif /* XMVB_AMD_H */
