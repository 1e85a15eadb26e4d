#!/usr/bin/env python3
"""VBSCF energy+gradient throughput on one MI355X.

Workload (medium config): synthetic RI problems n_bf=120, n_aux=360, n_act=10 VB orbitals with
10 electrons (5 alpha / 5 beta), 42 covalent Rumer structures expanded to 252 determinants;
orbitals resident in HBM. An evaluation is one full energy + analytic orbital gradient (orbital
normalisation, RI transform, determinant pairs, structure H/S, generalised eigensolve, backward)
through the engine's C ABI, FP64 throughout.

Three legs are timed:
  * plain: one problem per launch sequence (the plain FP64 GPU path);
  * batched (value): a pool of --batch resident problems (seeds 0..K-1) evaluated by one launch
    sequence per step; one step = K evaluations;
  * served: Poisson arrivals (--serve-rate requests/s for --serve-seconds) over a 64-point
    dissociation scan (inter-fragment h and B blocks scaled by exp(-0.5 r), r in [0.5, 4.0]),
    continuously batched onto two in-flight submissions (submit / poll / wait); request latency
    is completion minus arrival, host orbitals in and gradients out (PCIe included); run once with
    E+grad requests only and once with a second Poisson stream of full-convergence requests
    (P-LBFGS, converged in lockstep on their own pool and host thread) sharing the GPU;
  * converge: --converge full-convergence requests (projected L-BFGS on device-resident state,
    reference optimizer defaults) at scan points, reported as latency per request and per
    iteration next to the plain evaluation time.

Prints one JSON line: batched throughput (E+grad evaluations/s) beside the plain path's, request
p50/p99 latency of the served leg, the roofline of the dominant pair kernel of the batched leg (its
duration from HIP events recorded on the engine stream during the timed region), parity against
the FP64 CPU oracle and the oracle's own throughput on the host cores (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xmvb-cpp_amd"))
sys.path.insert(0, ROOT)

METRIC = "VBSCF E+grad evals/s served vs plain FP64 GPU; p50/p99 latency; |ΔE|≤1e-8 Eh"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "medium_pmc_traffic.json")

MEDIUM = dict(seed=0, n_bf=120, n_act=10, n_el=10, n_aux=360, structures="COV")


def det_pair_flops(m_alpha: int, m_beta: int) -> float:
    """Algorithmic FP64 flops of one ordered determinant pair's forward matrix element:
    per spin block LU (2/3 m^3) + inverse (2 m^3) + same-spin two-electron terms (6 C(m,2)^2),
    plus the opposite-spin coupling (2 m_alpha^2 m_beta^2)."""
    def block(m):
        c2 = m * (m - 1) / 2
        return 2.0 / 3.0 * m ** 3 + 2.0 * m ** 3 + 6.0 * c2 * c2
    return block(m_alpha) + block(m_beta) + 2.0 * m_alpha ** 2 * m_beta ** 2


def lane_pairs_evaluated(topo) -> int:
    """Determinant pairs the lane kernels evaluate per problem: one of each unordered pair {d, e}
    and, when every determinant has a spin-flip partner (alpha and beta strings exchanged), one of
    each couple {d, e} / {flip d, flip e} (equal blocks with the spins exchanged)."""
    nd = len(topo.det_alpha)
    unordered = nd * (nd + 1) // 2
    if topo.n_alpha != topo.n_beta:
        return unordered
    key = {(int(topo.alpha_masks[topo.det_alpha[d]]), int(topo.beta_masks[topo.det_beta[d]])): d for d in range(nd)}
    flip = [key.get((int(topo.beta_masks[topo.det_beta[d]]), int(topo.alpha_masks[topo.det_alpha[d]])))
            for d in range(nd)]
    if any(f is None for f in flip):
        return unordered
    # pairs that are their own partner: both determinants fixed by the flip, or {d, flip d}
    fixed = sum(1 for d in range(nd) if flip[d] == d)
    self_partner = fixed * (fixed + 1) // 2 + (nd - fixed) // 2
    return (unordered + self_partner) // 2


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return float("nan")
    k = (len(s) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def cpu_baseline(cfg, sp, topo, gpu_energy, gpu_grad, seconds: float):
    """The FP64 oracle (C++ restatement of the reference CPU path, OpenMP over all host cores),
    timed like the reference harness (warm-up 1, then repeated in-process evaluations)."""
    import numpy as np

    import oracle

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    args = (cfg["n_bf"], cfg["n_aux"], topo, sp.S, sp.h, sp.B, sp.C)
    ref = oracle.evaluate(*args)  # warm-up, also the parity reference
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.evaluate(*args)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds or n >= 1000:
            break
    parity = dict(abs_dE=abs(gpu_energy - ref["energy"]), grad_max_abs_err=float(np.abs(gpu_grad - ref["grad"]).max()),
                  oracle_energy=ref["energy"])
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    base = dict(value=n / dt, unit="E+grad evals/s", cores=cores, kind="port", cpu_model=cpu_model,
                sample=f"{n} full E+grad evaluations of the same medium-config problem in {dt:.1f} s "
                       f"(oracle/vb_oracle.cpp, OpenMP {cores} threads)")
    return base, parity


def serve(vb, ctx, scans, rate: float, seconds: float, max_batch: int, replicas: int = 1, seed: int = 7,
          converge_rate: float = 0.0, converge_iterations: int = 150):
    """Poisson arrivals of E+grad requests over the points of one or more bond-stretch scans,
    served by the engine's request loop (xmvb_amd.serving.Server: continuous batching per pool,
    admission under the context's HBM budget). scans: list of (name, topology, [SyntheticProblem
    per point]); each point is registered `replicas` times so that requests for one point can run
    in both of a pool's launches. With converge_rate > 0 a second Poisson stream of
    full-convergence requests (P-LBFGS from the point's start orbitals) goes to a
    ConvergenceWorker on its own host thread, sharing the device and the budget."""
    import numpy as np

    from xmvb_amd import serving

    def loader(sp):
        return lambda: (sp.n_bf, sp.n_aux, sp.S, sp.h, sp.B)

    n_points = sum(len(sps) for _, _, sps in scans)
    server = serving.Server(ctx, capacity=max(len(sps) for _, _, sps in scans) * replicas, max_batch=max_batch)
    for si, (_, topo, sps) in enumerate(scans):
        for p, sp in enumerate(sps):
            for k in range(replicas):
                server.register((si, p, k), topo, loader(sp))
    # direct reference energies: every point of a scan in one batched launch of a pool of its own
    e_ref, alone_ms = {}, {}
    for si, (_, topo, sps) in enumerate(scans):
        ref_pool = vb.Pool(ctx, topo, capacity=len(sps))
        for p, sp in enumerate(sps):
            ref_pool.load_synthetic(p, sp)
        e, _ = ref_pool.evaluate(list(range(len(sps))), [sp.C for sp in sps], gradient=False)
        # one problem of this size evaluated alone (E+grad, host orbitals in and gradient out, as the
        # served requests): the yardstick of the served latencies
        ts = []
        for rep in range(6):
            t = time.perf_counter()
            ref_pool.evaluate([rep % len(sps)], [sps[rep % len(sps)].C], gradient=True)
            ts.append((time.perf_counter() - t) * 1e3)
        alone_ms[si] = float(np.median(ts[1:]))
        ref_pool.close()
        for p in range(len(sps)):
            e_ref[(si, p)] = e[p]
    # admission before the timed window: one request per registered problem
    for key in list(server.res.problems):
        si, p, _k = key
        server.submit(key, scans[si][2][p].C)
    server.drain()
    loads_warm = server.loads

    worker = None
    if converge_rate > 0:
        worker = serving.ConvergenceWorker(ctx, capacity=max(len(sps) for _, _, sps in scans),
                                           max_iterations=converge_iterations)
        for si, (_, topo, sps) in enumerate(scans):
            for p, sp in enumerate(sps):
                worker.register((si, p), topo, loader(sp))
        worker.preload()  # resident before the window, like the served problems
        crng = np.random.default_rng(seed + 1)
        n_conv = max(1, int(converge_rate * seconds))
        c_arrivals = np.cumsum(crng.exponential(1.0 / converge_rate, n_conv))
        c_scan = crng.integers(0, len(scans), n_conv)
        c_pick = [(int(si), int(crng.integers(len(scans[si][2])))) for si in c_scan]

    rng = np.random.default_rng(seed)
    n_req = max(1, int(rate * seconds))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    scan_of = rng.integers(0, len(scans), n_req)
    pick = [(int(si), int(rng.integers(len(scans[si][2])))) for si in scan_of]
    reqs, launches0 = [], server.launches
    nxt, nxt_c, done = 0, 0, 0
    t0 = time.perf_counter()
    if worker is not None:
        worker.start()
    while done < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            si, p = pick[nxt]
            k = min(range(replicas), key=lambda j: server.res.problems[(si, p, j)].pending)
            reqs.append(server.submit((si, p, k), scans[si][2][p].C, arrival=t0 + arrivals[nxt]))
            nxt += 1
        if worker is not None:
            while nxt_c < n_conv and c_arrivals[nxt_c] <= now:
                si, p = c_pick[nxt_c]
                worker.submit((si, p), scans[si][2][p].C, arrival=t0 + c_arrivals[nxt_c])
                nxt_c += 1
        # the engine's dispatcher launches on its own thread; this loop only submits on time and
        # collects, waiting inside the engine until the next arrival
        gap = arrivals[nxt] - (time.perf_counter() - t0) if nxt < n_req else 2e-3
        done += len(server.pump(wait_us=int(max(0.0, min(gap, 2e-3)) * 1e6)))
    makespan = max(r.done for r in reqs) - t0
    converge = None
    if worker is not None:
        while nxt_c < n_conv:  # the rest of the convergence stream, at its own arrival times
            now = time.perf_counter() - t0
            if c_arrivals[nxt_c] <= now:
                si, p = c_pick[nxt_c]
                worker.submit((si, p), scans[si][2][p].C, arrival=t0 + c_arrivals[nxt_c])
                nxt_c += 1
            else:
                time.sleep(min(c_arrivals[nxt_c] - now, 5e-3))
        worker.finish()
        cl = [r.latency * 1e3 for r in worker.completed]
        its = [r.result["iterations"] for r in worker.completed]
        term = {}
        for r in worker.completed:
            term[r.result["termination"]] = term.get(r.result["termination"], 0) + 1
        converge = dict(requests=n_conv, offered_rate=converge_rate, max_iterations=converge_iterations,
                        latency_ms=dict(p50=percentile(cl, 0.5), p99=percentile(cl, 0.99), max=max(cl)),
                        mean_iterations=float(np.mean(its)), batches=len(worker.rounds),
                        mean_batch=float(np.mean(worker.rounds)), terminations=term,
                        makespan_s=time.perf_counter() - t0)
        worker.close()
    lat = [r.latency * 1e3 for r in reqs]
    queue_ms = [r.queue_ms for r in reqs]
    device_ms = [r.device_ms for r in reqs]
    host_ms = [r.latency * 1e3 - r.queue_ms - r.device_ms for r in reqs]
    assert all(r.status == 0 for r in reqs)
    err = max(abs(r.energy - e_ref[(r.key[0], r.key[1])]) for r in reqs)
    per_scan = {}
    for si, (name, topo, sps) in enumerate(scans):
        ls = [r.latency * 1e3 for r in reqs if r.key[0] == si]
        if ls:
            p50, p99 = percentile(ls, 0.5), percentile(ls, 0.99)
            per_scan[name] = dict(requests=len(ls), n_act=topo.n_act, n_bf=sps[0].n_bf, p50=p50, p99=p99,
                                  alone_ms=alone_ms[si], p50_over_alone=p50 / alone_ms[si],
                                  p99_over_alone=p99 / alone_ms[si])
    used, budget = ctx.memory()
    launches = server.launches - launches0
    out = dict(requests=n_req, offered_rate=rate, throughput=n_req / makespan, seconds=makespan,
               latency_ms=dict(p50=percentile(lat, 0.5), p99=percentile(lat, 0.99), max=max(lat)),
               latency_parts_ms=dict(
                   queue=dict(p50=percentile(queue_ms, 0.5), p99=percentile(queue_ms, 0.99)),
                   device=dict(p50=percentile(device_ms, 0.5), p99=percentile(device_ms, 0.99)),
                   host=dict(p50=percentile(host_ms, 0.5), p99=percentile(host_ms, 0.99)),
                   note="queue: arrival to launch (waiting for a free lane of the pool or for the slot); "
                        "device: the launch on the device (HIP events); host: the rest (copies, stream "
                        "order behind other launches, completion polling)"),
               mean_batch=n_req / max(1, launches), launches=launches,
               max_abs_dE_vs_direct=float(err),
               scan=f"{len(scans)} scan(s) x {len(scans[0][2])} points r in [0.5, 4.0], {replicas} resident "
                    f"replica(s) per point ({n_points * replicas} problems)",
               admission=dict(budget_gib=budget / 2**30, resident_gib=used / 2**30, loads=server.loads,
                              loads_before_timing=loads_warm, evictions=server.evictions),
               per_scan=per_scan, io="host orbitals in, host gradients out (PCIe included)",
               convergence_requests=converge)
    server.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed batched steps (K evaluations each)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=512, help="resident problems evaluated per batched step")
    ap.add_argument("--distinct", type=int, default=128,
                    help="distinct seeded problems behind the batch (slot s holds problem s mod this)")
    ap.add_argument("--plain-steps", type=int, default=300, help="timed single-problem evaluations")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate of the served leg")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="arrival window of the served leg (0: skip)")
    ap.add_argument("--serve-converge-rate", type=float, default=4.0,
                    help="Poisson rate of full-convergence requests mixed into the served leg (0: none)")
    ap.add_argument("--serve-converge-iterations", type=int, default=150,
                    help="iteration cap of a served convergence request")
    ap.add_argument("--serve-mixed", type=int, default=1,
                    help="also serve the mixed-size scans (one per n_act in {4,...,12}) under a 200 GB budget")
    ap.add_argument("--precision-steps", type=int, default=5,
                    help="timed evaluations of the precision form, FP64 and FP32 RI (0: skip)")
    ap.add_argument("--mixed-steps", type=int, default=30, help="timed batched steps with FP32 RI factors (0: skip)")
    ap.add_argument("--converge", type=int, default=3, help="full-convergence requests timed (0: skip)")
    ap.add_argument("--converge-iterations", type=int, default=500, help="iteration cap per convergence request")
    ap.add_argument("--converge-batch", type=int, default=128,
                    help="scan-point problems converged together in one pool (0: skip)")
    ap.add_argument("--converge-batch-iterations", type=int, default=100, help="iteration cap in the batched run")
    ap.add_argument("--mixed-batch", type=int, default=4096,
                    help="problems of the uneven-size batching workload evaluated per step (0: skip)")
    ap.add_argument("--mixed-steps-batching", type=int, default=3, help="timed steps of the batching workload")
    ap.add_argument("--autotune", type=int, default=1, help="autotune the pair-kernel launch shape of each pool")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="bounded CPU-baseline sample length")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle parity")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the engine has no "
              f"multi-GPU path in this build).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import xmvb_amd as vb

    if args.copy_calibration_mib > 0:
        # a plain copy kernel of known size: FETCH_SIZE/WRITE_SIZE of this dispatch calibrate the
        # counters used for the roofline traffic figure
        src = torch.ones(args.copy_calibration_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cfg = MEDIUM
    sp = vb.synthetic_problem(cfg["seed"], cfg["n_bf"], cfg["n_act"], cfg["n_aux"])
    topo = vb.Topology.build(cfg["structures"], cfg["n_act"], cfg["n_el"], 1)
    ctx = vb.Context(0)
    dev = torch.device("cuda:0")
    phases = ("orbital_prep_ms", "transform_ms", "ri_one_index_ms", "pair_forward_ms", "structure_ms", "eigensolve_ms",
              "pair_backward_ms", "pair_backward_kernel_ms", "gradient_tail_ms", "total_ms", "pair_forward_kernel_ms",
              "pair_backward_field_ms", "pair_backward_adjoint_ms")

    # ---- plain leg: one problem at a time ----
    prob = vb.Problem.from_synthetic(ctx, sp, topo)
    c_dev = torch.from_numpy(sp.C.copy()).to(dev)
    g_dev = torch.zeros_like(c_dev)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        prob.evaluate_device(c_dev.data_ptr(), g_dev.data_ptr())
    torch.cuda.synchronize()
    lat_plain = []
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        s = time.perf_counter()
        energy = prob.evaluate_device(c_dev.data_ptr(), g_dev.data_ptr())
        lat_plain.append((time.perf_counter() - s) * 1e3)
    torch.cuda.synchronize()
    plain_elapsed = time.perf_counter() - t0
    grad = g_dev.cpu().numpy()
    plain = dict(value=args.plain_steps / plain_elapsed, ms_per_eval=plain_elapsed * 1e3 / args.plain_steps,
                 latency_ms=dict(p50=percentile(lat_plain, 0.5), p99=percentile(lat_plain, 0.99)),
                 phase_ms={k: round(v, 5) for k, v in prob.phase_times().items()})
    prob.close()

    # ---- batched leg: K resident problems per launch sequence ----
    K = args.batch
    pool = vb.Pool(ctx, topo, capacity=K)
    orbitals = []
    # distinct problems generated on host threads (the generator releases the GIL); each slot holds
    # its own device copy of problem s mod D
    from concurrent.futures import ThreadPoolExecutor
    D = max(1, min(K, args.distinct))
    with ThreadPoolExecutor(max_workers=16) as ex:
        distinct = [sp] + list(ex.map(lambda s: vb.synthetic_problem(cfg["seed"] + s, cfg["n_bf"], cfg["n_act"],
                                                                     cfg["n_aux"]), range(1, D)))
    for s in range(K):
        sps = distinct[s % D]
        pool.load_synthetic(s, sps)
        orbitals.append(sps.C)
    cb_dev = torch.from_numpy(np.concatenate([o.ravel() for o in orbitals])).to(dev)
    gb_dev = torch.zeros_like(cb_dev)
    slots = list(range(K))
    # launch shape: the pool's autotuner picks the pair-kernel block targets on these slots
    tuned = pool.autotune(slots, orbitals, repeats=2) if args.autotune else None
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        pool.evaluate_device(slots, cb_dev.data_ptr(), gb_dev.data_ptr())
    torch.cuda.synchronize()
    sums = dict.fromkeys(phases, 0.0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        energies = pool.evaluate_device(slots, cb_dev.data_ptr(), gb_dev.data_ptr())
        lat.append((time.perf_counter() - s) * 1e3)
        pt = pool.phase_times()
        for k in phases:
            sums[k] += pt[k]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    mean = {k: v / args.steps for k, v in sums.items()}
    batch_grad0 = gb_dev[: cfg["n_act"] * cfg["n_bf"]].cpu().numpy().reshape(cfg["n_act"], cfg["n_bf"])
    batch_consistency = dict(abs_dE_slot0_vs_plain=abs(float(energies[0]) - energy),
                             grad_max_diff_slot0_vs_plain=float(np.abs(batch_grad0 - grad).max()))
    pool.close()

    # ---- rooflines of the single kernels timed on the engine stream (K problems per launch) ----
    n_det = topo.n_determinants
    lane_pairs = lane_pairs_evaluated(topo)
    nbp = cfg["n_bf"] * (cfg["n_bf"] + 1) // 2
    f_det = K * n_det * n_det * det_pair_flops(topo.n_alpha, topo.n_beta)  # ordered pairs, forward
    # (duration, roofline bound, algorithmic work, unit, what actually limits the kernel). The FP64
    # determinant-pair and eigensolver kernels are priced against the FP64 peak ("mfma": on MI355X
    # the FP64 vector and matrix peaks are the same 78.6 TFLOP/s) but are limited by something else,
    # named in "limiter": LDS gathers of the np x np kernel and FP64 LDS atomics for the pair
    # kernels, serial recurrences and block barriers for the eigensolver.
    kernels = {
        # backward ~ 2 F_det (two kernels over the same pairs): F blocks + S_act/h_act adjoints,
        # and the packed-kernel adjoint, one F_det each
        "k_pairs_lane_bwd_f": (mean["pair_backward_field_ms"], "mfma", f_det, "TFLOP/s",
                               "LDS gathers of the kernel and FP64 LDS atomics (VALU about a third busy)"),
        "k_pairs_lane_bwd_g": (mean["pair_backward_adjoint_ms"], "mfma", f_det, "TFLOP/s",
                               "FP64 LDS atomics of the same-spin adjoint"),
        "k_pairs_lane_fwd": (mean["pair_forward_kernel_ms"], "mfma", f_det, "TFLOP/s",
                             "LDS gathers of the kernel"),
        "k_ri_one_index": (mean["ri_one_index_ms"], "hbm",
                           K * 8.0 * cfg["n_aux"] * (nbp + cfg["n_act"] * cfg["n_bf"]), "GB/s", "HBM"),
        "k_gep_wave": (mean["eigensolve_ms"], "mfma", K * 10.0 * topo.n_structures ** 3, "TFLOP/s",
                       "latency: per-step LDS round trips, wave reductions and serial recurrences"),
    }
    traffic_table = {}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic_table = json.load(f).get("kernels", {})

    def roof(name):
        ms, bound, work, unit, limiter = kernels[name]
        if unit == "GB/s":
            achieved, peak = work / (ms * 1e-3) / 1e9, HBM_PEAK_GBS
        else:
            achieved, peak = work / (ms * 1e-3) / 1e12, FP64_PEAK_TFLOPS
        tr = traffic_table.get(name)
        traffic = None
        if tr and tr.get("problems_per_launch"):
            traffic = (tr["fetch_bytes"] + tr["write_bytes"]) * K / tr["problems_per_launch"]
        out = dict(kernel=name, bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak,
                   traffic=traffic, kernel_ms=ms, algorithmic_work=work, problems_per_launch=K, limiter=limiter,
                   timing="HIP events recorded on the evaluating stream around this kernel, mean over the timed steps")
        if traffic is not None:
            out["traffic_source"] = ("profiles/medium_pmc_traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                     "of this kernel in a separate run of the same launch shape (not measured "
                                     "in this run), scaled to this launch's problem count")
        if name.startswith("k_pairs_lane"):
            # the algorithmic work counts every ordered determinant pair; the kernel evaluates one
            # unordered pair of each spin-flip couple, and one evaluation does the work of the
            # pair's two ordered directions (about 2 F_det), so it executes a fraction
            # 2 evaluated / n_det^2 of the algorithmic flops
            out["pairs_algorithmic_per_problem"] = n_det * n_det
            out["pairs_evaluated_per_problem"] = lane_pairs
            out["executed_frac"] = achieved * 2.0 * lane_pairs / (n_det * n_det) / peak
        return out

    dominant = max(kernels, key=lambda k: kernels[k][0])
    roofline = roof(dominant)
    others = [roof(k) for k in kernels if k != dominant]

    # ---- the same batch with compensated FP32 RI factors (FP32 + an 8-bit correction of its rounding,
    # decoded exactly to FP64; FP64 arithmetic throughout): the mixed-precision engine ----
    mixed = None
    if args.mixed_steps > 0:
        ctx32 = vb.Context(0)
        ctx32.set_ri_precision(2)
        pool32 = vb.Pool(ctx32, topo, capacity=K)
        e64 = np.array([float(x) for x in energies])
        g64 = gb_dev.cpu().numpy()
        for s in range(K):
            pool32.load_synthetic(s, distinct[s % D])
        gb32 = torch.zeros_like(cb_dev)
        for _ in range(args.warmup):
            pool32.evaluate_device(slots, cb_dev.data_ptr(), gb32.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.mixed_steps):
            e32 = pool32.evaluate_device(slots, cb_dev.data_ptr(), gb32.data_ptr())
        torch.cuda.synchronize()
        el32 = time.perf_counter() - t0
        g32 = gb32.cpu().numpy()
        mixed = dict(value=K * args.mixed_steps / el32, unit="E+grad evals/s",
                     ri_storage="compensated fp32 (fp32 + 8-bit correction of its rounding error, 5 bytes per "
                                "element, at most 2^-31 relative), FP64 arithmetic",
                     problems_compared=K, distinct_problems=D,
                     max_abs_dE_vs_fp64=float(np.abs(np.asarray(e32) - e64).max()),
                     grad_max_err_vs_fp64=float(np.abs(g32 - g64).max()),
                     one_index_ms=pool32.phase_times()["ri_one_index_ms"],
                     phase_ms={k: round(v, 5) for k, v in pool32.phase_times().items()},
                     speedup_vs_plain=K * args.mixed_steps / el32 / plain["value"])
        pool32.close()
        ctx32.close()

    # ---- the precision form: one problem at n_bf=400, n_aux=1200, n_act=14 (429 structures), FP64
    # and with FP32 RI factors; |dE| and the gradient max-norm error of the FP32 run against FP64 ----
    precision = None
    if args.precision_steps > 0:
        psp = vb.synthetic_problem(0, 400, 14, 1200)
        ptopo = vb.Topology.build("COV", 14, 14, 1)
        runs = {}
        for label, ri in (("fp64", 0), ("fp32c_ri", 2)):
            pctx = vb.Context(0)
            if ri:
                pctx.set_ri_precision(ri)
            pprob = vb.Problem.from_synthetic(pctx, psp, ptopo)
            pe, pg = pprob.evaluate(psp.C)  # warm-up
            times = []
            for _ in range(args.precision_steps):
                t0 = time.perf_counter()
                pe, pg = pprob.evaluate(psp.C)
                times.append((time.perf_counter() - t0) * 1e3)
            runs[label] = dict(energy=float(pe), grad=pg, ms=percentile(times, 0.5),
                               phase_ms={k: round(v, 3) for k, v in pprob.phase_times().items()})
            pprob.close()
            pctx.close()
        precision = dict(workload="precision form: n_bf=400 n_aux=1200 n_act=14 (7a/7b), 429 covalent Rumer "
                                  "structures, 3432 determinants, one problem, host orbitals in",
                         fp64_ms_per_eval=runs["fp64"]["ms"], fp32c_ri_ms_per_eval=runs["fp32c_ri"]["ms"],
                         energy_fp64=runs["fp64"]["energy"],
                         abs_dE_fp32c_ri_vs_fp64=abs(runs["fp32c_ri"]["energy"] - runs["fp64"]["energy"]),
                         grad_max_err_fp32c_ri_vs_fp64=float(np.abs(runs["fp32c_ri"]["grad"] - runs["fp64"]["grad"]).max()),
                         phase_ms_fp64=runs["fp64"]["phase_ms"])
        del psp, runs

    served = None
    if args.serve_seconds > 0:
        rs = [0.5 + 3.5 * p / 63 for p in range(64)]
        scan = [vb.synthetic_problem(1000 + p, cfg["n_bf"], cfg["n_act"], cfg["n_aux"], scan_r=r)
                for p, r in enumerate(rs)]
        medium_scans = [("medium", topo, scan)]
        served = serve(vb, ctx, medium_scans, args.serve_rate, args.serve_seconds, max_batch=K, replicas=4)
        if args.serve_converge_rate > 0:
            served["with_convergence_requests"] = serve(
                vb, ctx, medium_scans, args.serve_rate, args.serve_seconds, max_batch=K, replicas=4,
                converge_rate=args.serve_converge_rate, converge_iterations=args.serve_converge_iterations)
        del scan, medium_scans
    served_mixed = None
    if args.serve_seconds > 0 and args.serve_mixed:
        # the serving form: one 64-point scan per active space n_act in {4,...,12} (n_bf drawn from
        # [40, 200], n_aux = 3 n_bf, one seed per scan, r scales the off-diagonal blocks), E+grad
        # requests spread over all scans and sizes, full-convergence requests beside them, all under
        # a 200 GB HBM budget
        from concurrent.futures import ThreadPoolExecutor
        srng = np.random.default_rng(2000)
        rs = [0.5 + 3.5 * p / 63 for p in range(64)]
        specs = [(na, int(srng.integers(40, 201))) for na in (4, 6, 8, 10, 12)]
        with ThreadPoolExecutor(max_workers=16) as ex:
            mixed_scans = [(f"n_act={na} n_bf={nb}", vb.Topology.build("COV", na, na, 1),
                            list(ex.map(lambda r, na=na, nb=nb: vb.synthetic_problem(30000 + na, nb, na, 3 * nb,
                                                                                     scan_r=r), rs)))
                           for na, nb in specs]
        ctx_s = vb.Context(0, hbm_budget_bytes=200 * 10**9)
        served_mixed = serve(vb, ctx_s, mixed_scans, args.serve_rate, args.serve_seconds, max_batch=64,
                             replicas=2, converge_rate=args.serve_converge_rate,
                             converge_iterations=args.serve_converge_iterations)
        ctx_s.close()
        del mixed_scans

    converge = None
    if args.converge > 0:
        runs = []
        for j in range(args.converge):
            r = 1.0 + 1.5 * j / max(1, args.converge - 1)
            spr = vb.synthetic_problem(cfg["seed"], cfg["n_bf"], cfg["n_act"], cfg["n_aux"], r)
            pr = vb.Problem.from_synthetic(ctx, spr, topo)
            pr.optimize(spr.C, max_iterations=2)  # first-use allocations outside the timing
            t0 = time.perf_counter()
            res = pr.optimize(spr.C, max_iterations=args.converge_iterations)
            dt = time.perf_counter() - t0
            runs.append(dict(r=r, iterations=res["iterations"], termination=res["termination"],
                             energy_start=float(res["trace"][0]), energy=res["energy"],
                             gradient_norm=res["gradient_norm"], forward_evaluations=res["energy_evaluations"],
                             backward_evaluations=res["gradient_evaluations"], ms=dt * 1e3,
                             ms_per_iteration=dt * 1e3 / max(1, res["iterations"])))
            pr.close()
        converge = dict(runs=runs, max_iterations=args.converge_iterations,
                        options="reference optimizer defaults: gradient_tolerance 1e-3, energy_tolerance 1e-7, "
                                "history 100, Armijo 1e-4",
                        plain_eval_ms=plain["ms_per_eval"],
                        iterations_per_s_single=sum(x["iterations"] for x in runs) / (sum(x["ms"] for x in runs) * 1e-3))
        if args.converge_batch > 0:
            # scan points r in [0.5, 4.0], several replicas each, converged in lockstep in one pool
            KB = args.converge_batch
            poolc = vb.Pool(ctx, topo, capacity=KB)
            starts = []
            for slot in range(KB):
                r = 0.5 + 3.5 * (slot % 64) / 63.0
                spr = vb.synthetic_problem(cfg["seed"] + slot // 64, cfg["n_bf"], cfg["n_act"], cfg["n_aux"], r)
                poolc.load_synthetic(slot, spr)
                starts.append(spr.C)
            poolc.optimize(list(range(KB)), starts, max_iterations=1)  # first-use allocations
            t0 = time.perf_counter()
            resb = poolc.optimize(list(range(KB)), starts, max_iterations=args.converge_batch_iterations)
            dtb = time.perf_counter() - t0
            its = sum(x["iterations"] for x in resb)
            term = {}
            for x in resb:
                term[x["termination"]] = term.get(x["termination"], 0) + 1
            converge["batched"] = dict(
                problems=KB, max_iterations=args.converge_batch_iterations, seconds=dtb, accepted_iterations=its,
                forward_evaluations=sum(x["energy_evaluations"] for x in resb), iterations_per_s=its / dtb,
                terminations=term, speedup_vs_single=(its / dtb) / converge["iterations_per_s_single"],
                workload="64-point scan x 2 seeds, medium size, lockstep rounds")
            poolc.close()

    mixed_batch = None
    if args.mixed_batch > 0:
        # uneven sizes: n_act uniform in {4,...,12} step 2 (all covalent Rumer structures), n_bf
        # uniform in [40, 200], n_aux = 3 n_bf; one pool per active space, all pools submitted
        # before any is waited for, so the buckets' launch sequences overlap on their own streams.
        # Problems are generated on host threads in chunks and loaded as they come, so host memory
        # holds one chunk of integrals at a time.
        rng = np.random.default_rng(4096)
        nb_total = args.mixed_batch
        acts = rng.choice([4, 6, 8, 10, 12], size=nb_total)
        bfs = rng.integers(40, 201, size=nb_total)
        buckets = {}
        for j, (na, nb) in enumerate(zip(acts, bfs)):
            buckets.setdefault(int(na), []).append((j, int(nb)))
        where = {}
        pools, orbs = {}, {}
        for na, members in buckets.items():
            pools[na] = vb.Pool(ctx, vb.Topology.build("COV", na, na, 1), capacity=len(members))
            orbs[na] = [None] * len(members)
            for slot, (j, nb) in enumerate(members):
                where[j] = (na, slot, nb)
        from concurrent.futures import ThreadPoolExecutor
        tg = time.perf_counter()
        resident_bytes = 0
        with ThreadPoolExecutor(max_workers=16) as ex:
            chunk = 64
            for c0 in range(0, nb_total, chunk):
                js = list(range(c0, min(nb_total, c0 + chunk)))
                made = list(ex.map(lambda j: vb.synthetic_problem(10000 + j, where[j][2], where[j][0], 3 * where[j][2]),
                                   js))
                for j, spm in zip(js, made):
                    na, slot, nb = where[j]
                    pools[na].load_synthetic(slot, spm)
                    orbs[na][slot] = spm.C
                    resident_bytes += spm.B.nbytes
                if (c0 // chunk) % 16 == 15:
                    print(f"batching workload: {c0 + len(js)} of {nb_total} problems generated and resident "
                          f"({resident_bytes / 2**30:.1f} GiB of RI factors, {time.perf_counter() - tg:.0f} s)",
                          file=sys.stderr, flush=True)
                del made
        gen_s = time.perf_counter() - tg

        tuned_buckets = {}
        if args.autotune:
            for na, pl in pools.items():  # on the bucket's own launch (the target is per launch)
                t = pl.autotune(list(range(len(orbs[na]))), orbs[na], repeats=1)
                tuned_buckets[str(na)] = dict(fwd_target=t["fwd_target"], bwd_target=t["bwd_target"])

        def mixed_step():
            tickets = {na: pools[na].submit(list(range(len(orbs[na]))), orbs[na]) for na in pools}
            return {na: pools[na].wait(t)[0] for na, t in tickets.items()}

        first = mixed_step()
        bad = [na for na, e in first.items() if not np.all(np.isfinite(e))]
        if bad:
            raise RuntimeError(f"batching workload: non-finite energies in buckets {bad}")
        per_bucket = {}
        msteps = args.mixed_steps_batching
        t0 = time.perf_counter()
        for _ in range(msteps):
            again = mixed_step()
        dtm = (time.perf_counter() - t0) / msteps
        same = max(float(np.max(np.abs(again[na] - first[na]))) for na in first)
        for na, pl in pools.items():
            per_bucket[str(na)] = dict(problems=len(orbs[na]), last_launch_ms=round(pl.phase_times()["total_ms"], 3),
                                       launch_shape=tuned_buckets.get(str(na)))
            pl.close()
        mixed_batch = dict(problems=nb_total, value=nb_total / dtm, unit="E+grad evals/s",
                           ms_per_step=dtm * 1e3, steps=msteps, buckets=per_bucket,
                           resident_ri_gib=resident_bytes / 2**30, setup_s=gen_s,
                           max_abs_dE_between_steps=same,
                           workload=f"{nb_total} problems per step: n_act uniform in {{4,6,8,10,12}} (COV singlet), "
                                    "n_bf uniform in [40,200], n_aux = 3 n_bf, one pool per n_act, host orbitals in "
                                    "and gradients out, FP64")

    value = K * args.steps / elapsed
    out = dict(metric=METRIC, value=value, unit="E+grad evals/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=elapsed * 1e3 / args.steps, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="medium: n_bf=120 n_aux=360 n_act=10 (5a/5b) 42 covalent Rumer structures, "
                                    f"252 determinants; {K} resident problems per batched step ({D} distinct "
                                    "seeded problems, each slot its own device copy), FP64 throughout, orbitals resident",
                           seed=cfg["seed"], n_bf=cfg["n_bf"], n_aux=cfg["n_aux"], n_act=cfg["n_act"],
                           n_structures=topo.n_structures, n_determinants=n_det, batch=K, path="batched-fp64",
                           launch_shape=(dict(fwd_target=tuned["fwd_target"], bwd_target=tuned["bwd_target"],
                                              tuned_ms=tuned["table"]) if tuned else "default")),
               latency_ms=(served["latency_ms"] if served else dict(p50=percentile(lat, 0.5),
                                                                    p99=percentile(lat, 0.99))),
               latency_source=("served leg: request completion minus Poisson arrival" if served
                               else "batched step"),
               batch_step_ms=dict(p50=percentile(lat, 0.5), p99=percentile(lat, 0.99)),
               plain_fp64=plain, speedup_vs_plain=value / plain["value"], served=served,
               served_mixed_sizes=served_mixed, mixed_precision=mixed, precision_form=precision,
               converge=converge, mixed_batch=mixed_batch,
               batch_consistency=batch_consistency, phase_ms={k: round(v, 5) for k, v in mean.items()},
               energy=energy, roofline=roofline, roofline_other_kernels=others)
    if not args.no_cpu:
        base, parity = cpu_baseline(cfg, sp, topo, energy, grad, args.cpu_seconds)
        out["cpu_baseline"] = base
        out["parity"] = parity
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
