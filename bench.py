#!/usr/bin/env python3
"""VBSCF energy+gradient throughput on one MI355X.

Workload (medium config): synthetic RI problems n_bf=120, n_aux=360, n_act=10 VB orbitals with
10 electrons (5 alpha / 5 beta), 42 covalent Rumer singlet structures expanded to 252
determinants; orbitals resident in HBM. An evaluation is one full energy + analytic orbital
gradient (orbital normalisation, RI transform, determinant pairs, structure H/S, generalised
eigensolve, backward) through the engine's C ABI, FP64 throughout.

Two legs are timed:
  * plain: one problem per launch sequence (the north star's plain FP64 GPU path);
  * batched (value): a pool of --batch resident problems (seeds 0..K-1) evaluated by one launch
    sequence per step; one step = K evaluations.

Prints one JSON line: batched throughput (E+grad evaluations/s) beside the plain path's, p50/p99
latency of both, the roofline of the dominant kernel of the batched leg (its duration from HIP
events recorded on the engine stream during the timed region), parity against the FP64 CPU
oracle and the oracle's own throughput on the host cores (cpu_baseline).
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

MEDIUM = dict(seed=0, n_bf=120, n_act=10, n_el=10, n_aux=360, structures="COV")


def pair_flops(m_alpha: int, m_beta: int) -> float:
    """Algorithmic FP64 flops of one ordered determinant pair's forward matrix element:
    per spin block LU (2/3 m^3) + inverse (2 m^3) + same-spin two-electron terms (6 C(m,2)^2),
    plus the opposite-spin coupling (2 m_alpha^2 m_beta^2)."""
    def block(m):
        c2 = m * (m - 1) / 2
        return 2.0 / 3.0 * m ** 3 + 2.0 * m ** 3 + 6.0 * c2 * c2
    return block(m_alpha) + block(m_beta) + 2.0 * m_alpha ** 2 * m_beta ** 2


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
    base = dict(value=n / dt, unit="E+grad evals/s", cores=cores, kind="port",
                sample=f"{n} full E+grad evaluations of the same medium-config problem in {dt:.1f} s "
                       f"(oracle/vb_oracle.cpp, OpenMP {cores} threads)")
    return base, parity


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed batched steps (K evaluations each)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=128, help="resident problems evaluated per batched step")
    ap.add_argument("--plain-steps", type=int, default=300, help="timed single-problem evaluations")
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
              "pair_backward_ms", "pair_backward_kernel_ms", "gradient_tail_ms", "total_ms")

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
    for s in range(K):
        sps = sp if s == 0 else vb.synthetic_problem(cfg["seed"] + s, cfg["n_bf"], cfg["n_act"], cfg["n_aux"])
        pool.load_synthetic(s, sps)
        orbitals.append(sps.C)
    cb_dev = torch.from_numpy(np.concatenate([o.ravel() for o in orbitals])).to(dev)
    gb_dev = torch.zeros_like(cb_dev)
    slots = list(range(K))
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

    # roofline of the dominant single kernel of the batched leg (K problems per launch)
    n_det = topo.n_determinants
    nbp = cfg["n_bf"] * (cfg["n_bf"] + 1) // 2
    fwd_flops = K * n_det * n_det * pair_flops(topo.n_alpha, topo.n_beta)
    candidates = {
        "k_pairs<backward>": (mean["pair_backward_kernel_ms"], "mfma", 2.0 * fwd_flops, "TFLOP/s"),
        "k_pairs<forward>": (mean["pair_forward_ms"], "mfma", fwd_flops, "TFLOP/s"),
        "k_ri_one_index": (mean["ri_one_index_ms"], "hbm",
                           K * 8.0 * cfg["n_aux"] * (nbp + cfg["n_act"] * cfg["n_bf"]), "GB/s"),
        "k_gep_lowest": (mean["eigensolve_ms"], "mfma", K * 10.0 * topo.n_structures ** 3, "TFLOP/s"),
    }
    kname = max(candidates, key=lambda k: candidates[k][0])
    ms, bound, work, unit = candidates[kname]
    if unit == "GB/s":
        achieved, peak = work / (ms * 1e-3) / 1e9, HBM_PEAK_GBS
    else:
        achieved, peak = work / (ms * 1e-3) / 1e12, FP64_PEAK_TFLOPS
    roofline = dict(kernel=kname, bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak,
                    traffic=None, kernel_ms=ms, algorithmic_work=work, problems_per_launch=K)

    value = K * args.steps / elapsed
    out = dict(metric=METRIC, value=value, unit="E+grad evals/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=elapsed * 1e3 / args.steps, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="medium: n_bf=120 n_aux=360 n_act=10 (5a/5b) 42 covalent Rumer structures, "
                                    f"252 determinants; {K} resident problems per batched step, FP64 throughout, "
                                    "orbitals resident",
                           seed=cfg["seed"], n_bf=cfg["n_bf"], n_aux=cfg["n_aux"], n_act=cfg["n_act"],
                           n_structures=topo.n_structures, n_determinants=n_det, batch=K, path="batched-fp64"),
               latency_ms=dict(p50=percentile(lat, 0.5), p99=percentile(lat, 0.99)),
               plain_fp64=plain, speedup_vs_plain=value / plain["value"], batch_consistency=batch_consistency,
               phase_ms={k: round(v, 5) for k, v in mean.items()}, energy=energy, roofline=roofline)
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
