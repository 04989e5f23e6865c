#!/usr/bin/env python3
"""Benchmark of the MI355X PairHMM forward (GCUPS, batched vs plain).

One step = one batched forward (FP32 pass + FP64 rescue + log10) over the
batched uneven workload: 10,000 regions, geometric(mean 80) reads of length
U{50..250} per region, 2-64 haplotypes of length read + U{50..400}, 0.5%
unrelated reads (synthetic, seed 1002).  Inputs are resident in HBM before
the timed region; the launch is one hipGraph replay.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed) peak
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak (the rescue pass)
FLOP_PER_CELL = 12        # 8 mul + 4 add per forward cell (pairhmm_impl.cpp:266-289)
FP_INSTR_PER_CELL = 8     # non-packed FP32 instructions of the bit-exact FMA split
FP_INSTR_PER_CELL_SH = 7  # the same with the I and D recurrences sharing M x indel prior
# adjoint, algorithmic work per cell: 12 forward + 24 reverse + 10 parameter
# accumulation (the kernel's checkpoint recomputation is not counted)
ADJ_FLOP_PER_CELL = 46

# HBM bytes per FP32 forward pass (all length buckets), from two rocprofv3
# --pmc runs of this script on the same workload: FETCH_SIZE x 2.000 +
# WRITE_SIZE x 1.000, the factors a 1 GiB device copy in the same run gives
# (--copy-calibration-mib 1024).  Keyed by (config, scale, semantics,
# autotuned, haplotype chains).
MEASURED_TRAFFIC = {(2, 1.0, "schedule", True, 1): 4.230e9, (2, 1.0, "schedule", True, 0): 2.239e10}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(batch, seconds=12.0, seed=7, all_core_seconds=6.0):
    """Time the reference's own intra kernel (compiled from its sources into
    oracle/_ref) on host core 0, one thread, over seeded random regions of
    the same workload until about `seconds` of CPU work, then on every core
    for about `all_core_seconds`; falls back to the oracle restatement (one
    core) when the reference build is unavailable here."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    order = np.random.default_rng(seed).permutation(len(batch.regions))

    def sample_regions():
        for k in order:
            h0, m, r0, n = (int(x) for x in batch.regions[k])
            c = int(batch.hap_len[h0:h0 + m].sum()) * int(batch.read_len[r0:r0 + n].sum())
            yield h0, m, r0, n, c

    ref = None
    for isa in ("avx512", "avx2"):
        try:
            ref = O.RefLib(isa)
            break
        except (FileNotFoundError, RuntimeError, OSError):
            continue
    cells = pairs = regions = 0
    t0 = time.perf_counter()
    for h0, m, r0, n, c in sample_regions():
        if ref is not None:
            for hh in range(h0, h0 + m):
                for rr in range(r0, r0 + n):
                    ref.log10(batch, hh, rr)
        else:
            hs = np.repeat(np.arange(h0, h0 + m, dtype=np.uint32), n)
            rs = np.tile(np.arange(r0, r0 + n, dtype=np.uint32), m)
            O.run_batch(batch, pairs=(hs, rs), nthreads=1)
        cells += c
        pairs += m * n
        regions += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    what = (f"intra::computeLikelihoods{ref.isa.upper()} per pair (reference sources, "
            "g++ -O3)" if ref is not None else "oracle restatement, scalar")
    res = {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": 1,
           "kind": "reference" if ref is not None else "port",
           "sample": f"{what}; {regions} random regions of the workload, {pairs} pairs, "
                     f"{cells:.3e} cells in {dt:.1f} s; host {_cpu_model()}"}
    if ref is not None and all_core_seconds > 0:
        # the same per-pair kernel on every host core of this process (at most
        # 16, a one-GPU host's share), threads over pairs
        nt = max(1, min(16, len(os.sched_getaffinity(0))))
        target = res["value"] * 1e9 * nt * all_core_seconds
        hs, rs = [], []
        c2 = 0
        for h0, m, r0, n, c in sample_regions():
            hs.append(np.repeat(np.arange(h0, h0 + m, dtype=np.uint32), n))
            rs.append(np.tile(np.arange(r0, r0 + n, dtype=np.uint32), m))
            c2 += c
            if c2 >= target:
                break
        t1 = time.perf_counter()
        ref.log10_batch(batch, (np.concatenate(hs), np.concatenate(rs)), nt)
        d2 = time.perf_counter() - t1
        res["all_cores"] = {"value": round(c2 / d2 / 1e9, 4), "cores": nt,
                            "sample": f"{len(hs)} random regions, {c2:.3e} cells in {d2:.1f} s"}
    return res


def oracle_check(batch, out, fl, r32, cfg, semantics, n_regions=100, max_cells=2.5e10, seed=11):
    """Outside the timed region: up to `n_regions` seeded random regions of
    the timed run (at most `max_cells` cells) against the oracle's
    restatement of the same semantics -- schedule_pairhmm routing with the
    run's group widths and idle bounds, or intra::computeLikelihoods per
    pair.  Reports the largest |d log10 L| over finite results and whether
    the deciding FP32 sums, the rescue set and the -inf set are identical."""
    import dataclasses
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from rodv_amd.engine import FLAG_INTER, FLAG_RESCUED, FLAG_UNDERFLOW
    regs = batch.regions
    npairs = regs[:, 1].astype(np.int64) * regs[:, 3].astype(np.int64)
    start = np.concatenate(([0], np.cumsum(npairs)[:-1]))
    sel, cells = [], 0
    for k in np.random.default_rng(seed).permutation(len(regs)):
        h0, m, r0, n = (int(x) for x in regs[k])
        c = int(batch.hap_len[h0:h0 + m].sum()) * int(batch.read_len[r0:r0 + n].sum())
        if sel and cells + c > max_cells:
            continue
        sel.append(int(k))
        cells += c
        if len(sel) >= n_regions:
            break
    sel = np.sort(np.array(sel))
    sub = dataclasses.replace(batch, regions=regs[sel])
    idx = np.concatenate([np.arange(start[k], start[k] + npairs[k]) for k in sel])
    nt = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    route = None
    if semantics == "schedule":
        o_out, o_fl, o_r32, route = O.schedule_batch(
            sub, use_double=bool(cfg.use_double), idle_f=cfg.max_idle_ratio_float,
            idle_d=cfg.max_idle_ratio_double, wf=cfg.simd_width_float, wd=cfg.simd_width_double,
            nthreads=nt)
        mask = FLAG_RESCUED | FLAG_UNDERFLOW | FLAG_INTER
    else:
        o_out, o_r32, _, o_fl = O.run_batch(sub, nthreads=nt)
        mask = FLAG_RESCUED | FLAG_UNDERFLOW
    g_out, g_fl, g_r32 = out[idx], fl[idx], r32[idx]
    n_ext = 0
    if cfg.extended_range:
        # the pairs the reference scores -inf: against the oracle's
        # extended-range forward of their route's arithmetic
        ext = np.isneginf(o_out)
        n_ext = int(ext.sum())
        if n_ext:
            e_in, _, _, _ = O.run_batch(sub, mode=O.MODE_EXTENDED, nthreads=nt)
            if route is not None:
                e_it, _, _, _ = O.run_batch(sub, mode=O.MODE_EXTENDED, semantics=O.INTER, nthreads=nt)
                e_in = np.where(route == O.ROUTE_INTER64, e_it, e_in)
            o_out = np.where(ext, e_in, o_out)
            mask &= ~FLAG_UNDERFLOW  # finite now
    dt = time.perf_counter() - t0
    fin = np.isfinite(o_out)
    err = float(np.max(np.abs(g_out[fin] - o_out[fin]))) if fin.any() else 0.0
    return {"regions": int(len(sel)), "pairs": int(len(idx)), "cells": int(cells),
            "max_abs_dlog10": err,
            "raw32_bit_exact": bool(np.array_equal(g_r32.view(np.uint32), o_r32.view(np.uint32))),
            "rescue_set_equal": bool(np.array_equal(g_fl & FLAG_RESCUED, o_fl & FLAG_RESCUED)),
            "flags_equal": bool(np.array_equal(g_fl & mask, o_fl & mask)),
            "neg_inf_equal": bool(np.array_equal(np.isneginf(g_out), np.isneginf(o_out))),
            "extended_pairs_checked": n_ext,
            "rescued": int(((o_fl & FLAG_RESCUED) != 0).sum()),
            "semantics": semantics, "oracle_threads": nt, "oracle_seconds": round(dt, 1)}


def dropin_latency(reps=20):
    """Per-call wall time of the drop-in as a reference caller sees it: the
    C++ caller tests/dropin/schedule_driver (pairhmm::schedule::schedule_pairhmm
    declared as the reference's header does, linked against the engine) on
    the single-region configurations, `reps` calls in a row in one child
    process (the first creates the engine's context; later calls reuse its
    device buffers); beside it the reference's own intra kernel
    (oracle/_ref, one host thread) on the same pairs."""
    from rodv_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "tests", "dropin"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import driver as D
    import oracle as O
    ref = None
    for isa in ("avx512", "avx2"):
        try:
            ref = O.RefLib(isa)
            break
        except (FileNotFoundError, RuntimeError, OSError):
            continue
    res = {}
    for cfg in (0, 1):
        b = synth.config(cfg)
        ok, _, us = D.run(*D.batch_lists(b), reps=reps)
        med = float(np.median(us[1:])) if len(us) > 1 else float(us[0])
        e = {"ok": ok, "pairs": b.num_pairs(), "cells": b.cells(), "calls": len(us),
             "call_us_median": round(med, 1), "call_us_first": round(float(us[0]), 1),
             "gcups_per_call": round(b.cells() / (med / 1e6) / 1e9, 2)}
        if ref is not None:
            t0 = time.perf_counter()
            ref.log10_batch(b, b.cross_pairs(), 1)
            dt = time.perf_counter() - t0
            e["reference_cpu_us"] = round(dt * 1e6, 1)
            e["reference_cpu"] = (f"intra::computeLikelihoods{ref.isa.upper()} per pair, 1 thread, "
                                  f"host {_cpu_model()}")
        res[f"config{cfg}"] = e
    return res


def valu_ceiling():
    """FP32 issue ceiling of this device (rodv_amd/libvalu_ceiling.so): lane
    operations per second of independent v_fma_f32, v_mul_f32 and of the
    forward cell's own 8-instruction mix and its 7-instruction shared-product
    mix, at 8 waves per SIMD; those divided by 8 and 7 are the cells/s a
    bit-exact non-packed forward can reach."""
    import ctypes
    path = os.path.join(ROOT, "rodv_amd", "libvalu_ceiling.so")
    if not os.path.exists(path):
        log("libvalu_ceiling.so not built; no issue ceiling")
        return None
    lib = ctypes.CDLL(path)
    lib.valu_ceiling.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_double)]
    out = {}
    for kind, name in ((0, "fma_lane_ops_per_s"), (1, "mul_lane_ops_per_s"),
                       (2, "cell_mix_lane_ops_per_s"), (3, "cell_mix_shared_lane_ops_per_s")):
        v = ctypes.c_double(0.0)
        rc = lib.valu_ceiling(0, kind, 8, 3, ctypes.byref(v))
        if rc:
            log(f"valu_ceiling kind {kind} failed ({rc})")
            return None
        out[name] = round(v.value, 1)
    out["cells_per_s"] = round(out["cell_mix_lane_ops_per_s"] / FP_INSTR_PER_CELL, 1)
    out["cells_per_s_shared"] = round(out["cell_mix_shared_lane_ops_per_s"] / FP_INSTR_PER_CELL_SH, 1)
    out["tflops_at_12_flop_per_cell"] = round(out["cells_per_s"] * FLOP_PER_CELL / 1e12, 2)
    out["tflops_at_12_flop_per_cell_shared"] = round(out["cells_per_s_shared"] * FLOP_PER_CELL / 1e12, 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint", type=int, default=1)
    ap.add_argument("--prior-table", type=int, default=1)
    ap.add_argument("--haplotype-chains", type=int, choices=(0, 1, 2, 3, 4), default=1,
                    help="chained items: 1 where a length bucket is large enough, short reads "
                         "two per lane group where that packs better (default), 2 always, "
                         "3 and 4 as 1 and 2 with one read per lane group, 0 one pair per slot")
    ap.add_argument("--shared-products", type=int, choices=(0, 1), default=1,
                    help="reads whose indel priors allow it take the shared-product kernels "
                         "(7 FP32 instructions per cell, the same bits)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="tune pairs per wavefront / strip width / waves per CU on a 4%% sample "
                         "of the workload's distribution before timing")
    ap.add_argument("--semantics", choices=("schedule", "intra"), default="schedule")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="persistent grid size per CU (0 = engine default)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this size (reference for PMC byte counters)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--extended-range", type=int, choices=(0, 1), default=0,
                    help="score pairs whose FP64 sum underflows again with renormalised strips")
    ap.add_argument("--dropin", type=int, default=1,
                    help="time schedule_pairhmm per call through the C++ reference-style caller "
                         "on the single-region configurations")
    ap.add_argument("--oracle-regions", type=int, default=100,
                    help="random regions of the timed run checked against the oracle "
                         "afterwards (0: no check)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (the forward shards by region with no exchange).", file=sys.stderr)
        sys.exit(2)

    import torch
    from rodv_amd import synth
    from rodv_amd.engine import Engine, default_config

    t0 = time.time()
    batch = synth.config(args.config, scale=args.scale)
    cells = batch.cells()
    log(f"workload config {args.config}: {batch.num_pairs()} pairs, {cells:.4e} cells "
        f"({time.time() - t0:.1f}s to generate)")
    eng = Engine(0)
    sem = 1 if args.semantics == "schedule" else 0
    tuning = None
    if args.autotune:
        t0 = time.time()
        sample = synth.config(args.config, scale=args.scale * 0.04)
        tn = eng.autotune(sample, default_config(prior_table=args.prior_table,
                                                 shared_products=args.shared_products))
        tuning = {"sample_pairs": sample.num_pairs(), "seconds": round(time.time() - t0, 1),
                  "classes": [{"max_read_len": e["max_read_len"],
                               "pairs_per_wave": e["pairs_per_wave"],
                               "max_rows_per_lane": e["max_rows_per_lane"],
                               "waves_per_cu": e["waves_per_cu"],
                               "gcups": round(e["gcups"], 1),
                               "default_gcups": round(e["default_gcups"], 1)}
                              for e in tn.entries() if e["pairs_per_wave"]]}
        log(f"autotune on {sample.num_pairs()} sample pairs: {tuning['seconds']} s")
        for c in tuning["classes"]:
            log(f"  R <= {c['max_read_len']}: {c['pairs_per_wave']} pairs/wave, "
                f"rows/lane cap {c['max_rows_per_lane']}, waves/CU {c['waves_per_cu']}: "
                f"{c['gcups']} GCUPS (built-in {c['default_gcups']})")
    t0 = time.time()
    run_cfg = default_config(use_graph=1, prior_table=args.prior_table,
                             waves_per_cu=args.waves_per_cu, semantics=sem,
                             haplotype_chains=args.haplotype_chains,
                             extended_range=args.extended_range,
                             shared_products=args.shared_products)
    plan = eng.plan(batch, run_cfg)
    log(f"plan + upload {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    if args.copy_calibration_mib:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
        # calibrate the counters of the PairHMM kernels in the same run
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {n * 4} bytes read, {n * 4} bytes written")
        del src, dst

    for _ in range(args.warmup):
        plan.execute(sh)
    torch.cuda.synchronize()
    log("warmup done")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    ev0.record(stream)
    evs[0].record(stream)
    for k in range(args.steps):
        plan.execute(sh)
        evs[k + 1].record(stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    median_ms = float(np.median([evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]))
    gcups = cells / (ms / 1e3) / 1e9
    log(f"batched: {ms:.2f} ms/step (median replay {median_ms:.2f} ms), {gcups:.1f} GCUPS")

    # per-kernel times with hipEvents on the launch stream
    profs = [plan.profile(sh) for _ in range(2)]
    fp32_ms = min(p["fp32_kernel_ms"] for p in profs)
    fp64_ms = min(p["fp64_kernel_ms"] for p in profs)
    st = profs[-1]
    out, fl, r32, _ = plan.read(raw=True)
    rescued = int(st["rescued_pairs"])
    rescued_cells = int(st["rescued_cells"])
    log(f"fp32 kernels {fp32_ms:.2f} ms, fp64 rescue {fp64_ms:.2f} ms, rescued {rescued} "
        f"({rescued_cells:.3e} cells), {int(st['stacked_pairs'])} pairs in two-read chained items, "
        f"{int(st['shared_pairs'])} in shared-product kernels, {int(st['num_buckets'])} FP32 buckets")
    parity = None
    if args.oracle_regions > 0:
        t0 = time.time()
        parity = oracle_check(batch, out, fl, r32, run_cfg, args.semantics, args.oracle_regions)
        log(f"oracle check ({time.time() - t0:.1f}s): {parity}")

    # plain: one launch per region (and length bucket)
    plain_ms = None
    if args.plain_steps > 0:
        plan.execute_plain(sh)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(args.plain_steps):
            plan.execute_plain(sh)
        ev1.record(stream)
        torch.cuda.synchronize()
        plain_ms = ev0.elapsed_time(ev1) / args.plain_steps
        out_p, fl_p = plan.read()
        same = bool(np.array_equal(out_p, out, equal_nan=True))
        log(f"plain: {plain_ms:.2f} ms/step ({plain_ms / ms:.2f}x batched), identical={same}")

    # adjoint: forward + checkpointed reverse sweep on the forward+adjoint
    # configuration (4,096 pairs, R = 150, H = 400, checkpoint every 32nd step)
    # (config 4 runs the adjoint of its own long reads: FP64, strips of 252
    # rows in sequence, renormalised between strips)
    adj = None
    if args.adjoint:
        acfg = 4 if args.config == 4 else 3
        b3 = synth.config(acfg)
        theta = [45.0, 10.0] + [1.0] * 8
        eng.adjoint(b3, theta, timing=True)
        times = [eng.adjoint(b3, theta, timing=True)[2] for _ in range(3 if acfg == 3 else 1)]
        adj_ms = min(times)
        adj_tflops = ADJ_FLOP_PER_CELL * b3.cells() / (adj_ms / 1e3) / 1e12
        peak = FP32_PEAK_TFLOPS if acfg == 3 else FP64_PEAK_TFLOPS
        adj = {"gcups": round(b3.cells() / (adj_ms / 1e3) / 1e9, 2), "kernel_ms": round(adj_ms, 3),
               "workload": ("config3: 4096 pairs, R=150, H=400, ckpt every 32 steps" if acfg == 3 else
                            "config4: 256 pairs, R 1,000-10,000, H <= 12,000, FP64, ckpt every 32 steps"),
               "cells": b3.cells(), "flop_per_cell": ADJ_FLOP_PER_CELL,
               "tflops": round(adj_tflops, 2),
               ("frac_of_fp32_peak" if acfg == 3 else "frac_of_fp64_peak"): round(adj_tflops / peak, 4)}
        log(f"adjoint (config {acfg}): {adj_ms:.2f} ms, {adj['gcups']} GCUPS")

    dropin = None
    if args.dropin:
        t0 = time.time()
        try:
            dropin = dropin_latency()
        except (OSError, RuntimeError, subprocess.SubprocessError) as e:
            dropin = {"error": str(e)[-500:]}
        log(f"drop-in schedule_pairhmm per call ({time.time() - t0:.1f}s): {dropin}")

    achieved = FLOP_PER_CELL * cells / (fp32_ms / 1e3) / 1e12
    achieved_step = FLOP_PER_CELL * cells / (ms / 1e3) / 1e12
    fp64_achieved = FLOP_PER_CELL * rescued_cells / (fp64_ms / 1e3) / 1e12 if fp64_ms > 0 else 0.0
    ceiling = valu_ceiling()
    if ceiling:
        cps = cells / (fp32_ms / 1e3)
        ceiling["forward_cells_per_s"] = round(cps, 1)
        # the workload's mix: shared-product pairs at 7 instructions per cell
        fs = float(st["shared_pairs"]) / max(1, batch.num_pairs())
        mix = 1.0 / (fs / ceiling["cells_per_s_shared"] + (1.0 - fs) / ceiling["cells_per_s"])
        ceiling["workload_cells_per_s"] = round(mix, 1)
        ceiling["frac_of_cell_mix_ceiling"] = round(cps / mix, 4)
        log(f"VALU ceiling: {ceiling}")
    traffic = MEASURED_TRAFFIC.get((args.config, args.scale, args.semantics, bool(args.autotune),
                                    args.haplotype_chains)) \
        if (args.waves_per_cu == 0 and args.prior_table) else None
    res = {
        "metric": "GCUPS fwd/adjoint, batched vs plain, 1 MI355X, |dlog10L|<=1e-4",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 rescue)",
        "data": "synthetic",
        "config": {"workload": f"config{args.config}: batched uneven regions"
                   if args.config == 2 else f"config{args.config}",
                   "pairs": batch.num_pairs(), "cells": cells, "scale": args.scale,
                   "parallelism": "single GPU, persistent length-bucketed hipGraph launch"},
        "plain_gcups": None if plain_ms is None else round(cells / (plain_ms / 1e3) / 1e9, 2),
        "batched_over_plain": None if plain_ms is None else round(plain_ms / ms, 3),
        "adjoint": adj,
        "dropin_latency": dropin,
        "semantics": args.semantics,
        "autotune": tuning,
        "parity": parity,
        "stacked_pairs": int(st["stacked_pairs"]),
        "shared_pairs": int(st["shared_pairs"]),
        "rescued_pairs": rescued,
        "rescued_cells": rescued_cells,
        "rescued_fraction": round(rescued / max(1, batch.num_pairs()), 6),
        "fp64_rescue_ms": round(fp64_ms, 3),
        "median_replay_ms": round(median_ms, 3),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                     "traffic_source": None if traffic is None else
                     "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this configuration, scaled by a "
                     "1 GiB device copy (a committed measurement, not this run's)",
                     # the whole step (FP32 pass, FP64 regroup and rescue, log10)
                     # over the driver-timed ms_per_step
                     "achieved_step": round(achieved_step, 3),
                     "frac_step": round(achieved_step / FP32_PEAK_TFLOPS, 4),
                     "fp64": {"bound": "valu", "achieved": round(fp64_achieved, 3),
                              "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(fp64_achieved / FP64_PEAK_TFLOPS, 4),
                              "cells": rescued_cells, "pairs": rescued,
                              "kernel_ms": round(fp64_ms, 3),
                              "kernel": "regroup_kernel + fwd_rescue(_route)_kernel + "
                                        "fwd_rescue_long_kernel <double>"},
                     "kernel": (("fwd_chain_kernel + fwd_tab_kernel <float> (all length buckets, "
                                 "one FP32 pass)" if args.haplotype_chains else
                                 "fwd_tab_kernel<float> (all length buckets, one FP32 pass)")
                                if args.prior_table else "fwd_kernel<float> (all length buckets)")
                     # reads beyond one 64-lane table strip (1,023 rows) take
                     # the strip-pipelined long-read kernel
                     + (" + fwd_long_kernel<float> (reads > 1,023 rows)"
                        if int(batch.read_len.max()) > 1023 else ""),
                     "kernel_ms": round(fp32_ms, 3),
                     "valu_issue_ceiling": ceiling},
    }
    if not args.no_cpu_baseline:
        t0 = time.time()
        res["cpu_baseline"] = cpu_baseline(batch, args.cpu_seconds)
        log(f"cpu baseline {res['cpu_baseline']} ({time.time() - t0:.1f}s)")
    plan.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
