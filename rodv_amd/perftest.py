"""GPU counterpart of roDV's PairHMM performance CLI (test/pairhmm_perftest.cpp:
271-488), same arguments and report:

    python -m rodv_amd.perftest [-v] [-i] [-h] <log_file> <max_idle_ratio_float> <max_idle_ratio_double>

For each region of an assemble-log dump (test/assemble_log_parser.cpp format)
it scores the region once with every pair as intra::computeLikelihoods and
once with schedule_pairhmm's routing (the two legs the reference times), both
on the GPU through the engine, checks the two within 1e-5 (the reference's
tolerance and skips: NaN/inf, |intra| < 1e-10, and with -i haplotypes or reads
containing N), and prints the times and the speedup.  It then scores all
regions as one batched plan (hipGraph replay, inputs resident) and prints its
GCUPS, which the reference's CLI has no counterpart for.  Exit status 0 when
every region passes, 1 otherwise or on a usage error.
"""
from __future__ import annotations

import getopt
import sys
import time

import numpy as np

from .batch import Batch, parse_assemble_log

TOLERANCE = 1e-5


def usage(prog):
    print(f"Usage: {prog} <log_file> <max_idle_ratio_float> <max_idle_ratio_double> [OPTIONS]")
    print("Arguments:")
    print("  log_file: Path to the log file containing regions")
    print("  max_idle_ratio_float: Maximum idle ratio for float precision (e.g., 0.5)")
    print("  max_idle_ratio_double: Maximum idle ratio for double precision (e.g., 0.7)")
    print("Options:")
    print("  -v, --verbose: Enable verbose output for schedule_pairhmm")
    print("  -i, --ignore-N-base: Skip comparison for reads containing N bases")
    print("  -h, --help: Show this help message")


def _has_n(seq: bytes) -> bool:
    return b"N" in seq or b"n" in seq


def compare(intra, sched, batch: Batch, ignore_n: bool, tol=TOLERANCE):
    """compareResults (pairhmm_perftest.cpp:190-245) on M x N matrices;
    returns (passed, message of the first mismatch)."""
    M, N = intra.shape
    for h in range(M):
        if ignore_n and _has_n(batch.hap(h)):
            continue
        for r in range(N):
            if ignore_n and _has_n(batch.read(r)):
                continue
            a, b = intra[h, r], sched[h, r]
            if not (np.isfinite(a) and np.isfinite(b)):
                continue
            if abs(a) < 1e-10:
                continue
            if abs(a - b) > tol:
                return False, (f"Error: Result mismatch for H{h}_R{r}\n  Intra result: {a}\n"
                               f"  Schedule result: {b}\n  Difference: {abs(a - b)}\n"
                               f"  Tolerance: {tol}")
    return True, ""


def main(argv=None) -> int:
    argv = sys.argv if argv is None else argv
    prog = argv[0]
    try:
        opts, args = getopt.gnu_getopt(argv[1:], "vih", ["verbose", "ignore-N-base", "help"])
    except getopt.GetoptError as e:
        print(f"{prog}: {e}", file=sys.stderr)
        usage(prog)
        return 1
    verbose = ignore_n = False
    for o, _ in opts:
        if o in ("-v", "--verbose"):
            verbose = True
        elif o in ("-i", "--ignore-N-base"):
            ignore_n = True
        elif o in ("-h", "--help"):
            usage(prog)
            return 0
    if len(args) < 3:
        print("Error: Missing required arguments", file=sys.stderr)
        usage(prog)
        return 1
    if len(args) > 3:
        print(f"Error: Unexpected argument: {args[3]}", file=sys.stderr)
        usage(prog)
        return 1
    log_file = args[0]
    try:
        idle_f, idle_d = float(args[1]), float(args[2])
    except ValueError as e:
        print(f"Error: Invalid number format: {e}", file=sys.stderr)
        usage(prog)
        return 1

    from .engine import Engine, default_config
    line = "=" * 42
    print(line)
    print("PairHMM Performance Test (MI355X engine)")
    print(line)
    print(f"Log file: {log_file}")
    print(f"Max idle ratio (float): {idle_f}")
    print(f"Max idle ratio (double): {idle_d}")
    print("Options:")
    print(f"  Verbose: {'Yes' if verbose else 'No'}")
    print(f"  Ignore N-base: {'Yes' if ignore_n else 'No'}")
    sched_cfg = default_config(semantics=1, max_idle_ratio_float=idle_f,
                               max_idle_ratio_double=idle_d, use_graph=0)
    print(f"Group widths of this host (float/double): {sched_cfg.simd_width_float}/"
          f"{sched_cfg.simd_width_double}")
    print(line)
    try:
        t0 = time.perf_counter()
        regions = parse_assemble_log(log_file, with_names=True)
        parse_s = time.perf_counter() - t0
    except (OSError, ValueError) as e:
        print(f"Error: Cannot parse log file: {e}", file=sys.stderr)
        return 1
    if not regions:
        print("Error: No regions found in log file", file=sys.stderr)
        return 1
    print(f"Parsed {len(regions)} region(s) in {parse_s:.6f} seconds")
    print(line)

    eng = Engine(0)
    intra_cfg = default_config(semantics=0, use_graph=0)
    # one untimed call so that device initialisation is not charged to region 1
    eng.forward(regions[0][1], intra_cfg)

    tot_intra = tot_sched = 0.0
    total_pairs = 0
    all_passed = True
    for k, (name, b) in enumerate(regions):
        M, N = len(b.hap_len), len(b.read_len)
        print(f"\nRegion {k + 1}/{len(regions)}: {name}")
        print(f"  Haplotypes: {M}")
        print(f"  Reads: {N}")
        total_pairs += M * N
        t0 = time.perf_counter()
        intra, _ = eng.forward(b, intra_cfg)
        t_intra = time.perf_counter() - t0
        tot_intra += t_intra
        print(f"  Intra time: {t_intra:.6f} seconds")
        t0 = time.perf_counter()
        sched, flags = eng.forward(b, sched_cfg)
        t_sched = time.perf_counter() - t0
        tot_sched += t_sched
        print(f"  Schedule time: {t_sched:.6f} seconds")
        if verbose:
            print(f"    inter pairs: {int(((flags & 8) != 0).sum())}, "
                  f"rescued in FP64: {int(((flags & 1) != 0).sum())}")
        ok, msg = compare(intra.reshape(M, N), sched.reshape(M, N), b, ignore_n)
        if ok:
            print(f"  Result check: PASSED (tolerance: {TOLERANCE})")
        else:
            print(msg, file=sys.stderr)
            print(f"  Result check: FAILED (tolerance: {TOLERANCE})")
            all_passed = False
            break
        if t_sched > 0:
            print(f"  Speedup: {t_intra / t_sched:.2f}x")

    # all regions as one plan: inputs resident in HBM, one hipGraph replay
    whole = Batch.concat([b for _, b in regions])
    plan = eng.plan(whole, default_config(semantics=1, max_idle_ratio_float=idle_f,
                                          max_idle_ratio_double=idle_d, use_graph=1))
    plan.execute()
    plan.read()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute()
    plan.read()
    batched_s = (time.perf_counter() - t0) / reps
    plan.close()
    cells = whole.cells()

    print(f"\n{line}")
    print("Summary")
    print(line)
    print(f"Total regions: {len(regions)}")
    print(f"Total pairs (M*N): {total_pairs}")
    print(f"Total intra time: {tot_intra:.6f} seconds")
    print(f"Total schedule time: {tot_sched:.6f} seconds")
    if tot_sched > 0:
        print(f"Overall speedup: {tot_intra / tot_sched:.2f}x")
    print(f"Batched (all regions, one plan): {batched_s:.6f} seconds, "
          f"{cells / batched_s / 1e9:.2f} GCUPS")
    print(f"Result check: {'PASSED' if all_passed else 'FAILED'}")
    print(line)
    return 0 if all_passed else 1


if __name__ == "__main__":
    sys.exit(main())
