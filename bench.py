#!/usr/bin/env python3
"""All-pairs Needleman-Wunsch throughput on one MI355X.

Workload: the log-normal protein set (4,000 sequences, lengths log-normal with
median 350 and sigma 0.8 clipped to [50, 5000], each from its own random parent
with substitution 0.3), all 7,998,000 pairs i<j plus the 4,000 self pairs the
distance matrix keeps on its diagonal, int32 scores, BLOSUM62, gap open 10,
gap extension 2. A step is one pass of the score kernels over the whole set,
with the sequences and the bucketed pair list already resident in HBM.

value      GCUPS of the batched form = sum_{i<j} L_i*L_j / wall seconds per step
plain      GCUPS of the plain drop-in form (one synchronous launch per bucket)
roofline   the dominant launch's integer VALU rate: 7 algorithmic int ops per
           cell (add S; sub and max for E; sub and max for F; max3 for H;
           H - go) x its cells / its HIP-event time, against 256 CU x 64 lanes
           x 2.4 GHz = 39.3 T int32 ops/s, or twice that for the packed 16-bit
           kernels (two cells per lane and instruction)
cpu_baseline  the oracle recurrence (oracle/liboracle.so, OpenMP over pairs)
           on a bounded random sample of the same pairs

config1    configuration 1 (1,000 proteins of ~200-600): plain and batched
           all-pairs GCUPS, NJ and progressive alignment times
long_pairs configuration 3 (256 sequences of 8,000-20,000): all-pairs GCUPS,
           NJ and progressive alignment times
traceback  the alignment form (flag kernels + traceback) on 20,000 random pairs
           of the set: flag bytes written (0.5 B per cell) per second of the
           flag kernels against HBM's 8 TB/s, and the traceback's time
served     the served form (continuous batching) on the job workload, a job
           being the whole run (scores, NJ tree, alignment): offline and loaded
           throughput, p50/p99 latency loaded and unloaded

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-serve]
"""
import argparse
import json
import os
import sys
import time

# One hardware queue per stream: the served form runs 17 streams (score
# rounds, NJ and merge batches); with HIP's default of 4 queues, streams
# sharing a queue serialise and a merge waits behind unrelated score rounds.
# Read when the runtime starts, so set before the first HIP call.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "24")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 39.3 T int32 VALU ops/s
PEAK_INT16_TOPS = 2 * PEAK_INT32_TOPS       # 78.6 T int16 ops/s (packed, 2 per lane)
OPS_PER_CELL = 7  # algorithmic int ops per DP cell, whatever the kernel issues


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter profile
    (profiles/counters_lognormal4000.md: FETCH_SIZE and WRITE_SIZE scaled by a
    plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "counters_lognormal4000.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic ") and " fetch_KiB=" in line and " write_KiB=" in line:
            name, rest = line[len("traffic "):].split(" fetch_KiB=")
            if name == kernel:
                fetch, write = rest.split(" write_KiB=")
                return int((float(fetch) + float(write)) * 1024)
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seqs, target_seconds=4.0, runs=3):
    """Oracle recurrence on the host cores over random samples of pairs: the
    median of `runs` samples of about `target_seconds` each. Threads: the CPU
    share the environment grants this process (OMP_NUM_THREADS; a GPU box
    gives each GPU's job its own share of the host), else every core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O

    S = O.blosum62()
    codes, offs = O.pack_seqs(seqs)
    n = len(seqs)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    rng = np.random.default_rng(123)
    # calibrate on a small sample, then size the measured samples
    pi = rng.integers(0, n, 256)
    pj = rng.integers(0, n, 256)
    keep = pi != pj
    pi, pj = pi[keep], pj[keep]
    _, secs = O.pair_list(codes, offs, pi, pj, S, threads=threads)
    rate = (L[pi] * L[pj]).sum() / max(secs, 1e-6)
    mean_cells = float((L.mean()) ** 2)
    npairs = int(max(512, min(2_000_000, rate * target_seconds / mean_cells)))
    vals, tot_cells, tot_secs, tot_pairs = [], 0, 0.0, 0
    for _ in range(runs):
        pi = rng.integers(0, n, npairs)
        pj = rng.integers(0, n, npairs)
        keep = pi != pj
        pi, pj = pi[keep], pj[keep]
        _, secs = O.pair_list(codes, offs, pi, pj, S, threads=threads)
        cells = int((L[pi] * L[pj]).sum())
        vals.append(cells / secs / 1e9)
        tot_cells += cells
        tot_secs += secs
        tot_pairs += len(pi)
    return {
        "value": round(float(np.median(vals)), 3),
        "unit": "GCUPS",
        "cores": threads,
        "nproc": nproc,
        "cpu_model": cpu_model(),
        "runs": [round(v, 3) for v in vals],
        "kind": "port",
        "sample": "median of %d samples of %d random pairs (i != j) of the same 4,000-sequence set, "
                  "%.3g cells and %.1f s wall in all" % (runs, len(pi), tot_cells, tot_secs),
    }


def traceback_bench(M, ctx, st, L, npairs=20000, seed=7):
    """The alignment form (flag kernels + traceback, mst_align_pairs) on random
    pairs of the same set: device time of the flag kernels and of the
    traceback from HIP events, and the flag bytes they write: 0.5 B per cell
    (a nibble: direction and two continuation bits), against HBM's 8 TB/s."""
    rng = np.random.default_rng(seed)
    n = len(L)
    pi = rng.integers(0, n, npairs).astype(np.int32)
    pj = rng.integers(0, n, npairs).astype(np.int32)
    keep = pi != pj
    pi, pj = pi[keep], pj[keep]
    st.align_pairs(pi[:64], pj[:64], L)  # warm
    t0 = time.perf_counter()
    st.align_pairs(pi, pj, L)
    wall = time.perf_counter() - t0
    stats = ctx.last_launch_stats()
    # the forward pass as a whole (-3000: the packed kernel's score tables and
    # every flag kernel, the packed and int32 ones overlapped on two streams)
    flag = [r for r in stats if r[0] == -3000]
    trace = [r for r in stats if r[0] == -3100]
    cells = int((L[pi].astype(np.int64) * L[pj]).sum())
    ms_flag = float(sum(r[3] for r in flag))
    ms_trace = float(sum(r[3] for r in trace))
    m_rows, n_cols = L[pj].astype(np.int64), L[pi].astype(np.int64)
    written = int((((m_rows + 511) // 512) * (n_cols + 63) * 64 * 4).sum())  # padded flag blocks
    alg_bytes = cells * 0.5
    gbs = alg_bytes / (ms_flag * 1e-3) / 1e9
    return {
        "workload": "%d random pairs (i != j) of the same set, %.3g cells" % (len(pi), cells),
        "flag_kernel_ms": round(ms_flag, 3),
        "flag_gcups": round(cells / (ms_flag * 1e-3) / 1e9, 1),
        "traceback_ms": round(ms_trace, 3),
        "flag_bytes_algorithmic": int(alg_bytes),
        "flag_bytes_written": written,
        "hbm": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(gbs / 8000.0, 4),
                "written_GBps": round(written / (ms_flag * 1e-3) / 1e9, 1)},
        "launches": [{"kernel": {-3000: "forward pass (all flag kernels)",
                                 -3001: "align_kernel<SEQ>", -3004: "align_pipe_kernel<SEQ, 4>",
                                 -3008: "align_tile_kernel<SEQ>", -3016: "flag16_kernel",
                                 -3017: "seq_table_kernel", -3100: "traceback_kernel"}[k],
                      "jobs": p, "cells": c,
                      "ms": round(m, 3)} for k, p, c, m in stats],
        "wall_s_with_host_strings": round(wall, 3),
    }


def config1_bench(M, ctx):
    """Configuration 1 (1,000 proteins of ~200-600 residues, the plain GPU
    drop-in's workload): all pairs in the plain form (one int32 launch per
    bucket) and in the batched form (device time of a resident pass), the NJ
    tree and the progressive alignment along it (wall time)."""
    seqs = M.synth(1, 0)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    cells = int((L.sum() ** 2 - (L * L).sum()) // 2)
    st = ctx.upload(seqs)
    st.all_pairs_resident(M.FORM_PLAIN)  # plan
    ms_p = min(st.all_pairs_resident(M.FORM_PLAIN) for _ in range(3))
    st.all_pairs_resident(M.FORM_BATCHED)  # plan and graph
    ms_b = min(st.all_pairs_resident(M.FORM_BATCHED) for _ in range(3))
    t0 = time.perf_counter()
    joins = st.nj()
    t1 = time.perf_counter()
    rows = st.msa(joins)
    t2 = time.perf_counter()
    st.free()
    return {
        "workload": "configuration 1: %d sequences, lengths %d..%d, %.3g cells (i<j)"
                    % (len(seqs), L.min(), L.max(), cells),
        "plain_ms": round(ms_p, 2),
        "plain_gcups": round(cells / (ms_p * 1e-3) / 1e9, 1),
        "batched_ms": round(ms_b, 2),
        "batched_gcups": round(cells / (ms_b * 1e-3) / 1e9, 1),
        "nj_ms": round((t1 - t0) * 1e3, 1),
        "msa_s": round(t2 - t1, 3),
        "msa_columns": len(rows[0]) if rows else 0,
    }


def long_pairs_bench(M, ctx):
    """Configuration 3 (256 sequences of 8,000-20,000 residues): all pairs in
    the batched form (device time of the resident pass), then NJ and the
    progressive alignment along its tree (wall time; 255 profile merges,
    the longest through the tiled flag kernel)."""
    seqs = M.synth(3, 0)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    cells = int((L.sum() ** 2 - (L * L).sum()) // 2)
    st = ctx.upload(seqs)
    ms = st.all_pairs_resident(M.FORM_BATCHED)  # first call: plan and one pass
    ms = st.all_pairs_resident(M.FORM_BATCHED)
    # packed launches (class < 0) count against the int16x2 peak, int32 ones
    # against the int32 peak
    stats = st.launch_stats(M.FORM_BATCHED)
    packed_cells = sum(r[2] for r in stats if r[0] < 0)
    int32_cells = sum(r[2] for r in stats if r[0] > 0)
    peak = PEAK_INT16_TOPS if packed_cells >= int32_cells else PEAK_INT32_TOPS
    t0 = time.perf_counter()
    joins = st.nj()
    t1 = time.perf_counter()
    rows = st.msa(joins)
    t2 = time.perf_counter()
    st.free()
    gcups = cells / (ms * 1e-3) / 1e9
    return {
        "workload": "configuration 3: %d sequences, lengths %d..%d, %.3g cells (i<j)"
                    % (len(seqs), L.min(), L.max(), cells),
        "all_pairs_ms": round(ms, 1),
        "all_pairs_gcups": round(gcups, 1),
        "all_pairs_roofline_frac": round(gcups * 1e9 * OPS_PER_CELL / (peak * 1e12), 4),
        "all_pairs_peak_tops": round(peak, 2),
        "all_pairs_launches": [{"class": int(r[0]), "pairs": int(r[1]), "cells": int(r[2]),
                                "ms": round(float(r[3]), 3)} for r in stats],
        "nj_ms": round((t1 - t0) * 1e3, 1),
        "msa_s": round(t2 - t1, 3),
        "msa_columns": len(rows[0]) if rows else 0,
    }


def served_bench(M, ctx, njobs=2000, load=0.75, seed=0, sets=(), what=None, unloaded_every=4):
    """Served form on the job workload (mestrado_amd.served.measure): njobs
    jobs, sizes power-law (exponent 1.5) on 8..512 sequences, each a
    tree-evolved family of lengths 100..2000; a job is the whole run
    (all-pairs scores, NJ tree, progressive alignment) unless `what` says
    otherwise; offline, unloaded (every `unloaded_every`-th job alone) and
    loaded (Poisson arrivals at `load` x the offline job rate)."""
    from mestrado_amd import served

    def note(msg):
        print("served: " + msg, file=sys.stderr, flush=True)

    jobs = served.job_workload(njobs, seed)
    note("%d jobs generated" % njobs)
    return served.measure(ctx, jobs, what=what, load=load, seed=seed, sets=sets,
                          unloaded_every=unloaded_every, note=note)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="time only the batched form (for a profile of its launches)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="a tunable of the batched form (mst_tune_set), repeatable")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-form measurement")
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--serve-set", action="append", default=[], metavar="KEY=VALUE",
                    help="a parameter of the served form (mst_server_set), repeatable")
    ap.add_argument("--serve-load", type=float, default=0.75,
                    help="arrival rate of the loaded run as a fraction of the offline job rate")
    ap.add_argument("--serve-what", choices=("msa", "tree", "scores"), default="msa",
                    help="what a served job computes (default: the whole run)")
    ap.add_argument("--serve-only", action="store_true", help="only the served-form measurement")
    ap.add_argument("--no-traceback", action="store_true", help="skip the alignment-form measurement")
    ap.add_argument("--no-long", action="store_true", help="skip the configuration 3 measurement")
    ap.add_argument("--long-only", action="store_true", help="only the configuration 3 measurement")
    ap.add_argument("--nj-only", action="store_true",
                    help="only the NJ guide tree of configurations 1 and 2 (after their batched all pairs)")
    ap.add_argument("--profile-set", action="store_true",
                    help="one pass of each profiled kernel (plain copy of 1 GiB, batched all pairs, "
                         "alignment form) for counter runs")
    ap.add_argument("--traceback-only", action="store_true",
                    help="only the alignment-form (flags + traceback) measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the all-pairs path runs on a single MI355X "
              "(sharding across GPUs is a design note, not built); use --gpus 1",
              file=sys.stderr)
        return 2

    import mestrado_amd as M

    seqs = M.synth(args.config, 0)
    L = np.array([len(s) for s in seqs], dtype=np.int64)
    cells = int((L.sum() ** 2 - (L * L).sum()) // 2)  # sum_{i<j} L_i L_j
    self_cells = int((L * L).sum())

    ctx = M.Context(device=0)
    if args.nj_only:
        out = {}
        for cfg in (1, 2):
            s = M.synth(cfg, 0)
            st = ctx.upload(s)
            st.all_pairs_resident(M.FORM_BATCHED)
            t0 = time.perf_counter()
            joins = st.nj()
            out["config%d" % cfg] = {"sequences": len(s), "joins": len(joins),
                                     "nj_ms": round((time.perf_counter() - t0) * 1e3, 1)}
            st.free()
        print(json.dumps({"nj": out}), flush=True)
        ctx.close()
        return 0
    if args.long_only:
        print(json.dumps({"long_pairs": long_pairs_bench(M, ctx)}), flush=True)
        ctx.close()
        return 0
    if args.serve_only:
        served = served_bench(
            M, ctx, args.serve_jobs, load=args.serve_load,
            sets=[(kv.split("=", 1)[0], float(kv.split("=", 1)[1])) for kv in args.serve_set],
            what={"msa": M.SERVE_MSA, "tree": M.SERVE_TREE, "scores": M.SERVE_SCORES}[args.serve_what])
        print(json.dumps({"served": served}), flush=True)
        ctx.close()
        return 0
    tunes = {}
    for kv in args.tune:
        k, v = kv.split("=", 1)
        ctx.tune_set(k, int(v))
        tunes[k] = int(v)
    st = ctx.upload(seqs)
    if args.profile_set:
        copy = ctx.copy_probe(1 << 30, 1)
        ms = st.all_pairs_resident(M.FORM_BATCHED)
        tb = traceback_bench(M, ctx, st, L)
        print(json.dumps({"copy_GBps": round(copy, 1), "batched_ms": round(ms, 3),
                          "flag_kernel_ms": tb["flag_kernel_ms"]}), flush=True)
        ctx.close()
        return 0
    if args.traceback_only:
        print(json.dumps({"traceback": traceback_bench(M, ctx, st, L)}), flush=True)
        ctx.close()
        return 0
    print("bench: %d sequences, %.4g cells (i<j), planning" % (len(seqs), cells), file=sys.stderr,
          flush=True)

    for _ in range(args.warmup):
        if not args.batched_only:
            st.all_pairs_resident(M.FORM_PLAIN)
        st.all_pairs_resident(M.FORM_BATCHED)
    print("bench: warmup done", file=sys.stderr, flush=True)

    def timed(form):
        walls, devs = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            devs.append(st.all_pairs_resident(form))
            walls.append(time.perf_counter() - t0)
        return walls, devs

    walls_b, devs_b = timed(M.FORM_BATCHED)
    walls_p, devs_p = ([], []) if args.batched_only else timed(M.FORM_PLAIN)
    print("bench: timed steps done", file=sys.stderr, flush=True)
    stats_b = st.launch_stats(M.FORM_BATCHED)  # one extra pass, per-launch events

    sec_b = float(np.mean(walls_b))
    sec_p = float(np.mean(walls_p)) if walls_p else None
    gcups_b = cells / sec_b / 1e9
    gcups_p = cells / sec_p / 1e9 if sec_p else None

    def kernel_of(cls):
        if cls <= -5000:
            return "score8_kernel<%d, true>" % (-5000 - cls), OPS_PER_CELL, PEAK_INT16_TOPS, "int16x2"
        if cls <= -4000:
            return "score8_kernel<%d, false>" % (-4000 - cls), OPS_PER_CELL, PEAK_INT16_TOPS, "int16x2"
        if cls <= -3000:
            return "score5_kernel<%d, 1, stacked>" % (-3000 - cls), OPS_PER_CELL, PEAK_INT16_TOPS, "int16x2"
        if cls <= -2000:
            return "score6_kernel<%d>" % (-2000 - cls), OPS_PER_CELL, PEAK_INT16_TOPS, "int16x2"
        if cls <= -1000:
            w, k = divmod(-cls - 1000, 100)
            return "score5_kernel<%d, %d>" % (k, w), OPS_PER_CELL, PEAK_INT16_TOPS, "int16x2"
        return "score_pairs_kernel<%d>" % cls, OPS_PER_CELL, PEAK_INT32_TOPS, "int32"

    peak_meas = {"int32": round(ctx.valu_peak(False), 2), "int16x2": round(ctx.valu_peak(True), 2)}
    cls, npairs, lcells, lms = max(stats_b, key=lambda r: r[3])
    kname, opc, peak, kind = kernel_of(cls)
    achieved = lcells * opc / (lms * 1e-3) / 1e12
    roof = {
        "bound": "valu",
        "kernel": kname,
        "achieved": round(achieved, 3),
        "peak": round(peak, 2),
        "unit": "Tops/s (%d algorithmic int ops per cell; %s VALU peak)" % (opc, kind),
        "frac": round(achieved / peak, 4),
        "peak_measured": peak_meas[kind],
        "traffic": profiled_traffic(kname),
        "traffic_source": "profiles/counters_lognormal4000.md (rocprofv3 FETCH_SIZE, WRITE_SIZE, copy-calibrated)",
        "launch_ms": round(lms, 3),
        "launch_cells": lcells,
    }
    launches = [{"kernel": kernel_of(c)[0], "pairs": p, "cells": ce, "ms": round(m, 3),
                 "gcups": round(ce / (m * 1e-3) / 1e9, 1) if m > 0 else None}
                for c, p, ce, m in stats_b]
    tb = None if (args.no_traceback or args.batched_only) else traceback_bench(M, ctx, st, L)
    nj_ms = None
    if not args.batched_only:
        t0 = time.perf_counter()
        st.nj()  # guide tree of the set from the resident scores
        nj_ms = round((time.perf_counter() - t0) * 1e3, 1)
    # the set's plans and graph (whose parallel branches hold a stream of
    # their own) are dropped before the other workloads
    st.free()
    tunables = dict(ctx.tuning, **tunes)
    config1 = None if (args.no_long or args.batched_only) else config1_bench(M, ctx)
    long_pairs = None if (args.no_long or args.batched_only) else long_pairs_bench(M, ctx)
    served = None
    if not (args.no_serve or args.batched_only):
        # the served loop on a context of its own: the streams the earlier
        # workloads left (the alignment form's second stream) otherwise share
        # the hardware queues with its 16 (offline ~300 against ~355 jobs/s)
        ctx.close()
        ctx = M.Context(device=0)
        served = served_bench(
            M, ctx, args.serve_jobs, load=args.serve_load,
            sets=[(kv.split("=", 1)[0], float(kv.split("=", 1)[1])) for kv in args.serve_set],
            what={"msa": M.SERVE_MSA, "tree": M.SERVE_TREE, "scores": M.SERVE_SCORES}[args.serve_what])

    line = {
        "metric": "all-pairs NW GCUPS (bit-exact), batched form",
        "value": round(gcups_b, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_b * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "lognormal-4000 all-pairs NW scores (BLOSUM62, go=10, ge=2)",
                   "sequences": len(seqs), "pairs_i_lt_j": len(seqs) * (len(seqs) - 1) // 2,
                   "cells_i_lt_j": cells, "self_cells": self_cells},
        "plain_gcups": round(gcups_p, 2) if gcups_p else None,
        "batched_over_plain": round(gcups_b / gcups_p, 3) if gcups_p else None,
        "device_ms_batched": round(float(np.mean(devs_b)), 3),
        "device_ms_plain": round(float(np.mean(devs_p)), 3) if devs_p else None,
        "roofline": roof,
        "batched_launches": launches,
        "tunables": tunables,
        "traceback": tb,
        "nj_ms": nj_ms,
        "config1": config1,
        "long_pairs": long_pairs,
        "served": served,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(seqs),
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
