"""Measurement of the served form (mst_server_*): a workload of jobs run
offline, alone (unloaded) and under Poisson arrivals (loaded), with p50/p99
latency and throughput. Used by bench.py and by the autotuner's search of
the server settings (mestrado_amd.autotune --serve)."""
import time

import numpy as np

from . import SERVE_MSA, SERVE_SCORES, SERVE_TREE, Server, job_sizes, synth


def job_workload(njobs, seed=0):
    """njobs jobs: sizes power-law (exponent 1.5) on 8..512 sequences, each a
    tree-evolved family of lengths 100..2000 from a per-job seed."""
    sizes = job_sizes(seed, njobs)
    return [synth(4, seed * 1000003 + k, int(n)) for k, n in enumerate(sizes)]


def measure(ctx, jobs, what=None, load=0.75, seed=0, sets=(), unloaded_every=4, note=None,
            budget_bytes=16 << 30):
    """Offline: every job submitted at once, drained. Unloaded: one job at a
    time, every `unloaded_every`-th job (a sample over all sizes, to bound the
    run time). Loaded: Poisson arrivals at `load` x the offline job rate;
    latency runs from a job's scheduled arrival to the end of its last stage.
    sets: (key, value) server settings (mst_server_set)."""
    if what is None:
        what = SERVE_MSA
    note = note or (lambda msg: None)
    njobs = len(jobs)
    jcells = []
    for j in jobs:
        L = np.array([len(x) for x in j], dtype=np.int64)
        jcells.append(int((L.sum() ** 2 - (L * L).sum()) // 2))
    total_cells = float(sum(jcells))
    srv = Server(ctx, budget_bytes)
    settings = dict(srv.settings)  # the tuning table's served settings
    for k, v in sets:
        srv.set(k, v)
        settings[k] = v
    # offline
    t0 = time.perf_counter()
    ids = [srv.submit(j, what) for j in jobs]
    srv.drain()
    t_off = time.perf_counter() - t0
    for jid in ids:
        srv.release(jid)
    srv.completed(1 << 16)
    note("offline %.2f s, %.3g cells/s estimate" % (t_off, srv.stats()["cells_per_s"]))
    # unloaded
    lat_u, info_u = [], []
    for j in jobs[::unloaded_every]:
        jid = srv.submit(j, what)
        srv.drain()
        info = srv.job(jid)
        lat_u.append(info["t_done"] - info["t_submit"])
        tt = srv.times(jid)
        info_u.append((lat_u[-1], info["n"], (tt[2] - tt[0], (tt[3] - tt[2]) if tt[3] else 0.0,
                                               (tt[4] - (tt[3] or tt[2]))), srv.job_stats(jid)))
        srv.release(jid)
        if len(lat_u) % 100 == 0:
            note("unloaded %d jobs" % len(lat_u))
    srv.completed(1 << 16)
    # loaded
    rng = np.random.default_rng(seed + 1)
    rate_off = njobs / t_off
    arr = np.cumsum(rng.exponential(1.0 / (load * rate_off), njobs))
    perm = rng.permutation(njobs)
    k, done, lat, first_of, stage_s, job_info = 0, 0, [], {}, [], []
    steps, step_s = 0, 0.0
    t0m = time.monotonic()
    t_last = t0m
    while done < njobs:
        now = time.monotonic() - t0m
        while k < njobs and arr[k] <= now:
            first_of[srv.submit(jobs[perm[k]], what)] = k
            k += 1
        ts = time.perf_counter()
        srv.step()
        step_s += time.perf_counter() - ts
        steps += 1
        for jid in srv.completed():
            info = srv.job(jid)
            lat.append(info["t_done"] - (t0m + arr[first_of[jid]]))
            tt = srv.times(jid)
            stage_s.append((tt[2] - tt[0], (tt[3] - tt[2]) if tt[3] else 0.0,
                            (tt[4] - (tt[3] or tt[2]))))
            js = srv.job_stats(jid)
            job_info.append((lat[-1], info["n"], stage_s[-1], js))
            t_last = max(t_last, info["t_done"])
            srv.release(jid)
            done += 1
            if done % 500 == 0:
                note("loaded %d jobs, %.2f s" % (done, time.monotonic() - t0m))
    t_load = t_last - t0m
    srv.close()
    # the slowest 1% of the jobs (those the p99 is about): their size and
    # where their time went, loaded and alone
    mean = lambda v: round(float(np.mean(v)), 3)

    def slowest(rows):
        slow = sorted(rows, key=lambda r: -r[0])[:max(1, len(rows) // 100)]
        return {
            "jobs": len(slow), "latency_ms": mean([r[0] * 1e3 for r in slow]), "n": mean([r[1] for r in slow]),
            "scores_ms": mean([r[2][0] * 1e3 for r in slow]), "tree_ms": mean([r[2][1] * 1e3 for r in slow]),
            "alignment_ms": mean([r[2][2] * 1e3 for r in slow]),
            "merge_batches": mean([r[3]["batches"] for r in slow]),
            "alignment_busy_ms": mean([r[3]["busy_s"] * 1e3 for r in slow]),
            "alignment_idle_ms": mean([r[3]["idle_s"] * 1e3 for r in slow]),
            "score_rounds": mean([r[3]["rounds"] for r in slow]),
            "first_round_ms": mean([r[3]["first_round_s"] * 1e3 for r in slow]),
            "admit_ms": mean([r[3]["admit_s"] * 1e3 for r in slow]),
        }
    ms = lambda v, q: round(float(np.percentile(v, q)) * 1e3, 3)
    stages = {SERVE_SCORES: "all-pairs scores", SERVE_TREE: "all-pairs scores + NJ tree",
              SERVE_MSA: "all-pairs scores + NJ tree + progressive alignment"}[what]
    return {
        "workload": "%d jobs, sizes power-law 1.5 on 8..512, lengths 100..2000, %.3g cells (i<j)"
                    % (njobs, total_cells),
        "job": stages,
        "unloaded_sample": "every %d-th job (%d jobs)" % (unloaded_every, len(lat_u)),
        "offline_jobs_per_s": round(rate_off, 1),
        "offline_gcups": round(total_cells / t_off / 1e9, 1),
        "load": load,
        "jobs_per_s": round(njobs / t_load, 1),
        "gcups": round(total_cells / t_load / 1e9, 1),
        "frac_of_offline": round((njobs / t_load) / rate_off, 3),
        "p50_ms": ms(lat, 50), "p99_ms": ms(lat, 99),
        "unloaded_p50_ms": ms(lat_u, 50), "unloaded_p99_ms": ms(lat_u, 99),
        "p99_over_unloaded_p99": round(float(np.percentile(lat, 99) / np.percentile(lat_u, 99)), 3),
        "loaded_stage_p50_ms": {name: ms([x[i] for x in stage_s], 50)
                                for i, name in enumerate(("scores", "tree", "alignment"))},
        "loaded_stage_p99_ms": {name: ms([x[i] for x in stage_s], 99)
                                for i, name in enumerate(("scores", "tree", "alignment"))},
        "loaded_slowest_1pct": slowest(job_info),
        "unloaded_slowest_1pct": slowest(info_u),
        "params": settings,
        "loaded_host": {"steps": steps, "ms_per_step": round(step_s / max(steps, 1) * 1e3, 4),
                        "loop_ms_per_step": round((t_last - t0m) / max(steps, 1) * 1e3, 4)},
    }
