"""Benchmark: Voronoi cells/s of the plain GPU path on one MI355X.

Workload (default): unit cube, 1,000,000 uniform f64 generators (seeded PCG32, seed 2), grid
100^3 bins, one SphereGeometry wall centred (0.5, 0.5, 0.5) with r = 0.5 (about 524k surviving
cells). One step = set_generators from HBM-resident points (wall filter, compaction, binning) plus
calculate() (every cell, full cells kept, as the reference's calculate stores them). The points
are uploaded once before the timed region.

Prints one JSON line. ``roofline`` is the FP64 roofline of the cell kernel (the path has no dense
contraction, so neither HBM nor matrix cores bound it): the algorithmic flops of the launch,
counted per cell by the oracle with the same rules, divided by the kernel time measured with HIP
events on the library's stream. Its ``issue`` object gives what does bound the kernel, the issue
rate of vector instructions against one wave64 instruction per SIMD every 4 cycles, from the
per-cell instruction counts of the committed profile. ``cpu_baseline`` times the CPU oracle (a C++
restatement of the reference's calculate) on the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Voronoi cells/s (f64, 1M uniform) batched vs plain GPU; p50/p99 serve latency"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--grid", type=int, default=100)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--cpu-threads", type=int, default=16)
    p.add_argument("--reduced", action="store_true", help="keep reductions only, not full cells")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="only run a device-to-device copy of MIB MiB (calibrates HBM byte counters)")
    p.add_argument("--no-batched", action="store_true", help="skip the batched vs plain comparison")
    p.add_argument("--batch-count", type=int, default=4096)
    p.add_argument("--batch-steps", type=int, default=3)
    p.add_argument("--no-serve", action="store_true", help="skip the served workload")
    p.add_argument("--serve-requests", type=int, default=100)
    p.add_argument("--serve-rate", type=float, default=2000.0)
    p.add_argument("--serve-budget-gb", type=float, default=64.0)
    p.add_argument("--serve-sweep-requests", type=int, default=30)
    p.add_argument("--serve-sweep-rates", type=str, default="5,50,500")
    p.add_argument("--no-clustered", action="store_true", help="skip the 4M clustered workload")
    p.add_argument("--only", choices=["batched", "clustered", "serve", "cpu", "scale"], default=None,
                   help="run one secondary workload alone and print its JSON object (for profiling)")
    p.add_argument("--no-plain", action="store_true",
                   help="--only batched / --only serve: skip the plain comparison legs")
    p.add_argument("--no-launch-tune", action="store_true",
                   help="keep the default launch shape instead of vd_*_autotune_launch's pick")
    p.add_argument("--no-scale", action="store_true", help="skip the 100M-generator tessellation")
    p.add_argument("--scale-n", type=int, default=100_000_000)
    p.add_argument("--clustered-n", type=int, default=4_000_000)
    p.add_argument("--clustered-targets", type=str, default="0.25,0.35,0.7",
                   help="generators-per-bin targets of the density rule tried for the clustered workload")
    return p.parse_args()


def targets_of(a) -> list:
    return [float(x) for x in a.clustered_targets.split(",") if x]


def serve_run(reqs, rate: float, budget_gb: float) -> dict:
    """One real-time run of the server over ``reqs``: every request is submitted once its arrival
    time has passed, the server is stepped (continuous batching) while work is pending;
    latency = finish - arrival on the monotonic clock."""
    from vorothree_amd import BoundingBox, Server
    from vorothree_amd import workloads as W

    unit = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    s = Server(hbm_budget=int(budget_gb * (1 << 30)), max_slots=64)
    # warm the kernels and allocator once with a small job (not counted)
    w = s.submit(unit, None, [], W.PCG32(999).uniform(5000, (0, 0, 0), (1, 1, 1)))
    while s.info(w)["state"] != Server.FINISHED:
        s.step()
    s.release(w)
    ids = [None] * len(reqs)
    nxt = 0
    t0 = time.monotonic()
    done = 0
    last = 0.0
    while done < len(reqs):
        now = time.monotonic() - t0
        if now - last > 10.0:
            print(f"serve {rate:g} req/s: t={now:.1f} s, {done} of {len(reqs)} finished", file=sys.stderr,
                  flush=True)
            last = now
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            r = reqs[nxt]
            ids[nxt] = s.submit(unit, None, [], r.points, iterations=r.iterations)
            nxt += 1
        p = s.pending()
        if p["queued"] or p["running"]:
            done += s.step()
        elif nxt < len(reqs):
            time.sleep(max(0.0, reqs[nxt].arrival - (time.monotonic() - t0)))
    wall = time.monotonic() - t0
    lat = []
    cells = 0
    grids = []
    for r, j in zip(reqs, ids):
        inf = s.info(j)
        lat.append(inf["finish"] - (t0 + r.arrival))
        cells += inf["total_cells"]
        grids.append(s.grid(j))
    lat = np.asarray(lat)
    st = s.stats()
    mem = s.memory()
    s.close()
    print(f"serve {rate:g} req/s: done in {wall:.2f} s", file=sys.stderr, flush=True)
    return {"rate_req_s": rate, "requests": len(reqs), "cells": int(cells), "wall_s": wall,
            "cells_per_s": cells / wall, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "latency_max_s": float(lat.max()),
            "server_steps": st["steps"], "device_bytes_peak": mem["peak"], "budget_bytes": mem["budget"],
            "_grids": grids}


def serve_plain(reqs, grids) -> dict:
    """The same requests without the server: one Tessellation per job, in arrival order, each run
    eagerly to completion (set_generators from the host, calculate, and for a Lloyd job relax +
    calculate for its remaining rounds), on the grid the server used for it."""
    from vorothree_amd import Algorithm3DGrid, BoundingBox, Tessellation, _hip

    unit = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    _hip.synchronize()
    t0 = time.perf_counter()
    cells = 0
    last = t0
    for k, (r, grid) in enumerate(zip(reqs, grids)):
        t = Tessellation(unit, Algorithm3DGrid(*grid, unit), keep_cells=False)
        t.set_generators(r.points)
        t.calculate()
        cells += t.count_cells()
        for _ in range(r.iterations - 1):
            t.relax()
            t.calculate()
            cells += t.count_cells()
        t.close()
        if time.perf_counter() - last > 10.0:
            last = time.perf_counter()
            print(f"serve plain: {k + 1} of {len(reqs)} jobs", file=sys.stderr, flush=True)
    _hip.synchronize()
    dt = time.perf_counter() - t0
    return {"cells": int(cells), "wall_s": dt, "cells_per_s": cells / dt}


def served(count: int, rate: float, budget_gb: float, sweep_count: int, sweep_rates, with_plain: bool = True) -> dict:
    """Served workload: the first ``count`` requests of a Poisson stream (rate req/s) of
    tessellations, 1e3..2e6 generators, 50% clustered, 30% Lloyd jobs of 10-50 iterations, jobs
    without a grid (the server picks it with vd_grid_for_points). Then a rate sweep over the first
    ``sweep_count`` requests of the same stream (same generators, arrival times rescaled)."""
    import dataclasses

    from vorothree_amd import workloads as W

    reqs = W.config5(count=max(count, sweep_count), rate=rate)
    print(f"serve: {len(reqs)} requests generated", file=sys.stderr, flush=True)
    main_run = serve_run(reqs[:count], rate, budget_gb)
    grids = main_run.pop("_grids")
    plain = None
    if with_plain:
        progress("serve: the same requests through the plain path, one tessellation per job")
        plain = serve_plain(reqs[:count], grids)
    sweep = []
    for r in sweep_rates:
        scaled = [dataclasses.replace(q, arrival=q.arrival * rate / r) for q in reqs[:sweep_count]]
        run = serve_run(scaled, r, budget_gb)
        run.pop("_grids")
        sweep.append(run)
    out = {"workload": f"first {count} requests of a Poisson stream at {rate:g} req/s: 1e3..2e6 "
                       "generators (log-uniform) in the unit cube, 50% uniform / 50% 8-cluster "
                       "sigma 0.05, 30% Lloyd jobs of 10-50 iterations; grid per job by "
                       "vd_grid_for_points at its default target (0.35 per bin); "
                       f"HBM budget {budget_gb:g} GiB, 64 slots; "
                       "jobs of up to 65,536 points in the small-job lane, larger ones stepped on a worker",
           "lloyd_jobs": int(sum(q.iterations > 1 for q in reqs[:count])),
           "note": "the offered load is far above one GPU's capacity at 2000 req/s, so its latency "
                   "is mostly queueing; the sweep lowers the rate on the first requests"}
    out.update(main_run)
    if plain is not None:
        out["plain"] = plain
        out["speedup_served_vs_plain"] = main_run["cells_per_s"] / plain["cells_per_s"]
    out["sweep"] = sweep
    return out


def seen_per_bin(points, grid) -> float:
    """Generators per bin as seen by the average generator (points in the unit cube): sum over
    bins of c^2, over n."""
    g = np.asarray(grid, dtype=np.int64)
    idx = np.minimum((points * g).astype(np.int64), g - 1)
    key = (idx[:, 2] * g[1] + idx[:, 1]) * g[0] + idx[:, 0]
    _, c = np.unique(key, return_counts=True)
    return float((c.astype(np.float64) ** 2).sum() / len(points))


def clustered(n: int, targets, steps: int, launch_tune: bool = True) -> dict:
    """Uneven density: n generators, 80% from 64 Gaussian clusters (sigma 0.02) + 20% uniform, in
    the unit cube, no walls. The grid is tuned over generators-per-bin targets of the density rule
    (vd_grid_for_points), each with its own launch shape (vd_tess_autotune_launch); the best pair
    is timed over ``steps`` calculate() calls. Each grid also reports the generators per bin seen
    by the average generator, sum(c^2) / n over its bins."""
    from vorothree_amd import Algorithm3DGrid, BoundingBox, Tessellation, grid_for_points, _hip
    from vorothree_amd import workloads as W

    w = W.config3(n=n)
    b = BoundingBox(w.lo, w.hi)
    dpts = _hip.DeviceArray(w.points)
    _hip.synchronize()
    tried = []
    best = None
    for tgt in targets:
        grid = grid_for_points(b, w.points, tgt)
        t = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=False)
        t.set_generators_device(dpts.ptr.value, n)
        t.calculate()                              # builds the shell prefix
        launch = t.autotune_launch() if launch_tune else None
        _hip.synchronize()
        c0 = time.perf_counter()
        t.calculate()
        dt = time.perf_counter() - c0
        tried.append({"per_bin_target": tgt, "grid": list(grid), "seen_per_bin": seen_per_bin(w.points, grid),
                      "launch": launch, "calc_ms": dt * 1e3})
        if best is None or dt < best[0]:
            best = (dt, tgt, grid, launch)
        t.close()
    _, tgt, grid, launch = best
    t = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=False)
    if launch:
        t.set_launch(**launch)
    t.set_generators_device(dpts.ptr.value, n)
    t.calculate()
    _hip.synchronize()
    c0 = time.perf_counter()
    for _ in range(steps):
        t.set_generators_device(dpts.ptr.value, n)
        t.calculate()
    _hip.synchronize()
    dt = (time.perf_counter() - c0) / steps
    st = t.stats()
    cells = t.count_cells()
    out = {"workload": f"unit cube, {n} f64 generators: 80% from 64 Gaussian clusters (sigma 0.02, "
                       "re-drawn inside the box) + 20% uniform, seed 3, no walls",
           "grid": list(grid), "per_bin_target": tgt, "tuning": tried, "launch": launch, "cells": cells,
           "ms_per_step": dt * 1e3, "cells_per_s": cells / dt,
           "path": "set_generators(device) + calculate(), reduced outputs",
           "cells_first_tier": int(st[0]), "cell_runs_larger_tiers": int(st[1]),
           "volume_sum": float(np.sum(t.volumes()))}
    t.close()
    return out


def scale(n: int, steps: int) -> dict:
    """One tessellation at HBM scale: n uniform generators drawn on the device by
    random_generators (seed 2026) in the unit cube, no walls, a grid of about one generator per
    bin, full cells kept. One calculate() allocates the arenas, then ``steps`` are timed; the
    device memory the handle holds is the drop in free device memory."""
    from vorothree_amd import Algorithm3DGrid, BoundingBox, Tessellation, grid_for, _hip

    b = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    grid = grid_for(b, n, 1.0)
    _hip.synchronize()
    free0 = _hip.mem_free()
    t = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=True, graphs=False)
    c0 = time.perf_counter()
    t.random_generators(n, seed=2026)
    _hip.synchronize()
    gen_s = time.perf_counter() - c0
    c0 = time.perf_counter()
    t.calculate()
    _hip.synchronize()
    first_s = time.perf_counter() - c0
    c0 = time.perf_counter()
    for _ in range(steps):
        t.calculate()
    _hip.synchronize()
    dt = (time.perf_counter() - c0) / steps
    st = t.stats()
    cells = t.count_cells()
    held = free0 - _hip.mem_free()
    out = {"workload": f"unit cube, {n} uniform f64 generators from random_generators (seed 2026), no walls, "
                       f"grid {grid[0]}^3, full cells kept",
           "generators": n, "cells": cells, "grid": list(grid), "steps": steps,
           "ms_per_step": dt * 1e3, "cells_per_s": cells / dt, "first_calculate_ms": first_s * 1e3,
           "random_generators_ms": gen_s * 1e3, "device_bytes_held": held,
           "arena_vertices": st[6], "arena_face_indices": st[7],
           "volume_sum": float(np.sum(t.volumes()))}
    t.close()
    return out


BATCH_TUNE = (0.25, 0.35, 0.5, 1.0, 2.0)


def batched_vs_plain(count: int, steps: int, tune=None, launch_tune: bool = True, plain_legs: bool = True) -> dict:
    """Mixed-size workload: ``count`` tessellations with log-uniform generator counts in
    [64, 65536], boxes with edges U[0.5, 2], 30% sphere walls, 20% plane walls. Grids: about 2
    generators per bin (``tune`` None), or the batch autotuner's pick among the generators-per-bin
    targets in ``tune`` (vd_batch_autotune); the plain path runs the same grids. Batched: all
    members in one vd_batch_calculate (wall filter, binning and one persistent cell launch for
    all). Plain: one Tessellation per member, each running its own set_generators + calculate
    launch sequence, eagerly; the same with graph replay of repeated sequences is reported beside
    it. All start from points already in device memory. The launch shape is the batch's
    vd_batch_autotune_launch pick (``launch_tune``); the plain tessellations keep the library's
    default shape (one tuned shape for thousands of small launches would not be theirs)."""
    from vorothree_amd import (Algorithm3DGrid, Batch, BoundingBox, PlaneGeometry, SphereGeometry,
                               Tessellation, Wall, _hip)
    from vorothree_amd import workloads as W

    ws = W.config4(count=count)

    def walls_of(w):
        out = []
        for wl in w.walls:
            g = SphereGeometry(wl[2], wl[3]) if wl[0] == "sphere" else PlaneGeometry(wl[2], wl[3])
            out.append(Wall(wl[1], g))
        return out

    b = Batch()
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), None if tune else w.grid, walls_of(w), w.points)
    b.upload()
    best = b.autotune(tune) if tune else None
    launch = b.autotune_launch() if launch_tune else None
    b.calculate()                       # warmup (also builds the shell prefixes)
    _hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.calculate()
    _hip.synchronize()
    tb = (time.perf_counter() - t0) / steps
    cells = int(sum(b.info(t)["cells"] for t in range(len(b))))
    grids = [b.info(t)["grid"] for t in range(len(b))]
    kern_b = float(b.stats()[4]) * 1e-6
    replays_b = int(b.stats()[8])
    b.close()

    allpts = np.concatenate([w.points for w in ws], axis=0)
    offs = np.concatenate([[0], np.cumsum([len(w.points) for w in ws])])
    dev = _hip.DeviceArray(allpts)

    def plain(graphs):
        ts = []
        for w, grid in zip(ws, grids):
            bb = BoundingBox(w.lo, w.hi)
            t = Tessellation(bb, Algorithm3DGrid(*grid, bb), keep_cells=False, graphs=graphs)
            for wl in walls_of(w):
                t.add_wall(wl)
            ts.append(t)

        def plain_step():
            for k, t in enumerate(ts):
                t.set_generators_device(dev.ptr.value + 24 * int(offs[k]), len(ws[k].points))
                t.calculate()

        for _ in range(2 if graphs else 1):
            plain_step()
        _hip.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plain_step()
        _hip.synchronize()
        tp = (time.perf_counter() - t0) / steps
        cells_p = int(sum(t.count_cells() for t in ts))
        for t in ts:
            t.close()
        return tp, cells_p

    # each tessellation of the workload is computed once, so the plain path launches eagerly; the
    # rerun of the same tessellations every step would let them replay graphs, reported apart
    tp = tpg = cells_p = None
    if plain_legs:
        progress(f"batched: {count} tessellations in one batch done; plain, one engine each")
        tp, cells_p = plain(False)
        progress("batched: plain done; plain with graph replay")
        tpg, _ = plain(True)
    dev.close()
    bins = sum(int(np.prod(g)) for g in grids)
    grid_note = (f"autotuned grids (vd_batch_autotune over {list(tune)} generators per bin; picked {best:g})"
                 if tune else "grids of about 2 generators per bin")
    return {"workload": f"{count} tessellations, generators log-uniform in [64, 65536], "
                        f"boxes U[0.5,2]^3, 30% sphere / 20% plane walls, {grid_note}",
            "per_bin": best if tune else 2.0, "launch": launch, "generators_per_bin": float(offs[-1]) / bins,
            "generators": int(offs[-1]), "cells": cells, "cells_plain": cells_p,
            "batched_cells_per_s": cells / tb, "batched_ms": tb * 1e3, "batched_first_round_ms": kern_b * 1e3,
            "batched_graph_replays": replays_b,
            "plain_cells_per_s": cells_p / tp if tp else None, "plain_ms": tp * 1e3 if tp else None,
            "speedup_batched_vs_plain": tp / tb if tp else None,
            "plain_graph_replay_cells_per_s": cells_p / tpg if tpg else None,
            "speedup_batched_vs_plain_graph_replay": tpg / tb if tpg else None, "steps": steps}


def cpu_table(threads: int) -> dict:
    """The CPU restatement of calculate() (oracle/, C++) on this host for the single-tessellation
    workloads: 1k in [0,10]^3 (grid 10^3), 1M uniform with the sphere wall (grid 100^3) and 4M
    clustered (grid 474^3, the density rule at 0.35 generators per bin). set_generators (wall filter
    and binning) runs on one thread as in the reference; calculate on ``threads`` threads and, for a
    bounded sample of every k-th cell, on one. Median of 3 calculates (1 for the 4M case)."""
    from oracle import OracleTess
    from vorothree_amd import workloads as W
    rows = []
    cases = [(W.config1(), 3), (W.config2(), 3), (W.config3(grid=(474, 474, 474)), 1)]
    for w, reps in cases:
        o = OracleTess(w.lo, w.hi, *w.grid)
        for wl in w.walls:
            o.add_wall_sphere(wl[1], wl[2], wl[3])
        c0 = time.perf_counter()
        o.set_generators(w.points)
        bdt = time.perf_counter() - c0
        ts = []
        for _ in range(reps):
            c0 = time.perf_counter()
            o.calculate(threads=threads)
            ts.append(time.perf_counter() - c0)
        cdt = float(np.median(ts))
        n = o.count_cells()
        _, _, _, work = o.reduced()
        step = max(1, n // 50_000)
        sub = np.arange(0, n, step, dtype=np.uint32)
        c0 = time.perf_counter()
        o.calculate(subset=sub, threads=1)
        sdt = time.perf_counter() - c0
        rows.append({"workload": w.name, "grid": list(w.grid), "generators": int(len(w.points)), "cells": int(n),
                     "threads": threads, "set_generators_ms": bdt * 1e3, "calculate_ms": cdt * 1e3,
                     "cells_per_s": n / cdt, "single_thread_cells_per_s": len(sub) / sdt,
                     "single_thread_sample": f"every {step}th cell ({len(sub)} cells)",
                     "flops_per_cell": float(work[:, 0].mean())})
        progress(f"cpu: {w.name} {n / cdt:.4g} cells/s on {threads} threads")
        del o
    return {"host_cpu": host_cpu(), "kind": "port (oracle/, the C++ restatement of calculate())", "rows": rows}


def copy_probe(mib: int) -> None:
    """One hipMemcpy device-to-device of a known size, so FETCH_SIZE / WRITE_SIZE counters read
    under rocprofv3 can be scaled to bytes."""
    from vorothree_amd import _hip
    n = mib * (1 << 20) // 8
    src = _hip.DeviceArray(np.ones(n))
    dst = _hip.DeviceArray(np.zeros(n))
    _hip.synchronize()
    rc = _hip.hip().hipMemcpy(dst.ptr, src.ptr, n * 8, 3)
    _hip.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "rc": rc}), flush=True)


def host_cpu() -> str:
    """CPU model of the host (first 'model name' of /proc/cpuinfo) and its logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def progress(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}), flush=True)
        sys.exit(2)
    if a.only == "cpu":
        print(json.dumps(cpu_table(a.cpu_threads)), flush=True)
        return
    if a.only == "batched":
        print(json.dumps(batched_vs_plain(a.batch_count, a.batch_steps, None, not a.no_launch_tune, not a.no_plain)),
              flush=True)
        return
    if a.only == "clustered":
        print(json.dumps(clustered(a.clustered_n, targets_of(a), 3, not a.no_launch_tune)), flush=True)
        return
    if a.only == "scale":
        print(json.dumps(scale(a.scale_n, 2)), flush=True)
        return
    if a.only == "serve":
        rates = [float(x) for x in a.serve_sweep_rates.split(",") if x]
        print(json.dumps(served(a.serve_requests, a.serve_rate, a.serve_budget_gb, a.serve_sweep_requests, rates,
                                not a.no_plain)),
              flush=True)
        return

    from vorothree_amd import Algorithm3DGrid, BoundingBox, SphereGeometry, Tessellation, Wall
    from vorothree_amd import _hip
    from vorothree_amd import workloads as W

    w = W.config2(n=a.n, grid=a.grid)
    b = BoundingBox(w.lo, w.hi)
    # eager launches, so the kernel time of the tier-0 launch is measured on its own
    t = Tessellation(b, Algorithm3DGrid(*w.grid, b), keep_cells=not a.reduced, graphs=False)
    t.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5)))
    dpts = _hip.DeviceArray(w.points)
    _hip.synchronize()
    launch = None
    if not a.no_launch_tune:
        t.set_generators_device(dpts.ptr.value, a.n)
        launch = t.autotune_launch()
        progress(f"plain: launch shape {launch}")

    def step():
        t.set_generators_device(dpts.ptr.value, a.n)
        t.calculate()

    for _ in range(a.warmup):
        step()
    _hip.synchronize()
    kern_us = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
        kern_us.append(float(t.stats()[4]))
    _hip.synchronize()
    dt = time.perf_counter() - t0
    st = t.stats()
    cells = t.count_cells()
    ms = 1000.0 * dt / a.steps
    value = cells / (dt / a.steps)
    kern_s = float(np.mean(kern_us)) * 1e-6

    # the same steps with the first round replayed from a captured graph (end-to-end rate only)
    tg = Tessellation(b, Algorithm3DGrid(*w.grid, b), keep_cells=not a.reduced, graphs=True)
    tg.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5)))
    if launch:
        tg.set_launch(**launch)
    for _ in range(max(2, a.warmup)):
        tg.set_generators_device(dpts.ptr.value, a.n)
        tg.calculate()
    _hip.synchronize()
    g0 = time.perf_counter()
    for _ in range(a.steps):
        tg.set_generators_device(dpts.ptr.value, a.n)
        tg.calculate()
    _hip.synchronize()
    gdt = time.perf_counter() - g0
    graph_rate = tg.count_cells() / (gdt / a.steps)
    graph_replays = int(tg.stats()[8])
    assert np.array_equal(tg.volumes().view(np.uint64), t.volumes().view(np.uint64))
    tg.close()

    # the same steps with the points handed over in host memory (PCIe upload inside the step):
    # reported beside the headline, never as it
    for _ in range(max(1, a.warmup)):
        t.set_generators(w.points)
        t.calculate()
    _hip.synchronize()
    h0 = time.perf_counter()
    for _ in range(a.steps):
        t.set_generators(w.points)
        t.calculate()
    _hip.synchronize()
    host_rate = t.count_cells() / ((time.perf_counter() - h0) / a.steps)

    progress(f"plain: {value:.4g} cells/s, graph replay {graph_rate:.4g} cells/s, host upload {host_rate:.4g} cells/s")
    # algorithmic work of one cell launch, counted by the oracle over the same cells
    cpu = None
    flops_per_cell = None
    if not a.no_cpu:
        from oracle import OracleTess
        o = OracleTess(w.lo, w.hi, *w.grid)
        o.add_wall_sphere(-1000, (0.5, 0.5, 0.5), 0.5)
        c0 = time.perf_counter()
        o.set_generators(w.points)            # wall filter + binning (single thread, as the reference)
        bdt = time.perf_counter() - c0
        c0 = time.perf_counter()
        o.calculate(threads=a.cpu_threads)
        cdt = time.perf_counter() - c0
        _, _, _, work = o.reduced()
        flops_per_cell = float(work[:, 0].mean())
        ncpu = o.count_cells()
        sub = np.arange(0, ncpu, max(1, ncpu // 50_000), dtype=np.uint32)
        c0 = time.perf_counter()
        o.calculate(subset=sub, threads=1)
        sdt = time.perf_counter() - c0
        cpu = {"value": ncpu / cdt, "unit": "cells/s", "cores": a.cpu_threads,
               "kind": "port",
               "sample": f"the whole workload: all {ncpu} cells of the timed "
                         f"tessellation, one calculate() on {a.cpu_threads} threads "
                         f"(std::thread, dynamic chunks of 64 cells); host nproc={os.cpu_count()}",
               "single_thread": {"value": len(sub) / sdt, "unit": "cells/s", "cores": 1,
                                 "sample": f"every {max(1, ncpu // 50_000)}th cell ({len(sub)} cells)"},
               "host_cpu": host_cpu(), "set_generators_ms": bdt * 1e3, "calculate_ms": cdt * 1e3,
               "flops_per_cell": None}
        cpu["flops_per_cell"] = flops_per_cell
        del o

    # algorithmic HBM bytes of the launch: per cell the generator, its bin item and bin offset read
    # (24 + 4 + 4), volume, centroid, face count, vertex count, status and arena offsets written
    # (8 + 24 + 4 + 4 + 4 + 24); the cell arena: 24 per vertex, 2 per face index, 1 + 4 per face
    nf = int(t.face_counts().sum())
    nv, ns = int(st[6]), int(st[7])
    vol = t.volumes()
    bytes_launch = cells * (24 + 4 + 4 + 8 + 24 + 4 + 4 + 4 + 24) + nv * 24 + ns * 2 + nf * 5
    traffic = None
    issue_counts = None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "k_cells_uniform1M_sphere_traffic.json")
    if os.path.exists(tfile) and a.n == 1_000_000 and a.grid == 100 and not a.reduced:
        with open(tfile) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        if prof.get("valu_insts_per_cell"):
            issue_counts = (prof["valu_insts_per_cell"], prof["salu_insts_per_cell"])
    roof = None
    if flops_per_cell is not None:
        achieved = cells * flops_per_cell / kern_s / 1e12
        roof = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                "kernel": ("k_cells<Tier1" if launch and launch["first_tier"] == 1
                           else f"k_cells<Tier0W<{launch['t0_waves'] if launch else 8}>")
                          + (", skip-ahead>" if launch and launch["first_probe"] else ">"),
                "kernel_ms": kern_s * 1e3,
                "flops_per_cell": flops_per_cell,
                "hbm": {"achieved": bytes_launch / kern_s / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": bytes_launch / kern_s / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_launch": bytes_launch}}
        if issue_counts:
            # what binds the kernel: vector-instruction issue. Peak: one wave64 VALU instruction
            # per SIMD every 4 cycles (1024 SIMDs at 2.4 GHz); the instructions per cell are the
            # SQ_INSTS_VALU / SQ_INSTS_SALU counts of the committed profile of this kernel
            valu_peak = 1024 * 2.4e9 / 4 / 1e9
            v_rate = issue_counts[0] * cells / kern_s / 1e9
            roof["issue"] = {"achieved": v_rate, "peak": valu_peak, "unit": "G VALU wave-instructions/s",
                             "frac": v_rate / valu_peak, "valu_per_cell": issue_counts[0],
                             "salu_per_cell": issue_counts[1],
                             "salu_frac": issue_counts[1] * cells / kern_s / 1e9 / valu_peak}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG32 uniform points)",
        "config": {"workload": "plain GPU, one tessellation: unit cube, 1M uniform f64 generators, "
                               "grid 100^3, sphere wall c=(0.5,0.5,0.5) r=0.5",
                   "generators": a.n, "cells": cells, "grid": list(w.grid),
                   "keep_cells": not a.reduced, "path": "set_generators(device) + calculate()",
                   "launch": launch or {"t0_waves": 7, "queue_grain": 1, "first_tier": 0, "first_probe": 0},
                   "launch_note": "tier-0 waves per SIMD, cells per queue atomic and first tier: "
                                  + ("vd_tess_autotune_launch's pick on this workload (coordinate "
                                     "descent over skip-ahead 0/1, waves 5/6/7/8, grain 1/2/4/8, first "
                                     "tier 0/1), before the warmup"
                                     if launch else "the library defaults"),
                   "cells_retried_large_tier": st[1], "shell_prefix_entries": st[3],
                   "graph_replay_cells_per_s": graph_rate, "graph_replays": graph_replays,
                   "host_upload_cells_per_s": host_rate,
                   "volume_sum": float(vol.sum())},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        progress(f"cpu baseline: {cpu['value']:.4g} cells/s on {a.cpu_threads} threads")
    if not a.no_batched:
        progress("batched: grids of about 2 generators per bin")
        out["batched"] = batched_vs_plain(a.batch_count, a.batch_steps, None, not a.no_launch_tune)
        progress("batched: autotuned grids")
        out["batched"]["autotuned"] = batched_vs_plain(a.batch_count, a.batch_steps, BATCH_TUNE,
                                                       not a.no_launch_tune)
    if not a.no_clustered:
        progress(f"clustered: {a.clustered_n} points")
        out["clustered"] = clustered(a.clustered_n, targets_of(a), 3, not a.no_launch_tune)
    if not a.no_scale:
        progress(f"scale: {a.scale_n} points")
        out["scale"] = scale(a.scale_n, 2)
    if not a.no_serve:
        rates = [float(x) for x in a.serve_sweep_rates.split(",") if x]
        out["serve"] = served(a.serve_requests, a.serve_rate, a.serve_budget_gb, a.serve_sweep_requests, rates)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
