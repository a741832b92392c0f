"""Benchmark: MGPCG pressure projection on one MI355X (Mcells/s at 1e-6 relative residual).

A step is one complete projection of the 384^3 dam-plus-spheres grid (divergence
-> ghost-fluid assembly -> FP16-preconditioned MGPCG with FP64 refinement to
||b - A p||/||b|| <= 1e-6 -> gradient subtract), inputs resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 384] [--precision fp16]
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth

# HBM bytes per launch of the timed kernel, measured with rocprofv3 FETCH_SIZE
# and WRITE_SIZE (each in its own run, scaled by a 1 GiB copy probe: reads 2x,
# writes 1x) -- profiles/uniform_384_sweep_and_tall_512_jacobi_hbm_traffic.md and
# profiles/uniform_384_sweep_pairs.md.
# The bench cannot read hardware counters itself; a configuration without a
# measurement reports null.
MEASURED_TRAFFIC_MB = {
    ("uniform", 384, "fp16", "smooth_fine", 2, False): 499.9,
    ("uniform", 384, "fp16", "smooth_fine", 2, True): 569.3,
    ("tall", 512, "fp32", "smooth_fine", 2): 286.3,
}


def sweep_launches(precision: str, sweeps: int, pairs: bool) -> int:
    """Finest-level sweep launches of one V-cycle (both sides): with pairs,
    two sweeps per launch (one more for an odd count per side)."""
    if pairs and precision != "fp32":
        return 2 * (sweeps // 2 + sweeps % 2)
    return 2 * sweeps


def kernel_bytes_per_cell(kernel: str, precision: str, sweeps: int = 2, pairs: bool = False) -> float:
    """Algorithmic (compulsory) HBM bytes per grid cell for one launch.

    smooth_fine: the red-black sweeps over the finest level, averaged over
    the launches one V-cycle times (per side `sweeps`).  Per sweep (the unit
    of work): read the packed face coefficients (4 values of the storage
    type, 4s B) and the FP32 right-hand side (4 B), write the new correction
    (s B); all but the first pre-smoothing sweep (zero guess) read the old
    correction (s B).  A launch of a pair of sweeps does two units of work
    and counts both (its own HBM traffic is about one sweep's: the measured
    traffic shows it).  The coarse correction is added by its own pass
    before the post-smoothing sweeps.
    residual_fine: read coefficients (4s), r (4), e (s); write t (s).
    cg_direction: read z (s), d_old (4), phi (4); write d (4), q (4).
    cg_update: read d, q, e, r (16); write e, r (8).
    Per cell of the liquid hulls (hull_cells): the finest level's passes
    neither read nor write the other cells.
    """
    s = 4 if precision == "fp32" else 2
    plain = 4 * s + 4 + 2 * s
    smooth = (2 * sweeps * plain - s) / sweep_launches(precision, sweeps, pairs) if sweeps > 0 else plain
    return {
        "smooth_fine": smooth,
        "cg_direction": s + 4 + 4 + 4 + 4,
        "cg_update": 24.0,
        "residual_fine": 4 * s + 4 + s + s,
    }[kernel]


def hull_cells(phi, tz: int = 32, ty: int = 8) -> int:
    """Cells of the finest level inside the liquid hulls of its 32 x 8 (z x y)
    column tiles: per tile, the planes from the first to the last x plane in
    which any of the tile's cells is liquid.  The hull-clamped passes (sweeps,
    residual, CG direction and update) read and write only these cells."""
    import numpy as np
    liq = np.asarray(phi) < 0
    nx, ny, nz = liq.shape
    py, pz = -ny % ty, -nz % tz
    inside = np.ones((ny + py, nz + pz), bool)
    if py or pz:
        liq = np.pad(liq, ((0, 0), (0, py), (0, pz)))
        inside[ny:, :] = False
        inside[:, nz:] = False
    t = liq.reshape(nx, (ny + py) // ty, ty, (nz + pz) // tz, tz).any(axis=(2, 4))
    has = t.any(axis=0)
    lo = np.where(has, t.argmax(axis=0), 0)
    hi = np.where(has, nx - 1 - t[::-1].argmax(axis=0), -1)
    planes = np.maximum(hi - lo + 1, 0)
    cols = inside.reshape((ny + py) // ty, ty, (nz + pz) // tz, tz).sum(axis=(1, 3))
    return int((planes * cols).sum())


def traffic_bytes(key):
    mb = MEASURED_TRAFFIC_MB.get(key)
    return None if mb is None else round(mb * 1e6)


def copy_probe(mib: int, reps: int) -> None:
    """A plain device copy of known size, for scaling the HBM byte counters
    (FETCH_SIZE / WRITE_SIZE) of a profiler run to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "reps": reps}))


def host_cpu() -> dict:
    """The host the CPU baseline ran on: processor model and logical CPUs."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def _median_solve(oracle, s, solver, runs=3):
    """Median wall time of `runs` FP64 projections of scene s (assembly ->
    divergence -> solve to 1e-6 -> gradient subtract; generation excluded)."""
    times, rep = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        _, _, _, rep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-6,
                                                 solver=solver)
        times.append(time.perf_counter() - t0)
    return statistics.median(times), rep


def cpu_baseline(n_all: int = 128, n_one: int = 64, runs: int = 3):
    """The reference's CPU MAC/PCG algorithm restated in C (oracle/fp_oracle.c)
    on bounded samples of the same workload (the dam + 8 spheres generator at
    reduced size): Jacobi-PCG (sparse-atlas-composite-projection.ts:1346-1457,
    the reference's own CPU solver) and an FP64 MG-PCG variant, each on all
    host cores and on one core, median of `runs`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from fluid_amd import scenes

    oracle.build()
    all_threads = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    s_all = scenes.dam_spheres_scene(n_all, seed=0)
    s_one = scenes.dam_spheres_scene(n_one, seed=0)
    out = {}
    try:
        for label, threads, s in (("all_cores", all_threads, s_all), ("one_core", 1, s_one)):
            oracle.set_threads(threads)
            for solver in ("jacobi", "mgpcg"):
                t, rep = _median_solve(oracle, s, solver, runs)
                out[f"{solver}_{label}"] = {"mcells_per_s": round(s.cells / t / 1e6, 4), "seconds": round(t, 3),
                                            "threads": threads, "grid": f"{s.shape[0]}^3",
                                            "iterations": rep["iterations"],
                                            "relative_residual": rep["relative_residual"]}
    finally:
        oracle.set_threads(all_threads)
    head = out["jacobi_all_cores"]
    return {
        "value": head["mcells_per_s"],
        "unit": "Mcells/s",
        "cores": all_threads,
        "kind": "port",
        "sample": f"dam+8 spheres generator at {n_all}^3 (seed 0; 1-core figures at {n_one}^3), one projection "
                  f"to 1e-6 relative residual, median of {runs}: Jacobi-PCG "
                  f"(sparse-atlas-composite-projection.ts:1346-1457 restated in C, FP64, OpenMP stencils) "
                  f"{head['iterations']} iterations in {head['seconds']:.2f} s; variants in 'variants'",
        "variants": out,
        "host": host_cpu(),
    }


def tall_cpu_baseline(n: int = 96):
    """The FP64 tall-cell oracle (oracle/fp_oracle.c, Jacobi-PCG over the
    expanded fine faces) on a bounded sample of the tall-cell workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from fluid_amd import scenes

    oracle.build()
    sc = scenes.tall_wave_scene(n, n, band=32, seed=0)
    L = sc.layout
    t = oracle.Tall(L.bases, sc.phi, L.fine_ny, L.band, L.h, L.max_delta)
    t0 = time.perf_counter()
    p, vel, b, rep = t.projection(sc.u, sc.v, sc.w, sc.dt, sc.rho, rel_tol=1e-6)
    dt = time.perf_counter() - t0
    return {
        "value": round(sc.samples / dt / 1e6, 4),
        "unit": "Msamples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"wave-surface tall-cell scene at {n}^2 columns x {n} rows, band 32 (seed 0), one projection, "
                  f"Jacobi-PCG to 1e-6 on the oracle's fine-face walk (scalar C): {rep['iterations']} "
                  f"iterations in {dt:.2f} s",
        "host": host_cpu(),
    }


def bench_tall(args, summary=False):
    """Restricted tall-cell workload: 512^2 columns, 512 fine rows, 32-row band."""
    import torch
    from fluid_amd import scenes
    from fluid_amd.projection import TallCellPressureProjection, default_options

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    sc = scenes.tall_wave_scene(args.tall_n, args.tall_n, band=32, seed=0)
    L = sc.layout
    log(f"tall scene {args.tall_n}^2 columns generated in {time.perf_counter() - t0:.1f} s, bases "
        f"{L.bases.min()}..{L.bases.max()}")
    phi, u, v, w = (torch.from_numpy(a).to(dev) for a in (sc.phi, sc.u, sc.v, sc.w))
    out = tuple(torch.empty_like(a) for a in (phi, u, v, w))
    g = TallCellPressureProjection.from_layout(L)
    common = dict(precision=args.precision, relative_tolerance=1e-6, pre_sweeps=args.sweeps,
                  post_sweeps=args.sweeps, graphs=0 if args.no_graphs else 1)
    opts = default_options(**common)
    steps = args.sub_steps if summary else args.steps
    for i in range(max(1, args.warmup)):
        r = g.project(phi, u, v, w, sc.dt, sc.rho, out=out, options=opts)
        log(f"tall warmup {i}: {r.report['device_ms']:.2f} ms iters={r.report['iterations']} "
            f"rel={r.report['relative_residual']:.2e}")
    torch.cuda.synchronize()
    lat, reports = [], []
    t_start = time.perf_counter()
    for i in range(steps):
        t1 = time.perf_counter()
        r = g.project(phi, u, v, w, sc.dt, sc.rho, out=out, options=opts)
        lat.append((time.perf_counter() - t1) * 1e3)
        reports.append(r.report)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    samples = sc.samples
    value = samples * steps / elapsed / 1e6
    rep = reports[-1]
    p_out = out[0].cpu().numpy()
    # per-launch time of the timed kernel: untimed projections after the timed
    # region, launched one by one (HIP events cannot split a replayed graph)
    timed = default_options(timed_kernel=args.timed_kernel, **common)
    kms = kl = 0
    for _ in range(2):
        r = g.project(phi, u, v, w, sc.dt, sc.rho, out=out, options=timed)
        kms += r.report["timed_kernel_ms"]
        kl += r.report["timed_kernel_launches"]
    avg_ms = kms / max(kl, 1)
    bps = {"smooth_fine": 20.0, "residual_fine": 16.0, "cg_direction": 24.0, "cg_update": 24.0}.get(
        args.timed_kernel, 0.0)
    achieved = bps * samples / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    remesh = bench_tall_remesh(g, sc, phi, u, v, w)
    g.destroy()
    err = None
    if not args.no_error_solve and args.tall_n <= 128:
        err = tall_oracle_error(sc, p_out)
    lat_sorted = sorted(lat)
    solve = {"iterations": rep["iterations"], "outer": rep["outer_iterations"], "levels": rep["levels"],
             "relative_residual": rep["relative_residual"], "max_div_after": rep["max_div_after"],
             "max_div_before": rep["max_div_before"], "device_ms": round(rep["device_ms"], 3)}
    roof = {"bound": "hbm", "kernel": args.timed_kernel, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_bytes(("tall", args.tall_n, "fp32", args.timed_kernel, args.sweeps)),
            "avg_launch_ms": round(avg_ms, 5), "bytes_per_sample": bps,
            "algorithmic_bytes": round(bps * samples), "launches": kl,
            "measured_in": "2 untimed projections after the timed region, kernels launched one by one"}
    if summary:
        return {"workload": f"tall-cell {args.tall_n}^2 columns x {args.tall_n} rows, band 32 (wave surface)",
                "msamples_per_s": round(value, 2), "ms": round(elapsed / steps * 1e3, 3), "steps": steps,
                "samples": samples, "equivalent_fine_mcells_per_s": round(L.equivalent_cells * steps / elapsed / 1e6, 1),
                "solve": solve, "roofline": roof, "remesh": remesh}
    line = {
        "metric": "MGPCG projection Mcells/s at 1e-6 rel. residual",
        "value": round(value, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 multigrid / fp32 CG / fp64 refinement",
        "data": "synthetic (seeded wave free surface, Gaussian slot velocities)",
        "config": {"workload": f"restricted tall-cell grid {args.tall_n}^2 columns x {args.tall_n} rows, band 32, "
                               "one projection to 1e-6 (cells = packed samples)",
                   "samples": samples, "equivalent_fine_cells": L.equivalent_cells,
                   "equivalent_mcells_per_s": round(L.equivalent_cells * steps / elapsed / 1e6, 1),
                   "liquid_samples": int(rep["liquid_samples"]), "sweeps": args.sweeps},
        "latency_ms": {"p50": round(statistics.median(lat), 3),
                       "p99": round(lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))], 3)},
        "solve": dict(solve, fp64_oracle_pressure_rel_l2_error=err),
        "roofline": roof,
        "remesh": remesh,
    }
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = tall_cpu_baseline()
        except Exception as e:
            line["cpu_baseline"] = {"error": repr(e)}
    return line


def bench_tall_remesh(g, sc, phi, u, v, w, calls: int = 6):
    """fp_tall_remesh on device arrays, ping-ponging between two sets of
    buffers: the surface alternately 4 fine rows lower and 4 rows higher than
    the current one (samples clamped at -5 h stay liquid), so every call moves
    most columns.  Host wall time of each call (the call returns when its
    outputs are written); the algorithmic bytes are one read and one write
    of phi, u, v and w plus the bases."""
    import torch
    h = sc.layout.h
    cur = [a.clone() for a in (phi, u, v, w)]
    nxt = [torch.empty_like(a) for a in cur]
    ms, moved = [], []
    for i in range(calls + 1):
        shift = (4 if i % 2 == 0 else -4) * h
        p = cur[0]
        cur[0] = torch.where(p <= -5 * h * (1 - 1e-6), p, (p + shift).clamp(-5 * h, 5 * h)).contiguous()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = g.remesh(*cur, out=nxt)
        if i > 0:  # the first call allocates the staging buffers
            ms.append((time.perf_counter() - t1) * 1e3)
            moved.append(res[4])
        cur, nxt = nxt, cur
    n = sum(int(a.numel()) for a in cur)
    algo = 2 * 4 * n + 2 * 4 * sc.layout.bases.size
    med = statistics.median(ms)
    return {"calls": calls, "median_ms": round(med, 3), "min_ms": round(min(ms), 3),
            "changed_columns": moved, "columns": int(sc.layout.bases.size),
            "algorithmic_bytes": algo, "achieved_gbs": round(algo / (med * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(algo / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "note": "device arrays, out of place; host wall time per call, including the "
                    "neighbour-limit convergence checks and the copy of the new bases to the host"}


def tall_oracle_error(sc, p_gpu):
    """Relative L2 error of a tall-cell pressure against the FP64 oracle solve."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    L = sc.layout
    t = oracle.Tall(L.bases, sc.phi, L.fine_ny, L.band, L.h, L.max_delta)
    op, _, _, _ = t.projection(sc.u, sc.v, sc.w, sc.dt, sc.rho, rel_tol=1e-11)
    return float(np.linalg.norm(p_gpu.astype(np.float64) - op) / np.linalg.norm(op))


def bench_batch(args):
    """96 uneven cubic grids (edges log-uniform in [32,192]) solved to 1e-6 each:
    one batched FP16-multigrid solve against the sequential per-grid FP32 loop
    (and the per-grid FP16 loop, to separate batching from precision); the
    pressure of every 8th grid is checked against the FP64 oracle."""
    import numpy as np
    import torch
    from fluid_amd import scenes
    from fluid_amd.projection import UniformPressureProjection, default_options, project_batch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ss = scenes.batch_scenes(args.batch_count, seed=0)
    cells = sum(s.cells for s in ss)
    log(f"{len(ss)} grids, {cells / 1e6:.1f} M cells, generated in {time.perf_counter() - t0:.1f} s")
    grids = [UniformPressureProjection(s.shape, s.h) for s in ss]
    reqs = []
    for s in ss:
        phi, u, v, w = (torch.from_numpy(a).to(dev) for a in (s.phi, s.u, s.v, s.w))
        out = tuple(torch.empty_like(a) for a in (phi, u, v, w))
        reqs.append(dict(phi=phi, u=u, v=v, w=w, dt=s.dt, rho=s.rho, out=out))
    mixed = default_options(precision="fp16", relative_tolerance=1e-6, persistent=1 if args.persistent else 0)
    plain = default_options(precision="fp32", relative_tolerance=1e-6)

    def timed(fn, steps):
        torch.cuda.synchronize()
        lat = []
        t = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            r = fn()
            lat.append((time.perf_counter() - t1) * 1e3)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps, lat, r

    def batched():
        return project_batch(grids, reqs, options=mixed)

    def loop(opts):
        return [g.project(r["phi"], r["u"], r["v"], r["w"], r["dt"], r["rho"], out=r["out"], options=opts)
                for g, r in zip(grids, reqs)]

    for _ in range(max(1, args.warmup)):
        batched()
    t_b, lat_b, res_b = timed(batched, args.batch_steps)
    log(f"batched fp16: {t_b * 1e3:.1f} ms per batch")
    p_b = [res_b[i].pressure.cpu().numpy() for i in range(0, len(grids), 8)]
    loop(plain)
    t_p, lat_p, res_p = timed(lambda: loop(plain), 1)
    log(f"per-grid fp32 loop: {t_p * 1e3:.1f} ms")
    loop(mixed)
    t_m, lat_m, res_m = timed(lambda: loop(mixed), 1)
    log(f"per-grid fp16 loop: {t_m * 1e3:.1f} ms")
    worst = max(r.report["relative_residual"] for r in res_b)
    flags = sorted({r.report["flags"] for r in res_b})
    for g in grids:
        g.destroy()
    # FP64 oracle error of the batched pressures (every 8th grid)
    errs = []
    for k, i in enumerate(range(0, len(ss), 8)):
        s = ss[i]
        op, _, _, _ = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-11)
        errs.append(float(np.linalg.norm(p_b[k].astype(np.float64) - op) / max(np.linalg.norm(op), 1e-300)))
    return {
        "workload": f"{len(ss)} cubic grids, edges log-uniform in [32,192] (seed 0), one batched solve",
        "cells": cells,
        "persistent_sweeps": bool(args.persistent),
        "mcells_per_s": round(cells / t_b / 1e6, 2),
        "ms": round(t_b * 1e3, 3),
        "steps": args.batch_steps,
        "dtype": "fp16 multigrid / fp32 CG / fp64 refinement",
        "plain_fp32_per_grid_loop": {"mcells_per_s": round(cells / t_p / 1e6, 2), "ms": round(t_p * 1e3, 3)},
        "fp16_per_grid_loop": {"mcells_per_s": round(cells / t_m / 1e6, 2), "ms": round(t_m * 1e3, 3)},
        "speedup_vs_plain_fp32_loop": round(t_p / t_b, 2),
        "worst_relative_residual": worst,
        "flags": flags,
        "fp64_oracle_pressure_rel_l2_error_max": max(errs),
        "oracle_checked_grids": len(errs),
        "latency_ms": {"p50": round(statistics.median(lat_b), 3), "max": round(max(lat_b), 3)},
    }


def bench_serve(args, duration=None):
    """Served stream: Poisson arrivals (rate, duration) of 20-step sessions of
    64^3 / 128^3 / 192^3 uniform grids and 256^2-column tall-cell grids (32-row
    band) with probabilities 0.4 / 0.3 / 0.2 / 0.1, admitted against a 200 GB HBM
    budget, warm-started steps, p50/p99 per-solve latency."""
    from fluid_amd.serve import ServeLoop
    duration = args.duration if duration is None else duration
    loop = ServeLoop(rate_hz=args.rate, duration_s=duration, steps=20, seed=0)
    stats = loop.run()
    stats["workload"] = (f"Poisson sessions at {args.rate}/s for {duration} s, 20 warm-started steps each, sizes "
                         "64^3:0.4 128^3:0.3 192^3:0.2 tall-cell 256^2 x 32-band:0.1, 200 GB HBM budget "
                         "(Mcells/s counts packed samples for tall-cell grids)")
    return stats


def bench_c0(args):
    """The CPU-runnable configuration: 64^3 corner dam break, one projection to
    1e-6, FP16 multigrid (and plain FP32), inputs resident."""
    import numpy as np
    import torch
    from fluid_amd import scenes
    from fluid_amd.projection import UniformPressureProjection, default_options
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    dev = torch.device("cuda:0")
    s = scenes.dam_break_scene(64, 0)
    phi, u, v, w = (torch.from_numpy(a).to(dev) for a in (s.phi, s.u, s.v, s.w))
    out = tuple(torch.empty_like(a) for a in (phi, u, v, w))
    g = UniformPressureProjection(s.shape, s.h)
    res = {}
    for prec in ("fp16", "fp32"):
        o = default_options(precision=prec, relative_tolerance=1e-6)
        for _ in range(3):
            g.project(phi, u, v, w, s.dt, s.rho, out=out, options=o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.sub_steps):
            r = g.project(phi, u, v, w, s.dt, s.rho, out=out, options=o)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.sub_steps * 1e3
        res[prec] = {"ms": round(ms, 4), "mcells_per_s": round(s.cells / ms / 1e3, 2),
                     "iterations": r.report["iterations"], "relative_residual": r.report["relative_residual"]}
    p = out[0].cpu().numpy()
    g.destroy()
    op, _, _, _ = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
    res["fp32"]["fp64_oracle_pressure_rel_l2_error"] = float(np.linalg.norm(p - op) / np.linalg.norm(op))
    res["workload"] = "64^3 corner dam break (CPU-runnable configuration), one projection to 1e-6"
    return res


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--timed-kernel", default="smooth_fine")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweeps", type=int, default=2, help="pre- and post-smoothing sweeps per level")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels one by one (profiling)")
    ap.add_argument("--persistent", action="store_true",
                    help="column-tile sweeps as persistent launches (a fixed grid walking each stage's table)")
    ap.add_argument("--no-error-solve", action="store_true", help="skip the FP64 oracle solve")
    ap.add_argument("--sweep-pairs", action="store_true",
                    help="finest-level sweeps two to a launch (fp_tuning.sweep_pairs)")
    ap.add_argument("--headline-only", action="store_true",
                    help="uniform workload: skip the C0 / batch / served / tall-cell summaries")
    ap.add_argument("--copy-probe-mib", type=int, default=0, help="only run a device copy of this size")
    ap.add_argument("--tall-n", type=int, default=512, help="tall: columns per side (and fine rows)")
    ap.add_argument("--workload", default="uniform", choices=["uniform", "batch", "serve", "tall"],
                    help="uniform: one 384^3 grid (plus summaries of the others); batch: 96 uneven grids in "
                         "one batched solve; serve: Poisson-arriving 20-step sessions through fp_submit/fp_poll")
    ap.add_argument("--batch-count", type=int, default=96)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--sub-steps", type=int, default=5, help="timed steps of the summary workloads")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: session arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: arrival window in seconds")
    ap.add_argument("--serve-summary-duration", type=float, default=15.0,
                    help="arrival window of the served stream in the default run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this build measures one GPU only; "
                                   "independent grids shard across GPUs by request routing, "
                                   "a single grid by z-slabs (see the design notes)"}))
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib, 5)
        return
    if args.workload == "batch":
        b = bench_batch(args)
        line = {"metric": "batched MGPCG projection Mcells/s at 1e-6 rel. residual (96 uneven grids)",
                "value": b["mcells_per_s"], "unit": "Mcells/s", "n_gpus": 1, "steps": args.batch_steps,
                "warmup": args.warmup, "ms_per_step": b["ms"], "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": b["dtype"], "data": "synthetic (seeded uneven dam boxes, uniform "
                "face noise)", "config": {"workload": b["workload"], "cells": b["cells"]}, "batch": b}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "serve":
        st = bench_serve(args)
        line = {"metric": "served projections: p50/p99 solve latency (ms) and Mcells/s at 1e-6",
                "value": st["latency_ms"]["p99"], "unit": "ms (p99)", "higher_is_better": False, "n_gpus": 1,
                "dtype": "fp16 multigrid / fp32 CG / fp64 refinement",
                "data": "synthetic (seeded dam boxes per session, uniform face noise, gravity steps)",
                "config": {"workload": st["workload"], "hbm_budget_gb": 200}, "serve": st}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "tall":
        print(json.dumps(bench_tall(args)), flush=True)
        return
    line = bench_uniform(args)
    if not args.headline_only:
        for key, fn in (("c0_64", lambda: bench_c0(args)), ("batch", lambda: bench_batch(args)),
                        ("served", lambda: bench_serve(args, args.serve_summary_duration)),
                        ("tall", lambda: bench_tall(args, summary=True))):
            t0 = time.perf_counter()
            try:
                line[key] = fn()
            except Exception as e:  # a summary never hides the headline
                line[key] = {"error": repr(e)}
            log(f"{key} summary done in {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the baseline is reported, never required
            line["cpu_baseline"] = {"error": repr(e)}
        log(f"cpu baseline done in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)


def bench_extrapolation(g, s, phi, out, args, step_options, calls: int = 5):
    """fp_extrapolate_velocity of the projected velocities of the headline grid
    (device tensors, out of place), after the timed projections: host wall
    time per call (the call returns when its outputs are written).  The
    algorithmic bytes are one read of phi, u, v, w and one write of u, v, w
    (28 B per cell).  Then fp_advect_velocity of the same velocities (both
    schemes), and whole velocity steps of the grid: semi-Lagrangian advection
    through the field's own extrapolation with gravity, then the projection
    with `step_options`.  CPU baseline: the oracle's scalar C restatement on
    the same generator at 128^3."""
    import torch
    src = [a.clone() for a in out[1:]]
    res = [torch.empty_like(a) for a in src]
    rep = g.extrapolate(*src, phi=phi, out=res)[3]  # allocates the extrapolation state
    ms = []
    for _ in range(calls):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rep = g.extrapolate(*src, phi=phi, out=res)[3]
        ms.append((time.perf_counter() - t1) * 1e3)
    med = statistics.median(ms)
    cells = s.cells
    algo = 28 * cells
    advection = {}
    for scheme in ("semi-lagrangian", "maccormack"):
        adv = [torch.empty_like(a) for a in src]
        g.advect(*src, s.dt, phi=phi, transport=res, scheme=scheme, out=adv)  # allocates on first use
        t_adv = []
        for _ in range(3):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            g.advect(*src, s.dt, phi=phi, transport=res, scheme=scheme, out=adv)
            t_adv.append((time.perf_counter() - t1) * 1e3)
        advection[scheme] = {"median_ms": round(statistics.median(t_adv), 3),
                             "mcells_per_s": round(cells / statistics.median(t_adv) / 1e3, 1)}
    # whole velocity steps: advect (extrapolating the current field), project
    state = [a.clone() for a in src]
    adv = [torch.empty_like(a) for a in src]
    p_out = torch.empty_like(phi)
    t_step, iters = [], []
    for i in range(4):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.advect(*state, s.dt, phi=phi, out=adv)
        r = g.project(phi, *adv, s.dt, s.rho, out=(p_out, *state), options=step_options)
        if i:  # the first step is a warmup
            t_step.append((time.perf_counter() - t1) * 1e3)
            iters.append(r.report["iterations"])
    advection["velocity_step"] = {
        "median_ms": round(statistics.median(t_step), 3),
        "mcells_per_s": round(cells / statistics.median(t_step) / 1e3, 1), "steps": len(t_step),
        "iterations": iters, "note": "semi-Lagrangian advection through the field's own extrapolation with gravity, "
        "then the projection (headline options), consecutive steps of one state"}
    line = {"calls": calls, "median_ms": round(med, 3), "min_ms": round(min(ms), 3),
            "mcells_per_s": round(cells / med / 1e3, 1), "passes": rep["passes"],
            "active_left": rep["active_left"], "levels": rep["levels"], "band_faces": rep["band_faces"],
            "unknown_faces": rep["unknown_faces"], "algorithmic_bytes": algo,
            "achieved_gbs": round(algo / (med * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(algo / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "advection": dict(advection, note="fp_advect_velocity of the same projected velocities through "
                              "their extrapolation, dt of the scene, gravity, device tensors out of place")}
    if not args.no_cpu_baseline:
        from fluid_amd import scenes
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        c = scenes.dam_spheres_scene(128, seed=0)
        t1 = time.perf_counter()
        oracle.extrapolate(c.u, c.v, c.w, c.h, phi=c.phi)
        dt = time.perf_counter() - t1
        line["cpu_baseline"] = {"value": round(c.cells / dt / 1e6, 2), "unit": "Mcells/s", "cores": 1,
                                "kind": "port", "sample": f"oracle/fp_oracle_extrap.c (scalar C) on the dam + 8 "
                                f"spheres generator at 128^3, one extrapolation in {dt:.2f} s"}
        tr = oracle.extrapolate(c.u, c.v, c.w, c.h, phi=c.phi)[:3]
        t1 = time.perf_counter()
        oracle.advect_velocity(c.u, c.v, c.w, c.h, c.dt, phi=c.phi, transport=tr, gravity=-9.80665)
        dt = time.perf_counter() - t1
        line["advection"]["cpu_baseline"] = {
            "value": round(c.cells / dt / 1e6, 2), "unit": "Mcells/s", "cores": 1, "kind": "port",
            "sample": f"oracle/fp_oracle_advect.c (scalar C), semi-Lagrangian through the extrapolated field, dam + 8 "
                      f"spheres generator at 128^3, one call in {dt:.2f} s"}
    return line


def bench_uniform(args):
    """The headline: one projection of the 384^3 dam + 8 spheres grid per
    step, FP16 multigrid, each CG + V-cycle iteration pair replayed as a
    hipGraph; then the plain FP32 MGPCG arm on the same grid, the smoother's
    per-launch time and the FP64 oracle's error, all outside the timed region."""
    import numpy as np
    import torch
    from fluid_amd import scenes
    from fluid_amd.projection import UniformPressureProjection, default_options

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    s = scenes.dam_spheres_scene(args.n, seed=0)
    log(f"scene {args.n}^3 generated in {time.perf_counter() - t0:.1f} s, liquid {(s.phi < 0).mean():.3f}")
    phi, u, v, w = (torch.from_numpy(a).to(dev) for a in (s.phi, s.u, s.v, s.w))
    out = tuple(torch.empty_like(a) for a in (phi, u, v, w))
    g = UniformPressureProjection(s.shape, s.h)
    pairs = args.sweep_pairs and args.precision != "fp32"
    if pairs:
        g.set_tuning(pairs=1)
    common = dict(precision=args.precision, relative_tolerance=1e-6, pre_sweeps=args.sweeps,
                  post_sweeps=args.sweeps, graphs=0 if args.no_graphs else 1,
                  persistent=1 if args.persistent else 0)
    opts = default_options(**common)
    for i in range(args.warmup):
        r = g.project(phi, u, v, w, s.dt, s.rho, out=out, options=opts)
        log(f"warmup {i}: {r.report['device_ms']:.2f} ms iters={r.report['iterations']} "
            f"rel={r.report['relative_residual']:.2e}")
    torch.cuda.synchronize()
    lat = []
    reports = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        r = g.project(phi, u, v, w, s.dt, s.rho, out=out, options=opts)
        lat.append((time.perf_counter() - t1) * 1e3)
        reports.append(r.report)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    cells = s.cells
    value = cells * args.steps / elapsed / 1e6
    rep = reports[-1]
    host_out = [a.cpu().numpy() for a in out]
    log(f"timed: {ms_per_step:.3f} ms per projection")
    # the smoother's per-launch time: untimed projections after the timed
    # region with kernels launched one by one (HIP events on the library's
    # stream cannot split a replayed graph)
    timed = default_options(timed_kernel=args.timed_kernel, **dict(common, graphs=0))
    kms = kl = 0
    for _ in range(2):
        r = g.project(phi, u, v, w, s.dt, s.rho, out=out, options=timed)
        kms += r.report["timed_kernel_ms"]
        kl += r.report["timed_kernel_launches"]
    avg_ms = kms / max(kl, 1)
    # the plain FP32 MGPCG arm on the same grid (graphs, same tolerance)
    plain = {}
    if args.precision != "fp32":
        o32 = default_options(**dict(common, precision="fp32"))
        g.project(phi, u, v, w, s.dt, s.rho, out=out, options=o32)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.sub_steps):
            r32 = g.project(phi, u, v, w, s.dt, s.rho, out=out, options=o32)
        torch.cuda.synchronize()
        ms32 = (time.perf_counter() - t1) / args.sub_steps * 1e3
        plain = {"ms": round(ms32, 3), "mcells_per_s": round(cells / ms32 / 1e3, 2), "steps": args.sub_steps,
                 "iterations": r32.report["iterations"], "relative_residual": r32.report["relative_residual"],
                 "p_fp32": out[0].cpu().numpy()}
    extrapolation = bench_extrapolation(g, s, phi, out, args, opts)
    g.destroy()
    bpc = kernel_bytes_per_cell(args.timed_kernel, args.precision, args.sweeps,
                                pairs and args.timed_kernel == "smooth_fine")
    # the timed finest-level passes touch only the cells of the liquid hulls
    active = hull_cells(s.phi)
    alg_bytes = bpc * active
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # error against the FP64 oracle (MG-PCG to 1e-10 on the same inputs)
    err = None
    if not args.no_error_solve:
        t1 = time.perf_counter()
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        op, (ou, ov, ow), _, orep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-10)
        rl = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))
        err = {"pressure_rel_l2": rl(host_out[0], op), "u_rel_l2": rl(host_out[1], ou),
               "v_rel_l2": rl(host_out[2], ov), "w_rel_l2": rl(host_out[3], ow),
               "oracle": f"oracle/fp_oracle.c FP64 MG-PCG on the same {args.n}^3 inputs to "
                         f"{orep['relative_residual']:.1e} relative residual"}
        if "p_fp32" in plain:
            plain["fp64_oracle_pressure_rel_l2_error"] = rl(plain["p_fp32"], op)
        log(f"oracle error computed in {time.perf_counter() - t1:.1f} s")
    plain.pop("p_fp32", None)
    lat_sorted = sorted(lat)
    return {
        "metric": "MGPCG projection Mcells/s at 1e-6 rel. residual",
        "value": round(value, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.precision} multigrid / fp32 CG / fp64 refinement",
        "data": "synthetic (seeded dam + 8 spheres level set, Gaussian face velocities)",
        "config": {"workload": f"uniform MAC grid {args.n}^3 dam+8 spheres, one projection to 1e-6",
                   "cells": cells, "liquid_cells": int(rep["liquid_samples"]), "precision": args.precision,
                   "sweeps": args.sweeps, "graphs": not args.no_graphs, "persistent_sweeps": bool(args.persistent),
                   "sweep_pairs": pairs},
        "latency_ms": {"p50": round(statistics.median(lat), 3),
                       "p99": round(lat_sorted[min(len(lat) - 1, int(0.99 * len(lat)))], 3)},
        "solve": {"iterations": rep["iterations"], "outer": rep["outer_iterations"],
                  "relative_residual": rep["relative_residual"], "max_div_after": rep["max_div_after"],
                  "max_div_before": rep["max_div_before"], "device_ms": round(rep["device_ms"], 3)},
        "fp64_oracle_error": err,
        "plain_fp32_mgpcg": plain or None,
        "extrapolation": {k: v for k, v in extrapolation.items() if k != "advection"},
        "advection": extrapolation.get("advection"),
        "roofline": {"bound": "hbm", "kernel": args.timed_kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_bytes(("uniform", args.n, args.precision, args.timed_kernel, args.sweeps, pairs)),
                     "avg_launch_ms": round(avg_ms, 5), "bytes_per_cell": bpc, "hull_cells": active,
                     "algorithmic_bytes": round(alg_bytes), "launches": kl,
                     "measured_in": "2 untimed projections after the timed region, kernels launched one by one"},
    }


if __name__ == "__main__":
    main()
