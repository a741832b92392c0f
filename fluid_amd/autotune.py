"""Autotuner of the uniform projection's launch shapes and sweep counts.

For a grid size class it times every candidate configuration on a seeded
scene of that size -- smoothing sweeps per side, the finest-level sweep
tile width along z (16 or 32 lanes), the number of workgroups a sweep is
split into (which sets the x chunk each workgroup marches) and whether the
finest level's sweeps run two to a launch (FP16/BF16) -- keeps the
fastest one that reaches the tolerance with no flags, and persists the
choice per size class to a JSON file.  ``apply`` reads the file back and
configures a grid (fp_grid_set_tuning) and the options of its projections.

Size classes are keyed by the edge of the cube with the same cell count,
rounded to the nearest class edge below.
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np

from . import projection as fpl

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_FILE = os.path.join(HERE, "tuning", "mi355x.json")
CLASS_EDGES = (32, 48, 64, 96, 128, 192, 256, 384, 512)

SWEEPS = (1, 2, 3)
TILES = (32, 16)
GROUPS = (4096, 8192, 16384)
PAIRS = (0, 1)


def size_class(dims) -> int:
    """The class edge of a grid: the largest class edge not above the edge of
    the cube with the same number of cells (the smallest edge below that)."""
    edge = float(np.prod([int(d) for d in dims])) ** (1.0 / 3.0)
    best = CLASS_EDGES[0]
    for e in CLASS_EDGES:
        if e <= edge * 1.0001:
            best = e
    return best


def candidates(dims, precision="fp16"):
    """Every (sweeps, tile_z, groups, pairs) the grid can run: 32-wide tiles
    need nz >= 32; sweep pairs need an FP16/BF16 multigrid."""
    out = []
    for sw, tz, gr, pa in itertools.product(SWEEPS, TILES, GROUPS, PAIRS):
        if tz == 32 and int(dims[2]) < 32:
            continue
        if pa and precision == "fp32":
            continue
        out.append({"sweeps": sw, "tile_z": tz, "groups": gr, "pairs": pa})
    return out


def time_config(grid, tensors, scene, cfg, precision="fp16", tol=1e-6, reps=3):
    """Median wall time (ms) of `reps` projections with a configuration, and
    the last report."""
    import torch
    grid.set_tuning(tile_z=cfg["tile_z"], groups=cfg["groups"], pairs=cfg.get("pairs", 0))
    opts = fpl.default_options(precision=precision, relative_tolerance=tol, pre_sweeps=cfg["sweeps"],
                               post_sweeps=cfg["sweeps"])
    phi, u, v, w, out = tensors
    grid.project(phi, u, v, w, scene.dt, scene.rho, out=out, options=opts)  # warm-up (plans, graphs)
    times = []
    rep = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = grid.project(phi, u, v, w, scene.dt, scene.rho, out=out, options=opts)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        rep = r.report
    return float(np.median(times)), rep


def tune(dims, scene=None, precision="fp16", tol=1e-6, reps=3, log=None):
    """Times every candidate on one grid of `dims` and returns the record of
    the fastest configuration that converges with no flags."""
    import torch
    from . import scenes
    if scene is None:
        n = int(dims[0])
        if tuple(int(d) for d in dims) != (n, n, n):
            raise ValueError("the default tuning scene is cubic; pass a scene for other shapes")
        scene = scenes.dam_spheres_scene(n, seed=0)
    dev = torch.device("cuda:0")
    phi, u, v, w = (torch.from_numpy(a).to(dev) for a in (scene.phi, scene.u, scene.v, scene.w))
    out = tuple(torch.empty_like(a) for a in (phi, u, v, w))
    g = fpl.UniformPressureProjection(scene.shape, scene.h)
    trials = []
    try:
        for cfg in candidates(scene.shape, precision):
            ms, rep = time_config(g, (phi, u, v, w, out), scene, cfg, precision, tol, reps)
            ok = rep["flags"] == 0 and rep["relative_residual"] <= tol
            trials.append(dict(cfg, ms=round(ms, 4), iterations=rep["iterations"], ok=ok))
            if log:
                log(f"{tuple(scene.shape)} {cfg} {ms:.3f} ms iterations {rep['iterations']} ok {ok}")
    finally:
        g.destroy()
    good = [t for t in trials if t["ok"]]
    if not good:
        raise RuntimeError(f"no configuration converged for {tuple(scene.shape)}")
    best = min(good, key=lambda t: t["ms"])
    default = next((t for t in trials if t["sweeps"] == 2 and t["tile_z"] == (32 if scene.shape[2] >= 32 else 16)
                    and t["groups"] == 8192 and not t.get("pairs", 0)), None)
    return {"dims": [int(d) for d in scene.shape], "precision": precision, "tolerance": tol,
            "best": {k: best.get(k, 0) for k in ("sweeps", "tile_z", "groups", "pairs")}, "best_ms": best["ms"],
            "default_ms": default["ms"] if default else None, "trials": trials}


def save(records, path=DEFAULT_FILE):
    """Writes {size class: record} (merged with what the file holds)."""
    data = load(path)
    for r in records:
        data[str(size_class(r["dims"]))] = r
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    return path


def load(path=DEFAULT_FILE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def choice_for(dims, path=DEFAULT_FILE):
    """The tuned (sweeps, tile_z, groups, pairs) of a grid's size class, or
    None."""
    rec = load(path).get(str(size_class(dims)))
    if not rec:
        return None
    best = dict(rec["best"])
    best.setdefault("pairs", 0)
    if best["tile_z"] == 32 and int(dims[2]) < 32:
        best["tile_z"] = 16
    return best


def apply(grid, path=DEFAULT_FILE, **opts):
    """Configures a grid with its size class's tuned launch shapes and returns
    options with the tuned sweep counts (defaults when the class is untuned)."""
    best = choice_for(grid.dims, path)
    if best is None:
        return fpl.default_options(**opts), None
    grid.set_tuning(tile_z=best["tile_z"], groups=best["groups"], pairs=best["pairs"])
    return fpl.default_options(pre_sweeps=best["sweeps"], post_sweeps=best["sweeps"], **opts), best


def main(argv=None):
    """python -m fluid_amd.autotune [--out FILE] EDGE...: tunes cubic grids of
    the given edges on the dam + 8 spheres scene and writes the choices."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("edges", type=int, nargs="+")
    ap.add_argument("--out", default=DEFAULT_FILE)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    fpl.load_library()
    recs = []
    for n in args.edges:
        rec = tune((n, n, n), reps=args.reps, log=lambda m: print("#", m, file=sys.stderr, flush=True))
        print(json.dumps({k: rec[k] for k in ("dims", "best", "best_ms", "default_ms")}), flush=True)
        recs.append(rec)
    print(save(recs, args.out), flush=True)


if __name__ == "__main__":
    main()
