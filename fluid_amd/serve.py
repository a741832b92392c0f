"""Request loop for served projections (host side of fp_submit / fp_poll).

A stream of sessions arrives by a Poisson process.  Each session owns one grid
for its lifetime -- a cubic uniform grid, or a restricted tall-cell grid
("tall<n>": n x n columns of n rows under a 32-row band) -- and runs a fixed
number of time steps; each step is one projection submitted to the context's
continuous-batching worker, warm-started from the previous step's pressure,
followed by a gravity impulse on the projected velocities (on the device) that
feeds the next step.  Grids are
admitted against an HBM budget: a session whose grid does not fit waits in the
admission queue until finished sessions free memory.  Latency is measured per
projection from submit to completion.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import projection as fpl
from . import scenes


@dataclass
class Session:
    sid: int
    n: object  # edge of a cubic grid, or "tall<n>"
    arrival: float
    steps_left: int
    scene: object = None
    grid: object = None
    tensors: tuple = ()
    ticket: object = None
    admitted: float | None = None
    latencies: list = field(default_factory=list)
    iterations: list = field(default_factory=list)
    vmask: object = None
    cells: int = 0


class ServeLoop:
    """Poisson arrivals of `steps`-step sessions drawn from `mix`
    ({edge or "tall<n>": probability}); returns latency / throughput
    statistics."""

    def __init__(self, rate_hz: float, duration_s: float, mix=None, steps: int = 20, seed: int = 0,
                 hbm_budget_bytes: int = 200 << 30, precision: str = "fp16", tolerance: float = 1e-6,
                 keep_first: bool = False):
        import torch
        self.torch = torch
        self.rate = rate_hz
        self.duration = duration_s
        self.mix = mix or {64: 0.4, 128: 0.3, 192: 0.2, "tall256": 0.1}
        self.steps = steps
        self.rng = np.random.default_rng(seed)
        self.ctx = fpl.Context(0, hbm_budget_bytes)
        self.opts_cold = fpl.default_options(precision=precision, relative_tolerance=tolerance)
        self.opts_warm = fpl.default_options(precision=precision, relative_tolerance=tolerance, warm_start=1)
        self.dev = torch.device("cuda:0")
        # keep_first: the first served step of each size class (its scene and
        # host copies of the outputs) is kept in self.first for checking
        self.keep_first = keep_first
        self.first = {}

    def _arrivals(self):
        t, out = 0.0, []
        sizes = list(self.mix)
        probs = np.array([self.mix[k] for k in sizes], float)
        probs /= probs.sum()
        sid = 0
        while True:
            t += self.rng.exponential(1.0 / self.rate)
            if t >= self.duration:
                return out
            n = sizes[int(self.rng.choice(len(sizes), p=probs))]
            out.append(Session(sid, n, t, self.steps))
            sid += 1

    def _admit(self, s: Session) -> bool:
        tall = isinstance(s.n, str)
        if s.scene is None:
            # generated once: a session waiting for admission keeps its scene
            fill = tuple(float(x) for x in np.random.default_rng(1000 + s.sid).uniform(0.2, 0.8, 3))
            if tall:
                n = int(s.n[4:])
                s.scene = scenes.tall_dam_scene(n, n, band=32, seed=2000 + s.sid, fill=fill)
            else:
                s.scene = scenes.dam_break_scene(s.n, seed=2000 + s.sid, fill=fill)
        try:
            if tall:
                s.grid = fpl.TallCellPressureProjection.from_layout(s.scene.layout, context=self.ctx)
            else:
                s.grid = fpl.UniformPressureProjection((s.n,) * 3, 1.0 / s.n, context=self.ctx)
        except fpl.FluidProjectionError as e:
            if e.code == fpl.ERR_BUDGET:
                return False  # over the HBM budget: wait until sessions finish
            raise  # any other refusal would never clear: fail the run

        torch = self.torch
        phi, u, v, w = (torch.from_numpy(a).to(self.dev) for a in (s.scene.phi, s.scene.u, s.scene.v, s.scene.w))
        out = (torch.empty_like(phi), torch.empty_like(u), torch.empty_like(v), torch.empty_like(w))
        s.tensors = (phi, u, v, w, out)
        s.cells = int(s.scene.phi.size)
        if tall:
            # gravity on every v slot that carries an interior face
            s.vmask = torch.from_numpy(s.scene.layout.slot_masks()[1].astype(np.float32)).to(self.dev)
        return True

    def _submit(self, s: Session, first: bool):
        phi, u, v, w, out = s.tensors
        s.ticket = s.grid.submit(phi, u, v, w, s.scene.dt, s.scene.rho, out=out,
                                 options=self.opts_cold if first else self.opts_warm)

    def _advance(self, s: Session, res):
        """Next step's input: the projected field plus a gravity impulse on y faces."""
        phi, u, v, w, out = s.tensors
        p, uo, vo, wo = out
        u.copy_(uo)
        v.copy_(vo)
        w.copy_(wo)
        if s.vmask is not None:
            v.sub_(s.vmask * (9.81 * s.scene.dt))
        else:
            v[:, 1:-1, :] -= 9.81 * s.scene.dt

    def run(self):
        pending = collections.deque(self._arrivals())
        waiting = collections.deque()  # arrived, not admitted
        active = []
        done = []
        cells_solved = 0
        solves = 0
        t0 = time.perf_counter()
        waits = []
        while pending or waiting or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission in arrival order while memory lasts
            while waiting:
                s = waiting[0]
                if not self._admit(s):
                    break
                waiting.popleft()
                s.admitted = time.perf_counter() - t0
                waits.append(s.admitted - s.arrival)
                self._submit(s, True)
                active.append(s)
            progressed = False
            for s in list(active):
                res = s.ticket.wait(0)
                if res is None:
                    continue
                progressed = True
                s.latencies.append(s.ticket.latency_ms)
                s.iterations.append(res.report["iterations"])
                if self.keep_first and s.steps_left == self.steps and s.n not in self.first:
                    self.first[s.n] = (s.scene, tuple(a.cpu().numpy() for a in s.tensors[4]), dict(res.report))
                if res.report["flags"] != 0 or res.report["relative_residual"] > self.opts_cold.relative_tolerance:
                    raise RuntimeError(f"session {s.sid} step failed: {res.report}")
                cells_solved += s.cells
                solves += 1
                s.steps_left -= 1
                if s.steps_left > 0:
                    self._advance(s, res)
                    self._submit(s, False)
                else:
                    active.remove(s)
                    s.grid.destroy()
                    s.tensors = ()
                    done.append(s)
            if not progressed:
                if active:
                    active[0].ticket.wait(1)
                elif pending and not waiting:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        lat = np.array([x for s in done for x in s.latencies], float)
        its = np.array([x for s in done for x in s.iterations], float)
        self.ctx.destroy()
        by_size = {}
        for n in sorted({s.n for s in done}, key=str):
            ls = np.array([x for s in done if s.n == n for x in s.latencies])
            its_n = np.array([x for s in done if s.n == n for x in s.iterations], float)
            label = f"{n}^3" if not isinstance(n, str) else f"{n} columns x 32 band"
            by_size[label] = {"solves": int(ls.size), "p50_ms": round(float(np.percentile(ls, 50)), 3),
                              "p99_ms": round(float(np.percentile(ls, 99)), 3),
                              "iterations_mean": round(float(its_n.mean()), 2)}
        return {
            "sessions": len(done),
            "solves": solves,
            "wall_s": round(wall, 3),
            "mcells_per_s": round(cells_solved / wall / 1e6, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                           "p99": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
                           "mean": round(float(lat.mean()), 3) if lat.size else None},
            "admission_wait_ms": {"p50": round(float(np.percentile(waits, 50)) * 1e3, 3) if waits else None,
                                  "max": round(max(waits) * 1e3, 3) if waits else None},
            "iterations_mean": round(float(its.mean()), 2) if its.size else None,
            "by_size": by_size,
        }
