"""Serving (fp_submit / fp_poll): continuous batching of uneven projections.

Bars: every served result is bitwise equal to the same projection run alone
through fp_project (the same per-grid arithmetic, whatever it shares rounds
with); requests submitted while others are running complete; warm-started
session steps reach the tolerance; a grid with running work refuses
destruction."""
import time

import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu

SIZES = [48, 32, 64, 37, 24, 40]


def scene(g, n):
    rng = np.random.default_rng(300 + g)
    fill = tuple(float(x) for x in rng.uniform(0.3, 0.8, 3))
    return scenes.dam_break_scene(n, seed=400 + g, fill=fill)


def test_served_equals_alone_with_staggered_arrivals(fpl):
    ss = [scene(g, n) for g, n in enumerate(SIZES)]
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in ss]
    try:
        alone = [g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16") for g, s in zip(grids, ss)]
        tickets = []
        for i, (g, s) in enumerate(zip(grids, ss)):
            tickets.append(g.submit(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16"))
            if i == 2:
                time.sleep(0.01)  # later requests arrive while the first ones run
        for t, a, s in zip(tickets, alone, ss):
            r = t.wait(60000)
            assert r is not None, "served projection did not complete"
            print(s.shape, "latency ms", round(t.latency_ms, 2), "iters", r.report["iterations"])
            assert r.report["flags"] == 0 and r.report["relative_residual"] <= 1e-6
            assert r.report["iterations"] == a.report["iterations"]
            for x, y in ((a.pressure, r.pressure), (a.u, r.u), (a.v, r.v), (a.w, r.w)):
                assert np.array_equal(x, y)
    finally:
        for g in grids:
            g.destroy()


def test_session_warm_start_steps(fpl):
    s = scene(7, 40)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    try:
        u, v, w = s.u, s.v, s.w
        iters = []
        for step in range(4):
            t = g.submit(s.phi, u, v, w, s.dt, s.rho, precision="fp16", warm_start=1 if step else 0)
            r = t.wait(60000)
            assert r is not None and r.report["flags"] == 0
            assert r.report["relative_residual"] <= 1e-6
            iters.append(r.report["iterations"])
            # next step: the projected field plus a fresh gravity impulse
            u, v, w = r.u, r.v.copy(), r.w
            v[:, 1:-1, :] -= 9.81 * s.dt
        print("iterations per step", iters)
        assert iters[-1] <= iters[0]
    finally:
        g.destroy()


def test_destroy_refused_while_running(fpl):
    s = scene(9, 64)
    g = fpl.UniformPressureProjection(s.shape, s.h)
    t = g.submit(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp32", relative_tolerance=1e-10)
    try:
        try:
            g.destroy()
            refused = False
        except fpl.FluidProjectionError:
            refused = True
        r = t.wait(60000)
        assert r is not None
        if not refused:
            pytest.skip("the projection finished before destroy was attempted")
    finally:
        g.destroy()


def test_tall_and_uniform_served_together_equal_alone(fpl):
    """Tall-cell and uniform requests interleaved on one worker: each result
    bitwise equal to its projection alone."""
    tall = [scenes.tall_dam_scene(32, 48, band=16, seed=501, fill=(0.6, 0.5, 0.4)),
            scenes.tall_wave_scene(24, 64, band=16, seed=502)]
    uni = [scene(10, 32), scene(11, 40)]
    grids = [fpl.TallCellPressureProjection.from_layout(t.layout) for t in tall] + \
            [fpl.UniformPressureProjection(s.shape, s.h) for s in uni]
    reqs = tall + uni
    try:
        alone = [g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16") for g, s in zip(grids, reqs)]
        order = [0, 2, 1, 3]
        tickets = {}
        for k, i in enumerate(order):
            g, s = grids[i], reqs[i]
            tickets[i] = g.submit(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16")
            if k == 1:
                time.sleep(0.01)
        for i, a in enumerate(alone):
            r = tickets[i].wait(60000)
            assert r is not None
            assert r.report["flags"] == 0 and r.report["relative_residual"] <= 1e-6
            assert r.report["iterations"] == a.report["iterations"]
            for x, y in ((a.pressure, r.pressure), (a.u, r.u), (a.v, r.v), (a.w, r.w)):
                assert np.array_equal(x, y)
    finally:
        for g in grids:
            g.destroy()


def test_served_stream_admission_and_oracle(fpl, oracle):
    """The served-stream loop (Poisson sessions of warm-started steps, mixed
    uniform and tall-cell sizes) under an HBM budget that holds a 64^3 grid
    and a 32^3 one at most: arrivals wait for admission, every step reaches the
    tolerance, and the first result of each size class matches the FP64
    oracle (pressure and velocities within 1e-5 relative L2)."""
    from fluid_amd.serve import ServeLoop

    mix = {32: 0.4, 48: 0.3, 64: 0.2, "tall64": 0.1}
    # a 64^3 session leaves room for one 32^3 grid beside it: others wait
    budget = fpl.grid_bytes((64, 64, 64)) + fpl.grid_bytes((32, 32, 32))
    assert fpl.tall_grid_bytes(64, 64, 64, 32) < budget
    loop = ServeLoop(rate_hz=40.0, duration_s=1.5, mix=mix, steps=4, seed=3, hbm_budget_bytes=budget,
                     keep_first=True)
    stats = loop.run()
    print(stats)
    assert stats["sessions"] >= 20
    assert stats["solves"] == 4 * stats["sessions"]
    assert stats["admission_wait_ms"]["max"] > 0.0
    assert set(loop.first) == set(mix)
    for n, (sc, outs, rep) in loop.first.items():
        p, u, v, w = outs
        if isinstance(n, str):
            L = sc.layout
            t = oracle.Tall(L.bases, sc.phi, L.fine_ny, L.band, L.h, L.max_delta)
            op, (ou, ov, ow), _, _ = t.projection(sc.u, sc.v, sc.w, sc.dt, sc.rho, rel_tol=1e-11)
        else:
            op, (ou, ov, ow), _, _ = oracle.uniform_projection(sc.phi, sc.u, sc.v, sc.w, sc.h, sc.dt, sc.rho,
                                                               rel_tol=1e-12)
        for a, b in ((p, op), (u, ou), (v, ov), (w, ow)):
            a = np.asarray(a, np.float64)
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
            assert err < 1e-5, (n, err)
