"""Batched projections (fp_project_batch): uneven grids in one solve.

Bars: each grid's batched result is bitwise equal to its projection alone
(every grid keeps its own block decomposition, scalars and reductions); each
reaches the requested FP64 relative residual; pressures match the FP64
oracle within 1e-5 relative L2 (small grids, where the oracle is quick)."""
import numpy as np
import pytest

from fluid_amd import scenes

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def requests(ss):
    return [dict(phi=s.phi, u=s.u, v=s.v, w=s.w, dt=s.dt, rho=s.rho) for s in ss]


# uneven edges: power-of-two lockstep, semi-coarsened (48, 40), odd (ceil
# ladder: 37), a single-level grid (3) and a grid below the fused-sweep size (12)
SIZES = [32, 48, 37, 40, 12, 64, 3, 24]


@pytest.fixture(scope="module")
def batch():
    out = []
    for g, n in enumerate(SIZES):
        rng = np.random.default_rng(100 + g)
        fill = tuple(float(x) for x in rng.uniform(0.3, 0.8, 3))
        out.append(scenes.dam_break_scene(n, seed=200 + g, fill=fill))
    return out


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_batch_equals_alone(fpl, batch, precision):
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in batch]
    try:
        res_b = fpl.project_batch(grids, requests(batch), precision=precision, relative_tolerance=1e-6)
        for g, s, rb in zip(grids, batch, res_b):
            ra = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision=precision, relative_tolerance=1e-6)
            print(s.shape, "iters", ra.report["iterations"], rb.report["iterations"])
            assert rb.report["flags"] == 0, rb.report
            assert rb.report["relative_residual"] <= 1e-6
            assert rb.report["iterations"] == ra.report["iterations"]
            for a, b in ((ra.pressure, rb.pressure), (ra.u, rb.u), (ra.v, rb.v), (ra.w, rb.w)):
                assert np.array_equal(a, b)
    finally:
        for g in grids:
            g.destroy()


def test_batch_matches_oracle(fpl, oracle, batch):
    small = [s for s in batch if s.shape[0] <= 40]
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in small]
    try:
        res = fpl.project_batch(grids, requests(small), precision="fp16", relative_tolerance=1e-6)
        for s, r in zip(small, res):
            op, vel, b, rep = oracle.uniform_projection(s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho, rel_tol=1e-12)
            err = rel_l2(r.pressure, op)
            print(s.shape, "liquid", int((s.phi < 0).sum()), "pressure rel L2 vs FP64", err)
            if (s.phi < 0).any():
                assert err < 1e-5
            for a, ob in zip((r.u, r.v, r.w), vel):
                assert rel_l2(a, ob) < 1e-5
    finally:
        for g in grids:
            g.destroy()


def test_batch_repeat_and_regroup(fpl, batch):
    """The same grids batched in another order and a sub-batch give the same bits."""
    grids = [fpl.UniformPressureProjection(s.shape, s.h) for s in batch]
    try:
        first = fpl.project_batch(grids, requests(batch), precision="fp16")
        order = list(reversed(range(len(batch))))
        second = fpl.project_batch([grids[i] for i in order], requests([batch[i] for i in order]), precision="fp16")
        for k, i in enumerate(order):
            assert np.array_equal(first[i].pressure, second[k].pressure)
        sub = fpl.project_batch(grids[1:3], requests(batch[1:3]), precision="fp16")
        for k in range(2):
            assert np.array_equal(first[1 + k].pressure, sub[k].pressure)
    finally:
        for g in grids:
            g.destroy()


def test_batch_rejects_mixed_contexts(fpl, batch):
    s = batch[0]
    c2 = fpl.Context(0)
    a = fpl.UniformPressureProjection(s.shape, s.h)
    b = fpl.UniformPressureProjection(s.shape, s.h, context=c2)
    try:
        with pytest.raises(fpl.FluidProjectionError):
            fpl.project_batch([a, b], requests([s, s]))
    finally:
        a.destroy()
        b.destroy()
        c2.destroy()
