"""The CM11a separating-wall solver (fp_options.solver = FP_SOLVER_CM11A) on
the GPU through the C ABI, against its CPU restatement
(oracle/fp_oracle_cm11a.c, pinned in tests/test_oracle_cm11a.py).

Bar: bit-exact.  Both run the reference's FP32 operation order without
contraction (D4-ordered sums, a double-single coarsest solve, correctly
rounded division), red-black passes are order independent, so the pressure,
the velocities and the projected residual agree bit for bit; at 384^3 (no
oracle run) the properties the LCP fixes: p >= 0 on the separating walls and
a small projected residual."""
import numpy as np
import pytest

from fluid_amd import scenes

from test_oracle_cm11a import hydrostatic, separating

pytestmark = pytest.mark.gpu

NONCONVERGENCE = 0x8


def noisy_dam(n, seed):
    s = scenes.dam_break_scene(n, seed)
    return s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho


CASES = {
    "hydrostatic16": lambda: hydrostatic(16),
    "hydrostatic12x16x8": lambda: hydrostatic((12, 16, 8)),
    "separating12": lambda: separating(12),
    "separating24x20x16": lambda: separating((24, 20, 16)),
    "dam32": lambda: noisy_dam(32, 1),
    "dam64": lambda: noisy_dam(64, 0),
    # deep hierarchy, rough velocities: the coarsest solves reach their cap
    # (reported as FP_FLAG_NONCONVERGENCE), deterministically on both sides
    "spheres96": lambda: (lambda s: (s.phi, s.u, s.v, s.w, s.h, s.dt, s.rho))(scenes.dam_spheres_scene(96, seed=0)),
}


def gpu_cm11a(fpl, phi, u, v, w, h, dt, rho, q=None, **opts):
    g = fpl.UniformPressureProjection(phi.shape, h)
    try:
        return g.project(phi, u, v, w, dt, rho, target_divergence=q, solver="cm11a", **opts)
    finally:
        g.destroy()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("open_top", [0, 1])
def test_bit_exact_against_the_oracle(fpl, oracle, name, open_top):
    phi, u, v, w, h, dt, rho = CASES[name]()
    P, (ou, ov, ow), orep = oracle.cm11a_project(phi, u, v, w, h, dt, rho, open_top=open_top)
    r = gpu_cm11a(fpl, phi, u, v, w, h, dt, rho, open_top=open_top)
    rep = r.report
    print(name, open_top, rep, orep)
    liq = phi < 0
    op = np.where(liq, P[1:-1, 1:-1, 1:-1], 0)
    assert np.array_equal(r.pressure, op)
    assert np.array_equal(r.u, ou) and np.array_equal(r.v, ov) and np.array_equal(r.w, ow)
    assert np.float32(rep["lcp_residual"]) == np.float32(orep["lcp_residual"])
    assert rep["coarse_sweeps"] == orep["coarse_sweeps"]
    assert rep["coarse_failures"] == orep["coarse_failed"]
    assert rep["flags"] == (NONCONVERGENCE if orep["coarse_failed"] else 0) and rep["levels"] == orep["levels"]
    assert (orep["coarse_failed"] > 0) == (name == "spheres96")
    assert rep["iterations"] == 7 and rep["liquid_samples"] == int(liq.sum())


def test_schedule_options_and_target_divergence(fpl, oracle):
    phi, u, v, w, h, dt, rho = separating(12)
    q = np.where(phi < 0, 0.25, 0).astype(np.float32)
    P, (ou, ov, ow), orep = oracle.cm11a_project(phi, u, v, w, h, dt, rho, q=q, full_cycles=2, v_cycles=9, pre=3,
                                                 post=3)
    r = gpu_cm11a(fpl, phi, u, v, w, h, dt, rho, q=q, full_cycles=2, v_cycles=9, lcp_sweeps=3)
    assert np.array_equal(r.pressure, np.where(phi < 0, P[1:-1, 1:-1, 1:-1], 0))
    assert np.array_equal(r.u, ou) and np.array_equal(r.v, ov) and np.array_equal(r.w, ow)
    assert r.report["iterations"] == 11


def test_device_tensors_in_place(fpl, oracle):
    torch = pytest.importorskip("torch")
    phi, u, v, w, h, dt, rho = separating(12)
    P, (ou, ov, ow), _ = oracle.cm11a_project(phi, u, v, w, h, dt, rho)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(a.copy()).to(dev) for a in (phi, u, v, w)]
    p = torch.zeros_like(t[0])
    g = fpl.UniformPressureProjection(phi.shape, h)
    try:
        g.project(*t, dt, rho, out=(p, t[1], t[2], t[3]), solver="cm11a")
        torch.cuda.synchronize()
        assert np.array_equal(t[1].cpu().numpy(), ou) and np.array_equal(t[3].cpu().numpy(), ow)
        assert np.array_equal(p.cpu().numpy(), np.where(phi < 0, P[1:-1, 1:-1, 1:-1], 0))
    finally:
        g.destroy()


def test_unsupported_forms_are_refused(fpl):
    phi, u, v, w, h, dt, rho = hydrostatic((9, 14, 11))  # no exact-halving hierarchy
    g = fpl.UniformPressureProjection(phi.shape, h)
    try:
        with pytest.raises(fpl.FluidProjectionError) as e:
            g.project(phi, u, v, w, dt, rho, solver="cm11a")
        assert e.value.code == fpl.ERR_LAYOUT
        with pytest.raises(fpl.FluidProjectionError) as e:
            g.project(phi, u, v, w, dt, rho, open_top=1)  # an open top needs the CM11a solver
        assert e.value.code == fpl.ERR_UNSUPPORTED
    finally:
        g.destroy()
    phi, u, v, w, h, dt, rho = hydrostatic(16)
    g = fpl.UniformPressureProjection(phi.shape, h)
    try:
        with pytest.raises(fpl.FluidProjectionError) as e:
            g.submit(phi, u, v, w, dt, rho, solver="cm11a")
        assert e.value.code == fpl.ERR_UNSUPPORTED
        with pytest.raises(fpl.FluidProjectionError):
            g.project(phi, u, v, w, dt, rho, solver="cm11a", lcp_sweeps=100)
        g.project(phi, u, v, w, dt, rho, solver="cm11a")
        g.project(phi, u, v, w, dt, rho)  # the MGPCG solver on the same grid afterwards
    finally:
        g.destroy()


def test_384_against_the_mgpcg_solution(fpl):
    """The 384^3 dam at rest after one gravity step (no active constraint):
    the CM11a schedule reaches the Neumann-wall solution that the MGPCG
    solver computes on the same grid to 1e-6."""
    phi, u, v, w, h, dt, rho = hydrostatic(384)
    g = fpl.UniformPressureProjection(phi.shape, h)
    try:
        a = g.project(phi, u, v, w, dt, rho)
        c = g.project(phi, u, v, w, dt, rho, solver="cm11a")
    finally:
        g.destroy()
    rep = c.report
    liq = phi < 0
    err = np.linalg.norm((c.pressure - a.pressure)[liq]) / np.linalg.norm(a.pressure[liq])
    print(rep, "rel. pressure difference", err)
    assert rep["flags"] == 0 and rep["coarse_failures"] == 0 and rep["levels"] == 8
    assert err < 1e-3
    assert (c.pressure[liq] > 0).all()
