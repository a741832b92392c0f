"""The N-API addon: the TypeScript host's boundary onto the C ABI."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(code):
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, cwd=ROOT, timeout=120)


def plugin_check(mode):
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "plugin_check.js"), mode], capture_output=True,
                       text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_method_satisfies_the_simulation_method_contract(fpl):
    """Every required member of SimulationMethod (lib/core/method-contract.ts:
    363-458): identity and copy, backend, a ResourcePluginDefinition
    (lib/core/resource-plugin.ts:33-41), quality labels for each GPUQuality,
    a parameter schema of NumberParamSpec / SelectParamSpec entries, the
    pressure mapping, presets that name declared parameters; the addon binds
    every C entry point the host needs."""
    out = plugin_check("contract")
    for k in ("id", "label", "shortLabel", "badge", "description", "detail", "backend", "pressureMapping"):
        assert out[k] == "string", k
    assert out["backend_value"] == "webgpu"
    res = out["resource"]
    assert res["lane"] in ("platform", "fluid", "svo", "optional") and res["blocks"] in ("viewport", "transport",
                                                                                        "nothing")
    assert set(res["provides"]) <= {"renderer", "live-scene", "fluid-authority", "water-presentation",
                                    "sparse-voxel-presentation", "optional-tooling"} and res["id"] and res["label"]
    assert out["qualityLabels"] == ["balanced", "high", "ultra"]
    keys = {p["key"] for p in out["params"]}
    assert all(p["ok"] and p["kind"] in ("number", "select") and p["tier"] in ("coarse", "fine")
               for p in out["params"])
    for q, values in out["presets"].items():
        assert set(values) <= keys, q
    assert out["normalized"] == {"precision": "fp16", "smoothingSweeps": 4, "layout": "uniform",
                                 "pressureSolver": "mgpcg", "velocityTransport": "semi-lagrangian"}
    assert out["createSolverAsync"] == "function"
    assert out["binding"] == sorted(["createContext", "destroyContext", "createGrid", "destroyGrid", "gridLevels",
                                     "project", "submit", "projectBatch", "bytesInUse", "lastError", "abiVersion",
                                     "remeshTallGrid", "extrapolateVelocity", "advectVelocity"])
    # the reference's default lattice sizing: 1.2 x 0.8 x 0.8 m at 2.5 cm
    assert out["lattice"]["nx"] == 48 and out["lattice"]["ny"] == 32 and out["lattice"]["nz"] == 32
    assert out["openTop"] == "RangeError"


def test_addon_abi_version(fpl):
    r = run_node("const m=require('./fluid_amd/node');console.log(m.binding.abiVersion())")
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "3"


@pytest.mark.gpu
def test_plugin_solver_runs_scene_steps_and_entry_points(fpl):
    """createSolverAsync on a SceneDescription (uniform and tall-cell layouts):
    advanceTo steps the fixed step, readStats resolves a GPUEulerianInfo with
    every required member (lib/core/webgpu-eulerian.ts:13-38), destroy is
    idempotent; submit and projectBatch through the addon equal project bit
    for bit; destroy with a projection in flight waits for it."""
    out = plugin_check("run")
    for layout in ("uniform", "tall-cell"):
        o = out[layout]
        assert o["stepped"] == [True, True, False], layout
        info = o["info"]
        for k in ("nx", "ny", "nz", "storedNy", "cellCount", "equivalentUniformCells", "compressionRatio",
                  "regularLayers", "maximumNeighborDelta", "gridKind", "pressureIterations", "allocatedBytes",
                  "quality"):
            assert k in info, (layout, k)
        assert info["gridKind"] == "uniform" and info["quality"] == "high"
        assert info["steps"] == 3 and info["pressureSolveConverged"] is True, (layout, info)
        assert info["pressureFlags"] == 0 and info["pressureRelativeResidual"] <= 1e-6, (layout, info)
        assert info["maxDivergenceAfter_s"] < 1e-3 * info["maxDivergenceBefore_s"], (layout, info)
        assert info["allocatedBytes"] > 0
        assert o["progress"][0] == "planning" and o["progress"][-1] == "attach"
    t = out["tall-cell"]["info"]
    assert t["storedNy"] == 34 and t["ny"] == 64 and t["regularLayers"] == 32
    assert t["compressionRatio"] == pytest.approx(64 / 34)
    # the uniform step advects on the GPU before each projection, either scheme
    for name in ("maccormack", "runtimeSwitch"):
        info = out[name]["info"]
        assert info["steps"] == (3 if name == "maccormack" else 2), (name, info)
        assert info["advectionDeviceMs"] > 0 and info["pressureFlags"] == 0, (name, info)
        assert info["pressureSolveConverged"] is True, (name, info)
        assert info["maxDivergenceAfter_s"] < 1e-3 * info["maxDivergenceBefore_s"], (name, info)
    assert out["uniform"]["info"]["advectionDeviceMs"] > 0
    for name in ("cm11aOpenTop", "cm11aClosed"):
        o = out[name]
        info = o["info"]
        assert o["stepped"] == [True, True] and info["steps"] == 3, (name, info)
        assert "CM11a" in info["pressureSolver"] and ("open top" in info["pressureSolver"]) == (name == "cm11aOpenTop")
        assert info["pressureFlags"] == 0 and info["pressureSolveConverged"] is True, (name, info)
        assert info["maxDivergenceAfter_s"] < 0.1 * info["maxDivergenceBefore_s"], (name, info)
    e = out["entryPoints"]
    assert e["submitEqualsProject"] and e["batchEqualsProject"] and e["flags"] == [0, 0, 0, 0]
    assert e["latencyMs"] > 0
    assert out["destroyWhilePending"] == {"flags": 0, "freed": True}


def test_addon_errors_are_js_errors_without_device(fpl):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    r = run_node("const m=require('./fluid_amd/node');"
                 "try{m.binding.createContext(0,0)}catch(e){console.log(e.code)}")
    assert r.returncode == 0 and r.stdout.strip() == "FP_ERR_2"


@pytest.mark.gpu
def test_addon_projection_matches_c_abi(fpl, tmp_path):
    from fluid_amd import scenes
    s = scenes.dam_break_scene(32, 3)
    for name, a in (("phi", s.phi), ("u", s.u), ("v", s.v), ("w", s.w)):
        a.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    nx, ny, nz = s.shape
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "project_files.js"), str(tmp_path), str(nx), str(ny),
                        str(nz), repr(s.h), repr(s.dt), repr(s.rho), "fp16"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout.strip().splitlines()[-1])["report"]
    assert rep["flags"] == 0 and rep["relativeResidual"] <= 1e-6
    g = fpl.UniformPressureProjection(s.shape, s.h)
    ref = g.project(s.phi, s.u, s.v, s.w, s.dt, s.rho, precision="fp16")
    g.destroy()
    p = np.fromfile(tmp_path / "pressure.out.f32", np.float32).reshape(s.shape)
    v = np.fromfile(tmp_path / "v.out.f32", np.float32).reshape(s.v.shape)
    assert np.abs(p - ref.pressure).max() <= 1e-6 * np.abs(ref.pressure).max()
    assert np.abs(v - ref.v).max() <= 1e-6


@pytest.mark.gpu
def test_addon_tall_remesh_matches_oracle(fpl, tmp_path):
    """TallCellPressureProjection.remesh through the addon: the re-laid
    fields and bases equal the CPU restatement bit for bit; refused while a
    projection is in flight; a second remesh changes nothing; the grid then
    projects."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tall_remesh as orm
    from remesh_cases import moved_scene
    lay, phi, u, v, w = moved_scene(seed=11, drop=20)
    lay.bases.astype(np.int32).tofile(tmp_path / "bases.i32")
    for name, a in (("phi", phi), ("u", u), ("v", v), ("w", w)):
        a.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "remesh_files.js"), str(tmp_path), str(lay.nx),
                        str(lay.fine_ny), str(lay.nz), str(lay.band), repr(lay.h), str(lay.max_delta)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    want = orm.remesh(lay.bases, phi, u, v, w, lay.fine_ny, lay.band, lay.max_delta, lay.h)
    assert out == {"refusedInFlight": True, "changed": want[5], "again": 0, "flags": 0, "basesKept": True}
    np.testing.assert_array_equal(np.fromfile(tmp_path / "bases.out.i32", np.int32).reshape(lay.bases.shape), want[0])
    for name, a in zip(("phi", "u", "v", "w"), want[1:5]):
        got = np.fromfile(tmp_path / f"{name}.out.f32", np.float32).reshape(a.shape)
        assert np.array_equal(got.view(np.uint32), a.view(np.uint32)), name


@pytest.mark.gpu
def test_addon_extrapolation_matches_oracle(fpl, oracle, tmp_path):
    """UniformPressureProjection.extrapolate through the addon: u, v, w and
    the report equal the CPU restatement bit for bit."""
    from fluid_amd import scenes
    s = scenes.dam_spheres_scene(40, seed=2)
    for name, a in (("phi", s.phi), ("u", s.u), ("v", s.v), ("w", s.w)):
        a.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    nx, ny, nz = s.shape
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "extrapolate_files.js"), str(tmp_path), str(nx),
                        str(ny), str(nz), repr(s.h)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rep = json.loads(r.stdout.strip().splitlines()[-1])
    uo, vo, wo, want = oracle.extrapolate(s.u, s.v, s.w, s.h, phi=s.phi)
    assert (rep["passes"], rep["activeLeft"], rep["levels"], rep["bandFaces"], rep["unknownFaces"]) == \
        (want["passes"], want["active_left"], want["levels"], want["band_known"], want["unknown"])
    for name, a in (("u", uo), ("v", vo), ("w", wo)):
        got = np.fromfile(tmp_path / f"{name}.out.f32", np.float32).reshape(a.shape)
        assert np.array_equal(got.view(np.uint32), a.view(np.uint32)), name


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,nu,no_slip", [("semi-lagrangian", 0.0, 0), ("maccormack", 1e-3, 1)])
def test_addon_advection_matches_oracle(fpl, oracle, tmp_path, scheme, nu, no_slip):
    """UniformPressureProjection.advect through the addon (through the field's
    own extrapolation, gravity, viscosity, both schemes): u, v, w equal the CPU
    restatement bit for bit; an unknown scheme is a RangeError."""
    from fluid_amd import scenes
    s = scenes.dam_spheres_scene(40, seed=4)
    for name, a in (("phi", s.phi), ("u", s.u), ("v", s.v), ("w", s.w)):
        a.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    nx, ny, nz = s.shape
    dt = 1.0 / 30.0
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "advect_files.js"), str(tmp_path), str(nx),
                        str(ny), str(nz), repr(s.h), repr(dt), scheme, repr(nu), str(no_slip)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1]) == {"refusedScheme": True, "timed": True}
    want = oracle.advect_velocity(s.u, s.v, s.w, s.h, np.float32(dt), phi=s.phi, gravity=np.float32(-9.80665),
                                  nu=np.float32(nu), scheme=scheme, no_slip=no_slip)
    for name, a in zip("uvw", want):
        got = np.fromfile(tmp_path / f"{name}.out.f32", np.float32).reshape(a.shape)
        assert np.array_equal(got.view(np.uint32), a.view(np.uint32)), name
