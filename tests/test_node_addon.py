"""The N-API addon: the TypeScript host's boundary onto the C ABI."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(code):
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, cwd=ROOT, timeout=120)


def test_addon_exports_and_method_shape(fpl):
    r = run_node("const m=require('./fluid_amd/node');"
                 "console.log(JSON.stringify({b:Object.keys(m.binding).sort(), id:m.mi355xProjectionMethod.id,"
                 "abi:m.binding.abiVersion(), fn:typeof m.mi355xProjectionMethod.createSolverAsync}))")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["b"] == sorted(["createContext", "destroyContext", "createGrid", "destroyGrid", "gridLevels",
                               "project", "lastError", "abiVersion"])
    assert out["id"] == "mi355x-projection" and out["abi"] == 2 and out["fn"] == "function"


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
