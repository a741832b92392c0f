"""The scene format: a SceneDescription (lib/core/model.ts:52-130) as the
inputs of a uniform grid (fluid_amd.scenes.scene_from_description), by the
same rules as the Node plugin's createSolverAsync.

Checks: the lattice of the reference's default scene (1.2 x 0.8 x 0.8 m at
5 cm: 24 x 16 x 16); the dam-break reservoir against its closed form (the
liquid cells are exactly the cell centres inside the box of height 0.92 and
footprint sqrt(fill / 0.92)); tank-fill levels; authored dimensions and
origin; the step parameters (nu = mu / rho, gravity, walls, top, tolerance);
the level set equal to the Node plugin's for the same scenes; refused forms.
"""
import copy
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from fluid_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
REFERENCE_DEFAULT = "/root/reference/lib/core/default-scene.json"

# the fluid-relevant fields of the reference's default scene (lib/core/default-scene.json)
DEFAULT = {
    "schemaVersion": "1.0.0", "sceneId": "interactive-water-box", "randomSeed": 20260712, "duration_s": 20,
    "container": {"width_m": 1.2, "height_m": 0.8, "depth_m": 0.8, "fillFraction": 0.22, "top": "closed",
                  "fluidWallMode": "no-slip"},
    "voxelDomain": {"finestCellSize_m": 0.05, "brickSize_cells": 8},
    "fluid": {"density_kg_m3": 998.2, "dynamicViscosity_Pa_s": 0.001002, "surfaceTension_N_m": 0.072,
              "gravity_m_s2": {"x": 0, "y": -9.80665, "z": 0}, "initialCondition": "dam-break"},
    "nominalResolution": {"length_m": 0.025},
    "numerics": {"fixedDt_s": 0.016, "maxDt_s": 0.016, "pressureRelativeTolerance": 1e-8,
                 "pressureMaxIterations": 1000},
    "rigidBodies": [],
}


def variant(**changes):
    s = copy.deepcopy(DEFAULT)
    for path, value in changes.items():
        node = s
        keys = path.split("__")
        for k in keys[:-1]:
            node = node[k]
        node[keys[-1]] = value
    return s


def test_default_scene_lattice_and_step_parameters():
    sc = scenes.scene_from_description(DEFAULT)
    assert sc.shape == (24, 16, 16) and sc.h == pytest.approx(0.05)
    assert sc.dt == 0.016 and sc.rho == 998.2
    m = sc.meta
    assert m["nu"] == 0.001002 / 998.2 and m["gravity"] == -9.80665
    assert m["no_slip"] and not m["open_top"] and m["relative_tolerance"] == 1e-8 and m["max_iterations"] == 1000
    assert not sc.u.any() and not sc.v.any() and not sc.w.any()
    assert sc.u.shape == (25, 16, 16) and sc.v.shape == (24, 17, 16) and sc.w.shape == (24, 16, 17)


def test_dam_break_box_closed_form():
    sc = scenes.scene_from_description(DEFAULT)
    n = sc.shape
    h = sc.h
    height, foot = 0.92, math.sqrt(0.22 / 0.92)
    c = [(np.arange(k) + 0.5) * h for k in n]
    inside = ((c[0][:, None, None] < foot * 1.2) & (c[1][None, :, None] < height * 0.8)
              & (c[2][None, None, :] < foot * 0.8))
    np.testing.assert_array_equal(sc.phi < 0, inside)
    # signed distance: the cell centre nearest the box's top face inside it
    j = int(height * 0.8 / h - 0.5)
    assert sc.phi[0, j, 0] == pytest.approx((j + 0.5) * h - height * 0.8, abs=1e-7)


def test_tank_fill_and_authored_box():
    tank = scenes.scene_from_description(variant(fluid__initialCondition="tank-fill"))
    level = 0.22 * 0.8
    y = (np.arange(16) + 0.5) * 0.05
    np.testing.assert_allclose(tank.phi[3, :, 5], (y - level).astype(np.float32), rtol=0, atol=1e-7)
    s = variant(fluid__initialDamBreakDimensions_m={"x": 0.3, "y": 0.4, "z": 0.2},
                fluid__initialDamBreakOrigin_m={"x": 0.6, "y": 0.0, "z": 0.1})
    lo, hi = scenes.dam_break_box(s)
    assert lo == pytest.approx((0.5, 0.0, 0.125)) and hi == pytest.approx((0.75, 0.5, 0.375))
    sc = scenes.scene_from_description(s)
    liquid = np.argwhere(sc.phi < 0)
    assert liquid[:, 0].min() == 12 and liquid[:, 0].max() == 17  # x in [0.6, 0.9) m
    assert liquid[:, 1].max() == 7 and liquid[:, 2].min() == 2 and liquid[:, 2].max() == 5


def test_json_text_path_and_reference_file(tmp_path):
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(DEFAULT))
    a = scenes.scene_from_description(str(p))
    b = scenes.scene_from_description(json.dumps(DEFAULT))
    np.testing.assert_array_equal(a.phi, b.phi)
    if os.path.exists(REFERENCE_DEFAULT):
        r = scenes.scene_from_description(REFERENCE_DEFAULT)
        np.testing.assert_array_equal(r.phi, a.phi)
        assert r.meta == a.meta and (r.dt, r.rho, r.h) == (a.dt, a.rho, a.h)


def test_refused_forms():
    with pytest.raises(ValueError):
        scenes.scene_from_description(variant(container__shape="sphere"))
    with pytest.raises(ValueError):
        scenes.scene_from_description(variant(container__height_m=0.83))  # non-cubic cells
    with pytest.raises(ValueError):
        scenes.scene_from_description(variant(fluid__initialCondition="droplet"))
    bad = copy.deepcopy(DEFAULT)
    del bad["numerics"]
    with pytest.raises(ValueError):
        scenes.scene_from_description(bad)


@pytest.mark.skipif(NODE is None, reason="node is not installed")
@pytest.mark.parametrize("name", ["default", "tank", "authored"])
def test_level_set_equals_the_node_plugin(tmp_path, name):
    s = {"default": DEFAULT, "tank": variant(fluid__initialCondition="tank-fill"),
         "authored": variant(fluid__initialDamBreakDimensions_m={"x": 0.3, "y": 0.4, "z": 0.2},
                             fluid__initialDamBreakOrigin_m={"x": 0.6, "y": 0.0, "z": 0.1})}[name]
    (tmp_path / "scene.json").write_text(json.dumps(s))
    code = ("const m=require('./fluid_amd/node');const fs=require('fs');"
            f"const s=JSON.parse(fs.readFileSync({json.dumps(str(tmp_path / 'scene.json'))},'utf8'));"
            "const lat=m.sceneLattice(s);const phi=m.uniformLevelSet(s,lat);"
            f"fs.writeFileSync({json.dumps(str(tmp_path / 'phi.f32'))},Buffer.from(phi.buffer));"
            "console.log(JSON.stringify(lat))")
    r = subprocess.run([NODE, "-e", code], capture_output=True, text=True, cwd=ROOT, timeout=60)
    assert r.returncode == 0, r.stderr
    lat = json.loads(r.stdout.strip())
    sc = scenes.scene_from_description(s)
    assert (lat["nx"], lat["ny"], lat["nz"]) == sc.shape and lat["h"] == sc.h
    node_phi = np.fromfile(tmp_path / "phi.f32", np.float32).reshape(sc.shape)
    # both evaluate in double and round to float32; the Euclidean norm of the
    # outside distance may differ in the last double bit (Math.hypot)
    assert np.abs(node_phi - sc.phi).max() <= 1e-7
    np.testing.assert_array_equal(node_phi < 0, sc.phi < 0)


@pytest.mark.gpu
def test_scene_steps_on_the_gpu(fpl, oracle):
    """Three velocity steps of the default scene through the Python API:
    fp_advect_velocity with the scene's gravity, viscosity and no-slip walls
    (bit-exact against the oracle on the first step), then the projection to
    the scene's tolerance (converged, divergence reduced)."""
    sc = scenes.scene_from_description(DEFAULT)
    m = sc.meta
    g = fpl.UniformPressureProjection(sc.shape, sc.h)
    u, v, w = sc.u, sc.v, sc.w
    try:
        for step in range(3):
            print(f"scene step {step}", flush=True)
            au, av, aw, _ = g.advect(u, v, w, sc.dt, phi=sc.phi, gravity=m["gravity"], viscosity=m["nu"],
                                     no_slip=m["no_slip"], open_top=m["open_top"])
            if step == 0:
                want = oracle.advect_velocity(u, v, w, sc.h, sc.dt, phi=sc.phi, gravity=m["gravity"], nu=m["nu"],
                                              no_slip=1)
                for a, b in zip((au, av, aw), want):
                    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            r = g.project(sc.phi, au, av, aw, sc.dt, sc.rho, relative_tolerance=m["relative_tolerance"],
                          max_iterations=m["max_iterations"])
            rep = r.report
            assert rep["flags"] == 0 and rep["relative_residual"] <= 1e-8, rep
            assert rep["max_div_after"] < 1e-4 * rep["max_div_before"], rep
            u, v, w = r.u, r.v, r.w
        assert np.abs(v).max() > 0  # the liquid is moving
    finally:
        g.destroy()
