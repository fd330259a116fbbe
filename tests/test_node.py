"""The Node-API layer: the addon loads and exposes the reference-shaped API
(CPU); the Cube Drop flow run from JavaScript equals the same flow run from
Python through the same C ABI (GPU)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "webgpu-fluid-simulation_amd", "node")
ADDON = os.path.join(NODE_DIR, "pbf_addon.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or addon not built")


@needs_node
def test_addon_exports_reference_api():
    js = ("const a=require(%r); const {PBF}=require(%r);"
          "console.log(JSON.stringify({fns:Object.keys(a), methods:Object.getOwnPropertyNames(PBF.prototype),"
          "version:PBF.version()}))") % (ADDON, os.path.join(NODE_DIR, "PBF.js"))
    out = json.loads(subprocess.run([NODE, "-e", js], capture_output=True, text=True, check=True).stdout)
    for m in ("reset", "setConfig", "voxelizeCube", "voxelizeSphere", "setParticlePosition", "setBoundaryData",
              "run", "start", "stop", "switch", "step", "readPositions", "readVelocities", "densityError",
              "voxelizeMesh", "setWarmStart", "autotune", "setTuning"):
        assert m in out["methods"], m
    assert out["version"].startswith("pbf-mi355x")


@needs_node
def test_failed_create_throws_the_engine_message():
    # parameter validation runs before any device call: the thrown Error
    # carries pbf_last_error(NULL); NaN counts saturate and are refused
    js = ("const {PBF}=require(%r); const out=[];"
          "for (const p of [{particleRadius: NaN}, {particleRadius: 0.025, capacity: NaN},"
          "                 {particleRadius: 0.025, gridMode: 7}]) {"
          "  try { new PBF(p); out.push('created'); } catch (e) { out.push(e.message); } }"
          "console.log(JSON.stringify(out))") % os.path.join(NODE_DIR, "PBF.js")
    out = json.loads(subprocess.run([NODE, "-e", js], capture_output=True, text=True, check=True).stdout)
    assert "particle_radius must be finite" in out[0]
    assert "capacity above" in out[1]
    assert "grid_mode must be" in out[2]


@needs_node
@pytest.mark.gpu
def test_node_cube_drop_equals_python(tmp_path):
    import scenes
    from pbf_host import PBF
    f = tmp_path / "p.bin"
    res = subprocess.run([NODE, os.path.join(NODE_DIR, "run_scene.js"), "3", str(f)], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    info = json.loads(res.stdout.strip().splitlines()[-1])
    assert info["particles"] == 41 ** 3
    assert info["error"] == "Illegal length of Particle Position Buffer!"
    pj = np.fromfile(f, np.float32).reshape(-1, 4)
    sim = PBF()
    sim.reset()
    sim.voxelizeCube((0.15, 0.35, 0.15), (0.65, 0.85, 0.65))
    sim.setConfig(scenes.REF_CONFIG)
    sim.setParticlePosition()
    sim.start()
    for _ in range(3):
        sim.run()
    assert np.array_equal(pj, sim.read_positions())
