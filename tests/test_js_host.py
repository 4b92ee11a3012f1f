"""The JavaScript host (worldinagrain_amd/js/liquid.js over the wig.node
Node-API addon) drives the HIP step exactly as the Python host does: the
same Params (mulberry32 seeds, frames), the same control, the same bytes."""
import hashlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "worldinagrain_amd", "js", "wig.node")


@pytest.mark.skipif(NODE is None, reason="no node")
def test_addon_loads_without_device():
    r = subprocess.run([NODE, "-e", "const l=require(process.argv[1]); console.log(typeof l.LiquidSimulation, l.deviceOk())",
                        os.path.join(ROOT, "worldinagrain_amd", "js", "liquid.js")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("function")


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="no node")
def test_js_host_matches_python_host(tmp_path):
    from tests.helpers import make_gpu
    from worldinagrain_amd.host import Simulation
    W, H, steps = 48, 40, 8
    cells = scenes.dam_break(W, H, block=(24, 28))
    path = tmp_path / "cells.bin"
    cells.tofile(path)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "dam_steps.js"), str(W), str(H), str(steps), str(path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    w = make_gpu(W, H, cells)
    sim = Simulation(w, W, H)
    for s in range(steps):
        sim.step()
        assert list(map(int, sim.last_stats)) == js["stats"][s], s
    for name, part in (("cells", abi.PART_CELLS), ("counters", abi.PART_COUNTERS), ("field", abi.PART_FIELD)):
        assert hashlib.sha256(w.download(part).tobytes()).hexdigest() == js["digest"][name], name
