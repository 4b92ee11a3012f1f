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


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="no node")
def test_js_late_reports_match_python_host(tmp_path):
    # each report taken a step late (the reference's mapAsync), adaptive
    # cycles following the late reports on both hosts
    from tests.helpers import make_gpu
    from worldinagrain_amd.host import Simulation
    W, H, steps = 48, 40, 8
    cells = scenes.dam_break(W, H, block=(24, 28))
    path = tmp_path / "cells.bin"
    cells.tofile(path)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "dam_steps.js"), str(W), str(H), str(steps), str(path),
                        "late"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    w = make_gpu(W, H, cells)
    sim = Simulation(w, W, H)
    got = []
    for s in range(steps):
        sim.step(read_report="late")
        if s > 0:
            got.append(list(map(int, sim.last_stats)))
    sim.take_report(w.report_take(sim.pending_report, wait=True))
    got.append(list(map(int, sim.last_stats)))
    assert got == js["stats"]
    for name, part in (("cells", abi.PART_CELLS), ("counters", abi.PART_COUNTERS), ("field", abi.PART_FIELD)):
        assert hashlib.sha256(w.download(part).tobytes()).hexdigest() == js["digest"][name], name


SNAP_JS = os.path.join(ROOT, "tests", "js", "snapshot_check.js")


@pytest.mark.skipif(NODE is None, reason="no node")
def test_snapshot_json_numbers_as_javascript_prints_them(tmp_path):
    # the JavaScript engine's own JSON.stringify of the same float64 values
    # pins snapshot.js_json: exponents, thresholds, zeros, extremes, randoms
    from worldinagrain_amd import snapshot
    rng = np.random.default_rng(3)
    vals = [0.0, -0.0, 1.0, -2.5, 0.1, 1 / 3, 1e21, 1e20, 123456789012345680000.0, 1e-6, 1e-7, 1.5e-7,
            0.000123, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 2.0 ** 53, 2.0 ** 53 + 2,
            float(np.float32(9.81 / 3600 / 0.01)), 9.81 / 3600 / 0.01, 1 / 120, 0.01]
    vals += list(rng.normal(size=200) * 10.0 ** rng.integers(-30, 30, size=200))
    vals += list(rng.integers(0, 2 ** 32, size=20).astype(float))
    path = tmp_path / "v.f64"
    np.asarray(vals, dtype=np.float64).tofile(path)
    r = subprocess.run([NODE, SNAP_JS, "numbers", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == snapshot.js_json({"v": [float(v) for v in vals]})


@pytest.mark.skipif(NODE is None, reason="no node")
def test_snapshot_files_between_python_and_javascript(tmp_path):
    import gzip
    from worldinagrain_amd import snapshot
    rng = np.random.default_rng(4)
    parts = {f"p{i}": rng.integers(0, 256, int(n), dtype=np.uint8).tobytes() for i, n in enumerate([0, 1, 5, 8, 1023])}
    header = {"format": 1, "width": 48, "height": 40, "elements": list(snapshot.ELEMENT_IDS),
              "state": {"cur": 0, "frameNo": 12, "shiftedFirst": True, "rng": 3141592653, "metresPerCell": 0.01,
                        "earthGravity": 9.81 / 3600 / 0.01, "fluidStats": {"active": 7, "viscous": 0},
                        "fluidCycles": 16, "dt": 1 / 60},
              "parts": [{"name": k, "bytes": len(v)} for k, v in parts.items()]}
    # Python's file, read by the JavaScript host
    f = tmp_path / "py.snapshot"
    f.write_bytes(snapshot.pack(header, parts))
    r = subprocess.run([NODE, SNAP_JS, "read", str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    assert got["header"] == json.loads(snapshot.js_json(header))
    assert got["digests"] == {k: hashlib.sha256(v).hexdigest() for k, v in parts.items()}
    # the JavaScript host's container, byte for byte the same as Python's
    (tmp_path / "parts").mkdir()
    for k, v in parts.items():
        (tmp_path / "parts" / (k + ".bin")).write_bytes(v)
    (tmp_path / "h.json").write_text(json.dumps(header))
    r = subprocess.run([NODE, SNAP_JS, "pack", str(tmp_path / "h.json"), str(tmp_path / "parts")], capture_output=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == bytes(snapshot.pack_payload(header, parts))
    h, p = snapshot.unpack(gzip.compress(r.stdout))
    assert h == json.loads(snapshot.js_json(header)) and {k: bytes(v) for k, v in p.items()} == parts


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="no node")
def test_js_snapshot_continues_in_python_host(tmp_path):
    # a snapshot the JavaScript host saved, loaded by the Python host, steps
    # on to what the JavaScript host's world reached
    from tests.helpers import make_gpu, compare_worlds
    from worldinagrain_amd import snapshot
    from worldinagrain_amd.host import Simulation
    W, H, steps = 48, 40, 3
    cells = scenes.dam_break(W, H, block=(24, 28))
    path = tmp_path / "cells.bin"
    cells.tofile(path)
    a, b = tmp_path / "a.snapshot", tmp_path / "b.snapshot"
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "snapshot_steps.js"), str(W), str(H), str(steps),
                        str(path), str(a), str(b)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout)["fieldSame"]
    sim = Simulation(make_gpu(W, H, scenes.empty_cells(W, H)), W, H)
    snapshot.load(sim, a.read_bytes())
    for _ in range(steps):
        sim.step()
    end = Simulation(make_gpu(W, H, scenes.empty_cells(W, H)), W, H)
    hb = snapshot.load(end, b.read_bytes())
    assert (sim.frame_no, sim.rng_state, sim.shifted_first, sim.fluid_cycles) == \
        (end.frame_no, end.rng_state, end.shifted_first, end.fluid_cycles)
    assert hb["state"]["frameNo"] == sim.frame_no
    d = compare_worlds(sim.world, end.world)
    assert not d, "; ".join(d)


@pytest.mark.gpu
@pytest.mark.skipif(NODE is None, reason="no node")
def test_js_air_projection_matches_oracle(tmp_path):
    # the air's projection from the JavaScript host (worldinagrain_amd/js/
    # air.js over the addon): three steps with its airCycles, every buffer
    # bit-identical to the oracle's stepped with the same counts
    from oracle.air import OracleAir
    from worldinagrain_amd import scenes
    from worldinagrain_amd.air import AirProjection
    W, H, steps = 120, 88, 3
    mg0, x0 = scenes.air_rooms(W, H, seed=13)
    o = OracleAir(W, H)
    o.load(mg0, x0)
    for name, a in (("mg", o.mg), ("cg", o.cg), ("status", o.status), ("mean", o.mean)):
        a.tofile(str(tmp_path / f"{name}.bin"))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "air_steps.js"), str(W), str(H), str(steps), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p = AirProjection(o)
    its = []
    for s in range(steps):
        if s > 0:
            o.mg[:8 * W * H] = mg0.reshape(-1)
        its.append(p.step())
    assert json.load(open(tmp_path / "its.json")) == its
    for name, a in (("mg", o.mg), ("cg", o.cg), ("status", o.status)):
        got = np.fromfile(str(tmp_path / f"{name}.out"), dtype=a.dtype)
        assert np.array_equal(got.view(np.uint32), a.view(np.uint32)), name
