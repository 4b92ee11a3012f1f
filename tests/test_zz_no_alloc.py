"""wig_air_project_static, the allocation-free air projection, against
wig_air_project on the same inputs, through the C ABI: every buffer the
same bits after each call.  The calls share the library's one static
working array, so they run in a row on it: a grid, then a smaller one
(what the first left in the array must not reach the second), then a grid
of one level, each over warm-started steps; and a grid past the array's
size is refused before anything runs."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import scenes
from worldinagrain_amd.air import GpuAir, PART_MG, PART_CG, PART_STATUS, PART_MEAN, STATIC_MAX_CELLS, project_static
from worldinagrain_amd.gpu import HipStream, WigError

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def parts(g):
    return [g.download(p) for p in (PART_MG, PART_CG, PART_STATUS)]


def same(a, b, what):
    for name, x, y in zip(("mg", "cg", "status"), parts(a), parts(b)):
        d = np.nonzero(x.view(np.uint32) != y.view(np.uint32))[0]
        assert d.size == 0, f"{what}: {name} differs at {d.size} words, first {d[:6]}"


def new_flow(g, W, H, seed, step):
    mg0, _ = scenes.air_rooms(W, H, seed=seed, flow_seed=2000 + step)
    mg = g.download(PART_MG)
    mg[:8 * W * H] = mg0.reshape(-1)
    g.upload(PART_MG, mg)


def test_static_projection_equals_the_allocating_one():
    stream = HipStream()
    try:
        # large, then smaller, then one level: the array keeps the last call's words
        for W, H, seed, steps in ((300, 200, 5, 3), (96, 80, 3, 3), (16, 12, 2, 2)):
            assert W * H <= STATIC_MAX_CELLS
            say(f"[no-alloc] {W}x{H} seed {seed}, {steps} steps")
            mg0, x0 = scenes.air_rooms(W, H, seed=seed)
            ref, buf = GpuAir(W, H), GpuAir(W, H)
            buf.set_stream(stream.handle)
            ref.load(mg0, x0)
            buf.load(mg0, x0)
            ptrs = [buf.device_ptr(p) for p in (PART_MG, PART_CG, PART_STATUS, PART_MEAN)]
            for step in range(steps):
                ref.project(64)
                project_static(*ptrs, W, H, 64, stream.handle)
                ref.sync()
                buf.sync()
                same(buf, ref, f"{W}x{H} step {step}")
                its = int(ref.download(PART_STATUS)[4])
                say(f"[no-alloc]   step {step}: {its} iterations, same bits")
                if W * H > 256:
                    assert its > 0
                new_flow(ref, W, H, seed, step)
                new_flow(buf, W, H, seed, step)
            buf.close()
            ref.close()
    finally:
        stream.close()


def test_static_projection_refuses_what_it_cannot_hold():
    g = GpuAir(8, 8)
    ptrs = [g.device_ptr(p) for p in (PART_MG, PART_CG, PART_STATUS, PART_MEAN)]
    with pytest.raises(WigError, match="262144"):
        project_static(*ptrs, 1024, 1024, 64)
    with pytest.raises(WigError, match="NULL"):
        project_static(None, *ptrs[1:], 8, 8, 64)
    with pytest.raises(WigError, match="4096"):
        project_static(*ptrs, 8, 8, 5000)
    g.close()
