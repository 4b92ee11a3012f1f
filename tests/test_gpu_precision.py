"""The V-cycle in 16-bit storage (IEEE half, bfloat16; arithmetic in float):
the outer PCG still solves to the relative tolerance, its pressure against
the FP64 CG oracle on the same matrix; and a batch mixing storages keeps
each world as it is stepped alone."""
import sys

import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd.host import Simulation, SolverSettings, step_batch
from tests.helpers import make_oracle, make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("storage", [abi.WIG_STORE_F16, abi.WIG_STORE_BF16])
def test_16bit_vcycle_pressure_against_fp64_cg(storage):
    """Four FP32 steps (bit-identical to the oracle), then one step with the
    16-bit V-cycle: its pressure within 1e-5 of the FP64 CG solution of the
    oracle's matrix for that step, and ||r|| <= 1e-6 ||b||."""
    W = H = 96
    cells = scenes.dam_break(W, H, block=(48, 70))
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = config_sim(wo, W, H, tol=1e-6), config_sim(wg, W, H, tol=1e-6)
    for _ in range(4):
        so.step()
        sg.step()
    assert not compare_worlds(wo, wg)
    wo.capture(True)
    so.step()
    sg.solver = SolverSettings.relative(1e-6, storage=storage)
    sg.step()
    b = wo.captured_b()
    x64, it, rel = wo.cg64(b, tol=1e-12)
    f = wg.download(abi.PART_FIELD)[:8 * W * H].reshape(-1, 8)
    fl = f[:, 2] > 0
    err = np.linalg.norm(f[fl, 0] - x64[fl]) / np.linalg.norm(x64[fl])
    s = wg.solve_stats()
    say(f"  storage {storage}: {s.cycles} iterations (FP32: {wo.solve_stats().cycles}), error vs FP64 CG {err:.3e}")
    assert err < 1e-5
    assert s.rr <= 1e-12 * s.bb
    assert s.cycles <= 2 * wo.solve_stats().cycles + 2


def test_batch_mixing_storages():
    from worldinagrain_amd.gpu import GpuBatch
    W, H = 160, 120
    cells, _ = scenes.discs_dam(W, H, discs=12, rmin=3, rmax=9)
    stores = [abi.WIG_STORE_F32, abi.WIG_STORE_F16, abi.WIG_STORE_BF16]
    inb, alone = [], []
    for st in stores:
        for pair in (inb, alone):
            w = make_gpu(W, H, cells)
            sim = config_sim(w, W, H)
            sim.solver = SolverSettings.relative(1e-6, storage=st)
            pair.append((w, sim))
    batch = GpuBatch([w for w, _ in inb])
    for s in range(4):
        step_batch(batch, [sim for _, sim in inb])
        for _, sim in alone:
            sim.step()
    for st, (wb, _), (wa, _) in zip(stores, inb, alone):
        d = compare_worlds(wb, wa)
        assert not d, f"storage {st}: " + "; ".join(d)
    # and the 16-bit worlds did run a different preconditioner
    assert compare_worlds(inb[0][0], inb[1][0])


def test_gpu_fp64_check_matches_oracle_cg64():
    """wig_cg64 (FP64 CG on the device, on the solve's packed matrix and its
    captured rhs) against the oracle's FP64 CG on the same step's matrix;
    and the errors it reports against what the oracle's solution gives."""
    W = H = 96
    cells = scenes.dam_break(W, H, block=(48, 70))
    wo, wg = make_oracle(W, H, cells), make_gpu(W, H, cells)
    so, sg = config_sim(wo, W, H, tol=1e-6), config_sim(wg, W, H, tol=1e-6)
    for _ in range(4):
        so.step()
        sg.step()
    wo.capture(True)
    wg.capture(True)
    so.step()
    sg.step()
    x_cpu, _, _ = wo.cg64(wo.captured_b(), tol=1e-12)
    chk, x_gpu = wg.cg64(tol=1e-12)
    f = wg.download(abi.PART_FIELD)[:8 * W * H].reshape(-1, 8)
    fl = f[:, 2] > 0
    d = np.linalg.norm(x_gpu[fl] - x_cpu[fl]) / np.linalg.norm(x_cpu[fl])
    err_cpu = np.linalg.norm(f[fl, 0] - x_cpu[fl]) / np.linalg.norm(x_cpu[fl])
    say(f"  FP64 CG: {chk.iterations} iterations, relres {chk.relres64:.2e}; GPU vs CPU x64 {d:.2e}; "
        f"solve err {chk.err:.3e} (vs CPU x64 {err_cpu:.3e}), solve relres {chk.relres:.2e}")
    assert chk.relres64 <= 1e-11
    assert d < 1e-9
    assert abs(chk.err - err_cpu) <= 1e-3 * err_cpu + 1e-12
    assert chk.relres < 1e-4
