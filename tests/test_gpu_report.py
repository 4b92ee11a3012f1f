"""Reports asked for without waiting (wig_report_async / wig_report_take,
as the reference maps its report buffer asynchronously): each equals the
waiting report of a twin world stepped alike; the two latest tickets are
valid and older ones refused; a take that need not wait returns the words
once they have landed; a host taking each report a step late steps the
world exactly as one that waits; and the batch's reports likewise."""
import numpy as np
import pytest

from worldinagrain_amd import abi, scenes
from worldinagrain_amd.gpu import WigError
from worldinagrain_amd.host import step_batch
from tests.helpers import make_gpu, compare_worlds, config_sim

pytestmark = pytest.mark.gpu

W, H = 64, 48


def twins():
    cells = scenes.dam_break(W, H, block=(30, 36))
    a, b = make_gpu(W, H, cells), make_gpu(W, H, cells)
    return (a, config_sim(a, W, H)), (b, config_sim(b, W, H))


def test_async_report_equals_waiting_report():
    (a, sa), (b, sb) = twins()
    tickets = []
    for s in range(4):
        sa.step(read_report=False)
        sb.step(read_report=False)
        tickets.append(a.report_async())
        want = b.report()
        got = a.report_take(tickets[-1], wait=True)
        assert np.array_equal(got, want), (s, got, want)
    # the two latest are valid; older ones are refused
    with pytest.raises(WigError):
        a.report_take(tickets[0])
    assert a.report_take(tickets[-2]) is not None
    a.sync()
    assert a.report_take(tickets[-1], wait=False) is not None     # landed: no wait needed
    with pytest.raises(WigError):
        a.report_take(tickets[-1] + 1)                              # not asked for yet


def test_late_reports_step_the_world_as_waiting_ones():
    (a, sa), (b, sb) = twins()
    prev = None
    for s in range(10):
        sa.step(read_report="late")
        sb.step()
        if prev is not None:
            assert np.array_equal(sa.last_stats, prev), s          # a step behind
        prev = sb.last_stats.copy()
    assert not compare_worlds(a, b)


def test_batch_async_reports():
    from worldinagrain_amd.gpu import GpuBatch
    (a, sa), (b, sb) = twins()
    cells = scenes.dam_break(32, 32, block=(16, 20))
    c, d = make_gpu(32, 32, cells), make_gpu(32, 32, cells)
    sc, sd = config_sim(c, 32, 32), config_sim(d, 32, 32)
    b1, b2 = GpuBatch([a, c]), GpuBatch([b, d])
    for s in range(3):
        step_batch(b1, [sa, sc], read_report=False)
        step_batch(b2, [sb, sd], read_report=False)
        t = b1.report_async()
        assert np.array_equal(b1.report_take(t), b2.report()), s
    b1.assign([c])
    with pytest.raises(WigError):
        b1.report_take(t)                                           # the worlds before the assign
