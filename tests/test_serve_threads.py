"""The serving loop's eviction and admission threads (serve.Evictor,
serve.Admitter), with stand-in worlds on the host: every eviction runs
once, the bytes in flight stay within the limit (one world past it may run
alone), waits return once a world's eviction is done, an eviction's failure
reaches the loop, and an admission waits for its session's eviction."""
import threading
import time

import pytest

from worldinagrain_amd import serve


class _World:
    def __init__(self):
        self.stream = None

    def set_stream(self, s):
        self.stream = s


class _Session:
    def __init__(self, sid, nbytes, ev, fail=False):
        self.sid, self.bytes, self.world, self.evicting = sid, nbytes, _World(), None
        self.ev, self.fail, self.evicted = ev, fail, 0

    def evict(self):
        with self.ev.cv:
            self.ev.peak = max(getattr(self.ev, "peak", 0), self.ev.inflight)
        time.sleep(0.005)
        if self.fail:
            raise OSError("copy failed")
        assert self.world.stream == "copy"
        self.evicted += 1
        self.world = None


def test_evictions_bounded_and_waited():
    ev = serve.Evictor("copy", 30)
    ss = [_Session(i, 10, ev) for i in range(8)] + [_Session(8, 50, ev)]    # the last alone past the limit
    for s in ss:
        ev.submit(s)
    ev.wait(ss[0])
    assert ss[0].evicted == 1 and ss[0].evicting is None
    ev.drain()
    assert all(s.evicted == 1 and s.world is None for s in ss)
    assert ev.peak <= 50 and ev.inflight == 0
    ev.close()


def test_eviction_failure_reaches_the_loop():
    ev = serve.Evictor("copy", 100)
    bad = _Session(0, 10, ev, fail=True)
    ev.submit(bad)
    with pytest.raises(RuntimeError, match="an eviction failed"):
        ev.wait(bad)
    with pytest.raises(RuntimeError):
        ev.close()


class _Admitted:
    def __init__(self, sid, log):
        self.sid, self.world, self.evicting, self.admitting, self.log = sid, None, None, None, log

    def make_resident(self, stream=None):
        self.log.append(("in", self.sid, stream))
        self.world = _World()
        if stream is not None:
            self.world.set_stream(stream)


def test_admission_waits_for_its_eviction():
    log = []
    ad = serve.Admitter()
    s = _Admitted(1, log)
    s.evicting = threading.Event()          # an eviction of the same session still in flight
    ad.submit(s)
    time.sleep(0.02)
    assert not s.admitting.is_set() and s.world is None
    log.append(("out", 1, None))
    s.evicting.set()
    s.admitting.wait(5)
    # (made on its own stream: no stream set, none shared with other copies)
    assert s.admitting.is_set() and s.world.stream is None
    assert [e[0] for e in log] == ["out", "in"]
    ad.wait_any(0.0)
    assert ad.inflight == 0
    ad.close()
