"""The serving loop: step requests from many sessions, each session a world
kept resident on the device between its requests, served by continuous
batching.

A request asks for `steps` consecutive liquid steps of one session's world.
The server runs rounds.  A round takes the requests that have arrived, in
arrival order, and makes their worlds resident: a world is admitted only if
it fits the HBM budget, after idle worlds (no request waiting) are evicted
to host memory, least recently used first; a request whose world cannot be
admitted waits for a later round.  Given a copy stream, evictions and
admissions run in threads of their own on it while the rounds go on
(Evictor, Admitter), and a world joins the first round after it is ready.  Then one step of every admitted
request's world runs in a single batch step (wig_batch_step).  Requests
that arrive while a round is on the device join the next round, so their
steps overlap with the others'.  A request is done when its world has run
its steps; its latency runs from its arrival to the end of that round.

Eviction takes the world's snapshot (snapshot.take: every part by the
reference's names, the solver's words included, and the simulation's
state) into host memory, without the gzip container, and frees the world;
admission loads it into a new world (snapshot.put).  Each world's results
are those of the same world stepped alone whatever it was batched with and
whenever it was evicted: the batch step is exact per world and eviction is
a copy.  Session.save() gives an idle session's world as a snapshot file's
bytes."""
import collections
import queue
import threading
import time

import numpy as np

from . import abi, snapshot
from .gpu import GpuWorld, GpuBatch
from .host import Simulation, BatchHost


def world_bytes(width, height, coarse_levels=0):
    """HBM a world of this size takes (its parts and the solve's arrays)."""
    c = width * height
    b = ((width + 7) // 8) * ((height + 7) // 8)
    cap = int(np.ceil(c * 4 * 1.25))
    f = (width + 1) * height + width * (height + 1)
    field = 16 * (2 * c + f + 2 * ((f + 3) // 4) + 2 * (c // 3 + 64))
    parts = 24 * c + 2 * (16 + 48 * cap) + 4 * (11 * c + 2 * f) + field + 4 * (5 * b + 9) + 16 * (c + b + 1) + 32 * b + 64
    # the packed fine level (73 B a cell, block-major: 64 cells a block), the
    # 16-bit V-cycle's levels (12 B a cell of each, about 4 B a fine cell)
    return parts + 73 * 64 * b + 4 * c + 4 * b + 2048


class Session:
    """A sandbox session: its world's size, its initial cells, its solver
    settings; its world on the device or its snapshot on the host; the
    host-side simulation state (frame, seed, report) that goes with it."""

    def __init__(self, sid, width, height, cells, solver_factory, kinds, coarse_levels=0, sim_kw=None):
        self.sid = sid
        self.width, self.height = width, height
        self.cells = cells
        self.kinds = kinds
        self.coarse_levels = coarse_levels
        self.solver_factory = solver_factory
        self.sim_kw = sim_kw or {}
        self.world = None
        self.snapshot = None
        self.sim = None
        self.last_used = -1
        self.bytes = world_bytes(width, height, coarse_levels)
        self.steps_run = 0
        self.evicting = None    # an eviction in flight: set when it is done
        self.admitting = None   # an admission in flight: set when the world is ready

    def make_resident(self, stream=None):
        w = GpuWorld(self.width, self.height, coarse_levels=self.coarse_levels, kinds=self.kinds)
        if stream is not None:
            w.set_stream(stream)
        if self.snapshot is None:
            w.upload(abi.PART_CELLS, self.cells)
            self.sim = Simulation(w, self.width, self.height, solver=self.solver_factory(), **self.sim_kw)
        else:
            self.sim.world = w
            snapshot.put(self.sim, *self.snapshot)
            self.snapshot = None
        self.world = w

    def evict(self):
        self.snapshot = snapshot.take(self.sim)
        self.world.close()
        self.world = None

    def save(self, level=6):
        """The session's world as a snapshot file's bytes (resident or not)."""
        if self.world is not None:
            return snapshot.save(self.sim, level=level)
        return snapshot.pack(*self.snapshot, level=level)


class Evictor:
    """Evictions off the serving loop: a thread of its own takes each idle
    world's snapshot on a copy stream of its own while the rounds run on the
    server's, then frees the world.  The device holds the worlds in flight
    beside the budget; past `limit` bytes of them the loop waits for the
    thread.  (The worlds evicted are idle: their last round is done, and no
    round uses them until their eviction is.)"""

    def __init__(self, stream, limit):
        self.stream, self.limit = stream, limit
        self.inflight = 0
        self.error = None
        self.cv = threading.Condition()
        self.q = queue.Queue()
        self.thread = threading.Thread(target=self._run, name="evictor", daemon=True)
        self.thread.start()

    def submit(self, s):
        with self.cv:
            while self.inflight > 0 and self.inflight + s.bytes > self.limit:
                self.cv.wait()
            self._check()
            self.inflight += s.bytes
        s.evicting = threading.Event()
        s.world.set_stream(self.stream)
        self.q.put(s)

    def wait(self, s):
        """Waits for s's eviction, if one is in flight."""
        if s.evicting is not None:
            s.evicting.wait()
            s.evicting = None
            with self.cv:
                self._check()

    def drain(self):
        """Waits for every eviction in flight."""
        with self.cv:
            while self.inflight > 0:
                self.cv.wait()
            self._check()

    def _check(self):
        if self.error is not None:
            raise RuntimeError("an eviction failed") from self.error

    def _run(self):
        while True:
            s = self.q.get()
            if s is None:
                return
            try:
                s.evict()
            except BaseException as e:      # (raised on the loop's next submit or wait)
                self.error = e
            with self.cv:
                self.inflight -= s.bytes
                self.cv.notify_all()
            s.evicting.set()

    def close(self):
        self.q.put(None)
        self.thread.join()
        self._check()


class Admitter:
    """Admissions off the serving loop: a thread of its own makes each
    session's world resident (created, and its cells or its snapshot
    uploaded) on the world's own stream while the rounds run; the loop
    takes the session into a round once it is ready.  (A session evicted and
    asked for again is admitted once its eviction is done.)"""

    def __init__(self):
        self.error = None
        self.inflight = 0
        self.cv = threading.Condition()
        self.q = queue.Queue()
        self.thread = threading.Thread(target=self._run, name="admitter", daemon=True)
        self.thread.start()

    def submit(self, s):
        self._check()
        s.admitting = threading.Event()
        with self.cv:
            self.inflight += 1
        self.q.put(s)

    def wait_any(self, timeout):
        """Waits (at most timeout s) for an admission in flight to be done."""
        with self.cv:
            if self.inflight > 0:
                self.cv.wait(timeout)
        self._check()

    def _check(self):
        if self.error is not None:
            raise RuntimeError("an admission failed") from self.error

    def _run(self):
        while True:
            s = self.q.get()
            if s is None:
                return
            try:
                if s.evicting is not None:
                    s.evicting.wait()
                s.make_resident()
            except BaseException as e:      # (raised on the loop's next submit or wait)
                self.error = e
            s.admitting.set()
            with self.cv:
                self.inflight -= 1
                self.cv.notify_all()

    def close(self):
        self.q.put(None)
        self.thread.join()
        self._check()


class Request:
    __slots__ = ("sid", "steps", "arrival", "left", "done_at")

    def __init__(self, sid, steps, arrival):
        self.sid, self.steps, self.arrival = sid, steps, arrival
        self.left = steps
        self.done_at = None


class Server:
    """Continuous batching over sessions under an HBM budget (bytes)."""

    def __init__(self, sessions, budget, max_round_cells=1 << 24, stream=None, copy_stream=None, evicting_limit=16 << 30):
        """copy_stream: evictions run in a thread of their own on this stream
        (Evictor), at most evicting_limit bytes of worlds in eviction at
        once, and admissions in another (Admitter), each world made and
        loaded on its own stream; None: each in the loop, in the round that
        needs it.  (A world joining a round orders the round's stream after
        the work the world's stream holds: on a stream shared with other
        copies, the round waited for those too.)"""
        self.sessions = {s.sid: s for s in sessions}
        self.budget = budget
        self.max_round_cells = max_round_cells
        self.stream = stream
        self.evictor = Evictor(copy_stream, evicting_limit) if copy_stream is not None else None
        self.admitter = Admitter() if copy_stream is not None else None
        self.resident_bytes = 0
        self.batch = None
        self.round_no = 0
        self.evictions = 0
        self.admissions = 0
        self.particle_steps = 0
        # a round: (start, end, worlds, cells, largest world's cells, cold: a
        # world made resident for it or on its first step)
        self.rounds = []

    def _reserve(self, s, busy):
        """Room in the budget for session s's world, evicting idle worlds (not
        in this round, with no request waiting, not on their way in or out)
        LRU first.  False: no room."""
        if s.bytes > self.budget:
            raise ValueError(f"session {s.sid}: a {s.width}x{s.height} world does not fit the budget")
        if self.resident_bytes + s.bytes > self.budget:
            idle = sorted((x for x in self.sessions.values()
                           if x.world is not None and x.evicting is None and x.admitting is None and x.sid not in busy),
                          key=lambda x: x.last_used)
            for x in idle:
                if self.resident_bytes + s.bytes <= self.budget:
                    break
                if self.evictor is not None:
                    self.evictor.submit(x)
                else:
                    x.evict()
                self.resident_bytes -= x.bytes
                self.evictions += 1
            if self.resident_bytes + s.bytes > self.budget:
                return False
        self.resident_bytes += s.bytes
        self.admissions += 1
        return True

    def _admit(self, s, busy):
        """Makes session s resident within the budget, in the loop."""
        if self.evictor is not None:
            self.evictor.wait(s)
        if s.world is not None:
            return True
        if not self._reserve(s, busy):
            return False
        s.make_resident(self.stream)
        return True

    def _ready(self, s, busy):
        """With the threads: whether s's world is resident and settled (its
        eviction or admission done); if it is not and none is on its way,
        its admission is started when the budget has room."""
        if s.evicting is not None and s.evicting.is_set():
            self.evictor.wait(s)
        if s.admitting is not None and s.admitting.is_set():
            s.admitting = None
            self.admitter.wait_any(0.0)          # (an admission's error, raised)
            s.world.set_stream(self.stream)
        if s.world is not None and s.evicting is None and s.admitting is None:
            return True
        if s.admitting is None and self._reserve(s, busy):
            self.admitter.submit(s)
        return False

    def run(self, arrivals, clock=time.perf_counter, on_round=None):
        """Serves `arrivals` [(time offset in s, session id, steps)], sorted
        by time, against the wall clock from now.  Returns the requests."""
        reqs = [Request(sid, steps, t) for t, sid, steps in arrivals]
        t0 = clock()
        nxt = 0
        queue = collections.deque()         # arrived, not done, in arrival order
        while nxt < len(reqs) or queue:
            now = clock() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                queue.append(reqs[nxt])
                nxt += 1
            if not queue:
                time.sleep(max(0.0, min(reqs[nxt].arrival - now, 0.002)))
                continue
            start = clock() - t0
            # this round: the queue's requests in order, one per session,
            # while their worlds fit the round's cells and the budget
            chosen, busy, cells = [], set(), 0
            admitted = self.admissions
            waiting = {r.sid for r in queue}
            for r in queue:
                if r.sid in busy:
                    continue
                s = self.sessions[r.sid]
                c = s.width * s.height
                if chosen and cells + c > self.max_round_cells:
                    continue
                protect = busy | (waiting if s.world is None or s.evicting is not None else set())
                if self.admitter is not None:
                    if not self._ready(s, protect):
                        continue
                elif not self._admit(s, protect):
                    continue
                chosen.append(r)
                busy.add(r.sid)
                cells += c
            if not chosen:
                if self.admitter is not None and self.admitter.inflight > 0:
                    self.admitter.wait_any(0.002)    # (worlds on their way in)
                    continue
                raise RuntimeError("no waiting request fits the HBM budget")
            sess = [self.sessions[r.sid] for r in chosen]
            # (with the threads, admissions are not the round's)
            cold = (self.admitter is None and self.admissions != admitted) or any(s.steps_run == 0 for s in sess)
            worlds = [s.world for s in sess]
            if self.batch is None:
                self.batch = GpuBatch(worlds)
                if self.stream is not None:
                    self.batch.set_stream(self.stream)
            else:
                self.batch.assign(worlds)
            BatchHost([s.sim for s in sess]).step(self.batch)
            end = clock() - t0
            self.round_no += 1
            for s in sess:
                s.last_used = self.round_no
                s.steps_run += 1
                self.particle_steps += int(s.sim.last_stats[abi.STAT_PARTICLES])
            self.rounds.append((start, end, len(chosen), cells, max(s.width * s.height for s in sess), cold))
            for r in chosen:
                r.left -= 1
                if r.left == 0:
                    r.done_at = end
            queue = collections.deque(r for r in queue if r.left > 0)
            if on_round:
                on_round(self)
        if self.evictor is not None:
            self.evictor.drain()     # (the evictions in flight counted in the run's time)
        self.elapsed = clock() - t0
        return reqs

    def close(self):
        if self.admitter is not None:
            self.admitter.close()
            self.admitter = None
        if self.evictor is not None:
            self.evictor.close()
            self.evictor = None
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for s in self.sessions.values():
            if s.world is not None:
                s.world.close()
                s.world = None

