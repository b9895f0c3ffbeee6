"""Continuous batching of solve requests on one GPU.

Each problem type (one engine configuration: dimension, kernel, EOS, boundaries) owns one
resident batch (spx_batch) and one host worker thread.  A lane is given an HBM budget in
bytes: it measures the device bytes per particle of its configuration on a probe run (state,
work arrays, neighbour and candidate lists, all of which the engine sizes by the batch's
particle capacity), and sizes its batch to the capacity that fits the budget with HEADROOM for
the lists' growth beyond the probe's.  A worker
admits queued requests while that capacity and the run slots allow, steps every active run of
its batch once, hands back the state of runs that reached their end time and frees their slots.
Requests that arrive while others are on the GPU join the next step; batches of different
types step concurrently (each on its own HIP stream; the engine's calls release the GIL).

Time-to-solution of a request runs from its submission to the moment its result is on the
host.  A run's result does not depend on what else shares its batch (the engine evolves
every run of a batch exactly as it would alone).
"""
import collections
import threading
import time

import numpy as np

from . import abi, engine

RESULT_FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H")


class Request:
    def __init__(self, kind, state, t_end, arrival=None):
        self.kind = kind
        self.state = state
        self.t_end = float(t_end)
        self.n = len(state["m"])
        self.arrival = arrival
        self.admitted = None
        self.done = None
        self.steps = 0
        self.result = None
        self.error = None
        self.event = threading.Event()

    @property
    def time_to_solution(self):
        return None if self.done is None else self.done - self.arrival


PROBE_STEPS = 12
PROBE_MAX_STEPS = 200000
# A lane's batch holds budget / (bytes per particle x HEADROOM) particles.  The probe measures one
# typical request, but a batch's neighbour and candidate lists are sized by the most crowded node
# over all of its runs and grow in steps of 1.25x, so a batch serving many runs reaches a growth
# step or so beyond the probe: on the 500-request Sod 2-D / Sod 3-D / Gresho trace the lanes'
# peaks were 1.21x, 1.19x and 1.36x the probes' bytes per particle.
HEADROOM = 1.5


def bytes_per_particle(cfg, probe_state, t_end=None, device=0):
    """Device bytes per particle of a batch of this configuration, measured on a probe batch sized
    to the probe state: the largest device footprint over the probe run divided by its particle
    count.  The neighbour and candidate lists are sized by the most crowded node of the batch, which
    changes as the flow evolves (a shock compresses, a vortex shears), so with `t_end` the probe
    runs to that time -- a typical request of the kind, end to end.  Without it the probe takes
    PROBE_STEPS steps: past the first steps (which size the lists), the group-size measurement and
    into graph-replayed steps (whose neighbour lists are sized to the candidate lists' capacity)."""
    n = len(probe_state["m"])
    b = engine.Batch(cfg, n, 1, device)
    try:
        rid = b.add_run(probe_state, 1.0e30 if t_end is None else float(t_end))
        peak = b.device_bytes()
        for k in range(PROBE_STEPS if t_end is None else PROBE_MAX_STEPS):
            b.step()
            peak = max(peak, b.device_bytes())
            if t_end is not None and b.status(rid)["state"] == engine.RUN_DONE:
                break
        else:
            if t_end is not None:
                raise RuntimeError(f"probe run did not reach t = {t_end} in {PROBE_MAX_STEPS} steps")
        return peak/n
    finally:
        b.close()


class _Lane:
    """One problem type: a batch sized by an HBM budget, its queue and its worker."""

    def __init__(self, kind, cfg, hbm_bytes, max_runs, probe_state, probe_t_end=None, device=0):
        self.kind = kind
        self.hbm_bytes = int(hbm_bytes)
        self.bytes_per_particle = bytes_per_particle(cfg, probe_state, probe_t_end, device)
        self.capacity = int(self.hbm_bytes/(self.bytes_per_particle*HEADROOM))
        if self.capacity < 1:
            raise ValueError(f"{kind}: an HBM budget of {hbm_bytes} bytes holds no particle")
        self.max_runs = int(max_runs)
        self.batch = engine.Batch(cfg, self.capacity, self.max_runs, device)
        self.peak_bytes = self.batch.device_bytes()
        self.queue = collections.deque()
        self.active = {}
        self.cv = threading.Condition()
        self.closed = False
        self.steps = 0
        self.thread = threading.Thread(target=self._loop, name=f"spx-serve-{kind}", daemon=True)

    def submit(self, req):
        if req.n > self.capacity:
            raise ValueError(f"request of {req.n} particles exceeds the {self.kind} batch capacity {self.capacity} "
                             f"({self.hbm_bytes} bytes of HBM)")
        with self.cv:
            self.queue.append(req)
            self.cv.notify()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify()

    def _admit(self):
        while self.queue:
            req = self.queue[0]
            used, slots = self.batch.usage()
            if slots >= self.max_runs or used + req.n > self.capacity:
                return
            try:
                rid = self.batch.add_run(req.state, req.t_end)
            except engine.SpxError as e:
                # an empty batch admits any request within its capacity: failing then is an error,
                # not a reason to wait (nothing on the GPU would free room)
                if e.code == abi.SPX_ERR_CAPACITY and self.active:
                    return
                raise
            self.queue.popleft()
            req.admitted = time.perf_counter()
            self.active[rid] = req

    def _loop(self):
        try:
            while True:
                with self.cv:
                    while not self.queue and not self.active and not self.closed:
                        self.cv.wait()
                    if self.closed and not self.queue and not self.active:
                        return
                    self._admit()
                if self.active:
                    self.batch.step()
                    self.steps += 1
                    self.peak_bytes = max(self.peak_bytes, self.batch.device_bytes())
                for rid in list(self.active):
                    st = self.batch.status(rid)
                    if st["state"] != engine.RUN_DONE:
                        continue
                    req = self.active.pop(rid)
                    req.result = {f: self.batch.field(rid, f) for f in RESULT_FIELDS}
                    req.steps = st["cycle"]
                    self.batch.remove_run(rid)
                    req.done = time.perf_counter()
                    req.event.set()
        except Exception as e:        # fail every waiting request loudly
            with self.cv:
                pending = list(self.queue) + list(self.active.values())
                self.queue.clear()
                self.active.clear()
            for req in pending:
                req.error = e
                req.event.set()
            raise


class Server:
    """Continuous-batching front end: one lane per problem type."""

    def __init__(self, lanes, device=0):
        """lanes: {kind: (cfg, HBM budget in bytes, run slots, probe state[, probe end time])}; the
        probe (a typical request of the kind, run to its end time when one is given) measures the
        configuration's bytes per particle."""
        self.lanes = {k: _Lane(k, spec[0], spec[1], spec[2], spec[3], spec[4] if len(spec) > 4 else None, device)
                      for k, spec in lanes.items()}
        for lane in self.lanes.values():
            lane.thread.start()

    def submit(self, kind, state, t_end):
        req = Request(kind, state, t_end, time.perf_counter())
        self.lanes[kind].submit(req)
        return req

    def close(self):
        for lane in self.lanes.values():
            lane.close()
        for lane in self.lanes.values():
            lane.thread.join()
            lane.batch.close()


def solve_alone(cfg, state, t_end):
    """The same request as one plain spx_run (the reference's one-problem-per-process model)."""
    r = engine.make_run(cfg, state)
    r.initialize()
    steps, _ = r.advance(t_end, 10**8)
    out = {f: r.field(f) for f in RESULT_FIELDS}
    r.close()
    return out, steps


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else float("nan")
