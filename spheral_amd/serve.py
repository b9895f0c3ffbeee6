"""Continuous batching of solve requests on one GPU.

Each problem type (one engine configuration: dimension, kernel, EOS, boundaries) owns one
resident batch (spx_batch) and one host worker thread.  A worker admits queued requests
while the batch's particle budget and run slots allow, steps every active run of its batch
once, hands back the state of runs that reached their end time and frees their slots.
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


class _Lane:
    """One problem type: a batch, its queue and its worker."""

    def __init__(self, kind, cfg, capacity, max_runs, device=0):
        self.kind = kind
        self.capacity = int(capacity)
        self.max_runs = int(max_runs)
        self.batch = engine.Batch(cfg, self.capacity, self.max_runs, device)
        self.queue = collections.deque()
        self.active = {}
        self.cv = threading.Condition()
        self.closed = False
        self.steps = 0
        self.thread = threading.Thread(target=self._loop, name=f"spx-serve-{kind}", daemon=True)

    def submit(self, req):
        if req.n > self.capacity:
            raise ValueError(f"request of {req.n} particles exceeds the {self.kind} batch capacity {self.capacity}")
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
        """lanes: {kind: (cfg, particle capacity, run slots)}."""
        self.lanes = {k: _Lane(k, cfg, cap, slots, device) for k, (cfg, cap, slots) in lanes.items()}
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
