"""Continuous batching: requests that arrive while others run are admitted into the resident
batches, and every result equals the same request solved alone (bit-identical).  Lanes are sized
by an HBM budget in bytes; the budgets here are small on purpose (a few thousand particles), so
admission has to wait for slots and memory to free up."""
import time

import numpy as np
import pytest

from spheral_amd import abi, problems, serve

pytestmark = pytest.mark.gpu


def budget_for(cfg, probe, particles):
    """The HBM budget (bytes) of a lane that holds `particles` particles of this configuration."""
    return int(serve.bytes_per_particle(cfg, probe)*serve.HEADROOM*(particles + 0.5))


def test_served_results_equal_alone(gpu):
    print("[case] serve mixed trace", flush=True)
    kinds = {"sod2d": lambda n: problems.sod_tube(2, n), "sod3d": lambda n: problems.sod_tube(3, n),
             "gresho2d": lambda n: problems.gresho_2d(n)}
    trace = [("sod2d", 900, 0.01), ("gresho2d", 1000, 0.004), ("sod3d", 1200, 0.006),
             ("sod2d", 1500, 0.008), ("gresho2d", 700, 0.003), ("sod3d", 800, 0.004)]
    cfgs = {k: f(1000)[0] for k, f in kinds.items()}
    probes = {k: f(1000)[1] for k, f in kinds.items()}
    # small budgets so that admission has to wait for slots and memory to free up
    srv = serve.Server({k: (cfgs[k], budget_for(cfgs[k], probes[k], 2600), 2, probes[k]) for k in kinds})
    for k, lane in srv.lanes.items():
        assert lane.capacity == 2600, k
    handles = []
    for k, n, te in trace:
        handles.append((k, n, te, srv.submit(k, kinds[k](n)[1], te)))
        time.sleep(0.02)
    for *_, h in handles:
        assert h.event.wait(120)
        assert h.error is None
    for k, lane in srv.lanes.items():
        assert lane.peak_bytes <= lane.hbm_bytes, k
    srv.close()
    for k, n, te, h in handles:
        ref, steps = serve.solve_alone(cfgs[k], kinds[k](n)[1], te)
        assert h.steps == steps
        for f in serve.RESULT_FIELDS:
            assert np.array_equal(h.result[f], ref[f]), (k, n, f)


@pytest.mark.parametrize("slots,trace", [
    (1, [(900, 0.004), (700, 0.003)]),         # one slot: the second request waits for the first
    (2, [(800, 0.004), (800, 0.004)]),         # identical requests finish on the same step
])
def test_lane_admits_after_every_run_finished(gpu, slots, trace):
    """Every active run of a lane finishing in the same step leaves the batch empty with requests
    queued: the removed runs' slots and particles count as free, so the lane admits the queue
    instead of waiting on a step that never comes."""
    print(f"[case] serve slots={slots} trace={trace}", flush=True)
    cfg, probe = problems.sod_tube(2, 1000)
    srv = serve.Server({"sod2d": (cfg, budget_for(cfg, probe, 2000), slots, probe)})
    assert srv.lanes["sod2d"].capacity == 2000
    handles = [srv.submit("sod2d", problems.sod_tube(2, n)[1], te) for n, te in trace]
    handles.append(srv.submit("sod2d", problems.sod_tube(2, 600)[1], 0.002))
    for h in handles:
        assert h.event.wait(120)
        assert h.error is None
    lane = srv.lanes["sod2d"]
    assert lane.batch.usage() == (0, 0)
    srv.close()
    for (n, te), h in zip(trace + [(600, 0.002)], handles):
        ref, steps = serve.solve_alone(cfg, problems.sod_tube(2, n)[1], te)
        assert h.steps == steps
        assert np.array_equal(h.result["mass density"], ref["mass density"])


def test_probe_to_end_time_bounds_the_lane(gpu):
    """The lists grow as the flow evolves (the shock compresses the gas), so a lane sized from a
    probe run to the request's end time stays within its HBM budget over whole requests, and that
    probe measures at least what a 12-step probe does."""
    print("[case] serve probe to t_end", flush=True)
    cfg, probe = problems.sod_tube(2, 2000)
    short = serve.bytes_per_particle(cfg, probe)
    full = serve.bytes_per_particle(cfg, probe, 0.15)
    assert full >= short
    budget = int(full*serve.HEADROOM*(4000 + 0.5))
    srv = serve.Server({"sod2d": (cfg, budget, 4, probe, 0.15)})
    lane = srv.lanes["sod2d"]
    assert lane.capacity == 4000
    handles = [srv.submit("sod2d", problems.sod_tube(2, 2000)[1], 0.15) for _ in range(2)]
    for h in handles:
        assert h.event.wait(300)
        assert h.error is None
    assert lane.peak_bytes <= lane.hbm_bytes
    srv.close()
