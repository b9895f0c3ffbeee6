"""Graph-replayed steps and the measured neighbour-group sizes, through the C ABI.

* A run whose steady CheapSynchronousRK2 steps are replayed from captured HIP graphs equals the
  same run stepped eagerly, bit for bit, at every check.  The graphs are replayed many times
  between checks, with the program's own eager steps (candidate-list rebuilds, the step after a
  field read) and host synchronisations in between: 3-D periodic FP64, 2-D periodic mixed, and a
  mixed batch whose runs finish at different times (the composition changes under the graphs).
* Fixed group sizes (4, 4, 4) and (8, 8, 8) and the measured choice give the same bits; the
  measurement happens, and bad sizes are refused.
"""
import numpy as np
import pytest

from spheral_amd import abi, engine, problems

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H")


def same(get_a, get_b):
    for f in FIELDS:
        a, b = get_a(f), get_b(f)
        assert np.array_equal(a, b), f


@pytest.mark.parametrize("name", ["taylor_green_16^3_fp64", "kelvin_helmholtz_2d_4k_mixed"])
def test_graph_steps_equal_eager_steps(gpu, name):
    print(f"[case] graph vs eager {name}", flush=True)
    if name.startswith("taylor"):
        cfg, st = problems.taylor_green_3d(16)
    else:
        cfg, st = problems.kelvin_helmholtz_2d(4096, precision=abi.SPX_PREC_MIXED)
    a = gpu.make_run(cfg, st)          # graphs on (the default)
    b = gpu.make_run(cfg, st)
    b.set_step_graphs(False)
    a.initialize()
    b.initialize()
    for block in range(6):
        for _ in range(9):
            da, db = a.step(), b.step()
            assert da == db, block
        same(a.field, b.field)
    sa, sb = a.step_stats(), b.step_stats()
    print(f"  graphs on: {sa}\n  graphs off: {sb}", flush=True)
    assert sa["graph_replays"] >= 30 and 2 <= sa["graph_captures"] <= 12
    assert sb["graph_replays"] == 0 and sb["eager_steps"] == 54
    a.close()
    b.close()


def test_batch_graph_steps_equal_eager_steps(gpu):
    """Six mixed-precision Kelvin-Helmholtz runs with staggered end times: the batch's active set
    shrinks while graphs are being replayed."""
    print("[case] batch graph vs eager", flush=True)
    sizes = [4096, 6000, 9000, 4096, 12000, 5000]
    ends = [0.004, 0.012, 0.02, 0.008, 0.03, 0.016]
    cfg, _ = problems.kelvin_helmholtz_2d(4096, precision=abi.SPX_PREC_MIXED)
    states = [problems.kelvin_helmholtz_2d(n)[1] for n in sizes]
    cap = sum(len(s["m"]) for s in states)
    A = gpu.Batch(cfg, capacity=cap, max_runs=len(states))
    B = gpu.Batch(cfg, capacity=cap, max_runs=len(states))
    B.set_step_graphs(False)
    ia = [A.add_run(s, t) for s, t in zip(states, ends)]
    ib = [B.add_run(s, t) for s, t in zip(states, ends)]
    steps = 0
    while True:
        na, nb = A.step(), B.step()
        assert na == nb
        steps += 1
        if na == 0:
            break
        assert steps < 4000
        if steps % 25 == 0:
            for ra, rb in zip(ia, ib):
                same(lambda f: A.field(ra, f), lambda f: B.field(rb, f))
    for ra, rb in zip(ia, ib):
        assert A.status(ra) == B.status(rb)
        same(lambda f: A.field(ra, f), lambda f: B.field(rb, f))
    sa = A.step_stats()
    print(f"  {steps} batch steps, graphs on: {sa}", flush=True)
    assert sa["graph_replays"] >= steps//3
    A.close()
    B.close()


def test_group_sizes_keep_result_bits(gpu):
    print("[case] neighbour-group sizes", flush=True)
    cfg, st = problems.taylor_green_3d(14)
    runs = []
    for sizes in ((4, 4, 4), (8, 8, 8), None):
        r = gpu.make_run(cfg, st)
        r.set_step_graphs(False)
        if sizes:
            r.set_group_sizes(sizes)
        r.initialize()
        runs.append(r)
    for _ in range(12):
        dts = [r.step() for r in runs]
        assert dts[0] == dts[1] == dts[2]
    for r in runs[1:]:
        same(runs[0].field, r.field)
    fixed, tuned = runs[0].step_stats(), runs[2].step_stats()
    assert fixed["group_tunings"] == 0 and (fixed["group_filter"], fixed["group_gradh"], fixed["group_update"]) == (4, 4, 4)
    assert tuned["group_tunings"] == 1
    assert all(tuned[k] in (4, 8) for k in ("group_filter", "group_gradh", "group_update"))
    with pytest.raises(engine.SpxError):
        runs[0].set_group_sizes((5, 4, 4))
    for r in runs:
        r.close()
