"""Batched runs through the C ABI (spx_batch_*).

A batch steps many independent runs together, each with its own dt, cell grid and
end time.  The bar is bit-equality: a run evolved inside a batch -- next to runs of
other sizes, runs that finish earlier, runs removed and admitted mid-flight -- gives
exactly the state it gives alone as an spx_run.  Also: the mixed-precision mode
against the FP64 oracle (density and energy L1 <= 1e-5)."""
import numpy as np
import pytest

from spheral_amd import abi, problems

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H")


def alone(gpu, cfg, st, t_end):
    r = gpu.make_run(cfg, st)
    r.initialize()
    k, t = r.advance(t_end, 100000)
    return r, k, t


def assert_same_as_alone(gpu, b, rid, cfg, st, t_end):
    s = b.status(rid)
    assert s["state"] == gpu.RUN_DONE
    r, k, t = alone(gpu, cfg, st, t_end)
    assert s["cycle"] == k and s["t"] == t and s["last_dt"] == r.time()[2]
    for f in FIELDS:
        assert np.array_equal(b.field(rid, f), r.field(f)), f
    assert np.array_equal(b.conservation(rid), r.conservation())


def run_to_completion(b, limit=20000):
    steps = 0
    while b.step() > 0:
        steps += 1
        assert steps < limit
    return steps


@pytest.mark.parametrize("precision", [abi.SPX_PREC_FP64, abi.SPX_PREC_MIXED], ids=["fp64", "mixed"])
def test_batch_2d_runs_equal_alone(gpu, precision):
    sizes, t_ends = [1500, 3000, 2200], [0.02, 0.05, 0.035]
    print(f"[case] batch kh2d precision={precision} sizes={sizes}", flush=True)
    cfg, _ = problems.kelvin_helmholtz_2d(sizes[0], precision=precision)
    states = [problems.kelvin_helmholtz_2d(n, precision=precision)[1] for n in sizes]
    b = gpu.Batch(cfg, capacity=sum(len(s["m"]) for s in states), max_runs=4)
    ids = [b.add_run(s, te) for s, te in zip(states, t_ends)]
    assert b.usage() == (sum(len(s["m"]) for s in states), 3)
    run_to_completion(b)
    for rid, st, te in zip(ids, states, t_ends):
        assert_same_as_alone(gpu, b, rid, cfg, st, te)


def test_finished_runs_retire_in_place(gpu):
    """Runs finishing at different steps: after the first reorder the batch orders its runs by
    estimated remaining steps, so later finishers are the tail of the active range and retire
    without moving the others.  Every run still equals itself alone."""
    sizes, t_ends = [1200, 2500, 1800, 3000, 900, 2100], [0.01, 0.03, 0.02, 0.04, 0.015, 0.025]
    print(f"[case] batch retire in place sizes={sizes}", flush=True)
    cfg, _ = problems.kelvin_helmholtz_2d(sizes[0], precision=abi.SPX_PREC_MIXED)
    states = [problems.kelvin_helmholtz_2d(n, precision=abi.SPX_PREC_MIXED)[1] for n in sizes]
    b = gpu.Batch(cfg, capacity=sum(len(s["m"]) for s in states), max_runs=6)
    ids = [b.add_run(s, te) for s, te in zip(states, t_ends)]
    run_to_completion(b)
    st = b.step_stats()
    print(f"[case] reorders {st['full_reorders']}, in-place retirements {st['in_place_retirements']}", flush=True)
    assert st["in_place_retirements"] >= 1
    for rid, s, te in zip(ids, states, t_ends):
        assert_same_as_alone(gpu, b, rid, cfg, s, te)


def test_batch_3d_runs_equal_alone(gpu):
    print("[case] batch tg3d 16^3 + 12^3", flush=True)
    cfg, s16 = problems.taylor_green_3d(16)
    _, s12 = problems.taylor_green_3d(12)
    b = gpu.Batch(cfg, capacity=16**3 + 12**3, max_runs=2)
    a = b.add_run(s16, 0.004)
    c = b.add_run(s12, 0.002)
    run_to_completion(b)
    assert_same_as_alone(gpu, b, a, cfg, s16, 0.004)
    assert_same_as_alone(gpu, b, c, cfg, s12, 0.002)


def test_batch_admission_removal_and_capacity(gpu):
    print("[case] batch admission/removal", flush=True)
    cfg, sa = problems.kelvin_helmholtz_2d(2000, precision=abi.SPX_PREC_FP64)
    _, sb = problems.kelvin_helmholtz_2d(1200, precision=abi.SPX_PREC_FP64)
    _, sc = problems.kelvin_helmholtz_2d(1100, precision=abi.SPX_PREC_FP64)
    na, nb, nc = len(sa["m"]), len(sb["m"]), len(sc["m"])
    assert nc <= nb
    b = gpu.Batch(cfg, capacity=na + nb, max_runs=2)
    a = b.add_run(sa, 0.04)
    r2 = b.add_run(sb, 0.04)
    with pytest.raises(gpu.SpxError) as e:           # no slot left
        b.add_run(sc, 0.03)
    assert e.value.code == abi.SPX_ERR_CAPACITY
    for _ in range(3):
        b.step()
    b.remove_run(r2)
    c = b.add_run(sc, 0.03)                            # reuses the slot after squeezing run r2 out
    assert c == r2
    assert b.usage() == (na + nc, 2)
    with pytest.raises(gpu.SpxError):
        b.status(5)
    run_to_completion(b)
    assert_same_as_alone(gpu, b, a, cfg, sa, 0.04)
    assert_same_as_alone(gpu, b, c, cfg, sc, 0.03)


def test_mixed_precision_accuracy_vs_fp64_oracle(gpu, oracle_mod):
    """FP32 pair arithmetic on double-differenced offsets with double accumulation: density
    and specific energy stay within L1 1e-5 (relative) of the FP64 oracle after 5 steps."""
    print("[case] mixed tg3d 24^3", flush=True)
    cfg, st = problems.taylor_green_3d(24)
    cfg.precision = abi.SPX_PREC_MIXED
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for _ in range(5):
        dg, do = g.step(), o.step()
        assert abs(dg - do) <= 1e-4*do
    for f in ("mass density", "specific thermal energy"):
        a, r = g.field(f), o.field(f)
        l1 = np.abs(a - r).sum()/np.abs(r).sum()
        print(f"  {f}: L1 {l1:.3e}", flush=True)
        assert l1 <= 1e-5, f
    v, vo = g.field("velocity"), o.field("velocity")
    assert np.abs(v - vo).max() <= 1e-4*np.abs(vo).max()


def test_evrard_mixed_precision_vs_fp64_oracle(gpu, oracle_mod):
    """Hydro + TreeGravity (Evrard collapse, 20k particles): mixed-precision pair math with
    the FP64 tree stays within L1 1e-5 of the FP64 oracle over 5 steps."""
    print("[case] evrard mixed 20000", flush=True)
    cfg, st = problems.evrard_3d(20000, precision=abi.SPX_PREC_MIXED)
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for _ in range(5):
        dg, do = g.step(), o.step()
        assert abs(dg - do) <= 1e-4*do
    for f in ("mass density", "specific thermal energy"):
        a, r = g.field(f), o.field(f)
        l1 = np.abs(a - r).sum()/np.abs(r).sum()
        print(f"  {f}: L1 {l1:.3e}", flush=True)
        assert l1 <= 1e-5, f
