"""The ensemble and serving workloads' problems against the FP64 oracle, through the C ABI.

* The served problem kinds (2-D and 3-D Sod tubes with reflecting walls, the periodic Gresho
  vortex; tests/functional/Hydro/Sod, tests/functional/Hydro/GreshoVortex/GreshoVortex.py), each
  as a plain FP64 run: every step's dt within 1e-10 of the oracle's, the state after 8 steps
  within 1e-9 of the state's scale or twice the problem's one-ulp roundoff floor, whichever is
  larger (the floor decides only the vortex's specific energy: see roundoff_floor).
* A batch of 8 Kelvin-Helmholtz runs (tests/functional/Hydro/KelvinHelmholtz/
  KelvinHelmholtz-2d.py) of 4k-16k particles: in FP64 every run tracks its oracle run at the
  same bars; in mixed precision the density and specific-energy L1 against the FP64 oracle stay
  within 1e-5 after 20 steps.
* The full-size ensemble: 256 runs of 4k-256k particles (log-uniform, seed 7) to t = 0.1 in one
  mixed-precision batch -- every run's total energy conserved to 1e-10 relative, a sample of runs
  bit-equal to the same runs alone.  (Measured: at most 1.4e-11 over 124-422 steps; the same
  runs in plain FP64 drift up to 9e-12, most of it in the first step, so the figure is the FP64
  accumulation's roundoff, not the FP32 pair terms.)
"""
import numpy as np
import pytest

from spheral_amd import abi, problems

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "mass density", "specific thermal energy", "H")


def oracle_run(oracle_mod, cfg, st):
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    return o


def assert_state_close(get_g, o, tol=1e-9, floor=None):
    """Every field within tol of its scale, or within twice the problem's own roundoff floor
    (floor[f]: how far the oracle moves when the initial positions move by one ulp) where that
    is larger."""
    cs = o.field("sound speed").max()
    for f in FIELDS:
        a, b = get_g(f), o.field(f)
        sc = max(np.abs(b).max(), cs if f == "velocity" else 0.0)
        bar = max(tol*sc, 2.0*floor[f]) if floor else tol*sc
        assert np.all(np.abs(a - b) <= bar), (f, float(np.abs(a - b).max()), bar)


def roundoff_floor(oracle_mod, cfg, st, steps):
    """Largest change of each field after `steps` oracle steps when every initial position moves
    by one ulp.  The compatible-energy partition weighs a pair by max(eps, DepsDt sgn du): where
    DepsDt is zero up to roundoff (the Gresho vortex's rigidly rotating core, div v = 0, on the
    first step) the weights, and so the specific energies, follow the roundoff of DepsDt.  Any two
    summation orders -- the engine's and the oracle's, or the oracle on a one-ulp move -- differ
    there by the same amount (3e-8 of an energy of 8.7 on the 4000-node vortex)."""
    o, p = oracle_run(oracle_mod, cfg, st), oracle_run(oracle_mod, cfg, dict(st, x=np.nextafter(st["x"], np.inf)))
    for _ in range(steps):
        o.step()
        p.step()
    return {f: float(np.abs(o.field(f) - p.field(f)).max()) for f in FIELDS}


SERVED = [
    ("sod_2d", lambda: problems.sod_tube(2, 3000, precision=abi.SPX_PREC_FP64)),
    ("sod_3d", lambda: problems.sod_tube(3, 6000, precision=abi.SPX_PREC_FP64)),
    ("gresho_2d", lambda: problems.gresho_2d(4000, precision=abi.SPX_PREC_FP64)),
]


@pytest.mark.parametrize("name,make", SERVED, ids=[s[0] for s in SERVED])
def test_served_problem_kinds_track_oracle(gpu, oracle_mod, name, make):
    print(f"[case] {name}", flush=True)
    cfg, st = make()
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_run(oracle_mod, cfg, st)
    for k in range(8):
        dg, do = g.step(), o.step()
        assert abs(dg - do) <= 1e-10*do, k
    assert_state_close(g.field, o, floor=roundoff_floor(oracle_mod, cfg, st, 8))
    cg, co = g.conservation(), o.conservation()
    assert abs((cg[4] + cg[5]) - (co[4] + co[5])) <= 1e-12*abs(co[4] + co[5])


def kh_sizes(k, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.round(np.exp(rng.uniform(np.log(4096.0), np.log(16384.0), k))).astype(int)


@pytest.mark.parametrize("precision", [abi.SPX_PREC_FP64, abi.SPX_PREC_MIXED], ids=["fp64", "mixed"])
def test_kelvin_helmholtz_batch_vs_oracle(gpu, oracle_mod, precision):
    sizes = kh_sizes(8)
    steps = 10 if precision == abi.SPX_PREC_FP64 else 20
    print(f"[case] kh batch precision={precision} sizes={sizes.tolist()} steps={steps}", flush=True)
    cfg, _ = problems.kelvin_helmholtz_2d(int(sizes[0]), precision=precision)
    states = [problems.kelvin_helmholtz_2d(int(n), precision=precision)[1] for n in sizes]
    b = gpu.Batch(cfg, capacity=sum(len(s["m"]) for s in states), max_runs=len(states))
    ids = [b.add_run(s, 10.0) for s in states]
    for _ in range(steps):
        assert b.step() == len(states)
    cfg64, _ = problems.kelvin_helmholtz_2d(int(sizes[0]), precision=abi.SPX_PREC_FP64)
    for rid, st in zip(ids, states):
        o = oracle_run(oracle_mod, cfg64, st)
        for _ in range(steps):
            o.step()
        s = b.status(rid)
        assert s["cycle"] == steps
        t_o = o.time()[0]
        if precision == abi.SPX_PREC_FP64:
            assert abs(s["t"] - t_o) <= 1e-10*t_o
            assert_state_close(lambda f: b.field(rid, f), o)
        else:
            assert abs(s["t"] - t_o) <= 1e-4*t_o
            for f in ("mass density", "specific thermal energy"):
                a, r = b.field(rid, f), o.field(f)
                l1 = np.abs(a - r).sum()/np.abs(r).sum()
                print(f"  run {rid} n={len(st['m'])} {f}: L1 {l1:.3e}", flush=True)
                assert l1 <= 1e-5, (rid, f)


def test_full_size_ensemble_properties(gpu):
    """256 runs, N log-uniform in [4096, 262144] (PCG64 seed 7), to t = 0.1 in one mixed batch."""
    rng = np.random.Generator(np.random.PCG64(7))
    targets = np.round(np.exp(rng.uniform(np.log(4096.0), np.log(262144.0), 256))).astype(int)
    cfg, _ = problems.kelvin_helmholtz_2d(int(targets[0]), precision=abi.SPX_PREC_MIXED)
    states = [problems.kelvin_helmholtz_2d(int(n))[1] for n in targets]
    total = sum(len(s["m"]) for s in states)
    print(f"[case] ensemble 256 runs, {total} particles", flush=True)
    b = gpu.Batch(cfg, capacity=total, max_runs=256)
    ids = [b.add_run(s, 0.1) for s in states]
    e0 = {}
    steps = 0
    while True:
        if steps == 0:
            for rid in ids:
                c = b.conservation(rid)
                e0[rid] = c[4] + c[5]
        if b.step() == 0:
            break
        steps += 1
        assert steps < 5000
    print(f"  {steps} batch steps", flush=True)
    for rid in ids:
        s = b.status(rid)
        assert s["state"] == gpu.RUN_DONE and s["t"] >= 0.1*(1.0 - 1e-10)
        c = b.conservation(rid)
        assert abs((c[4] + c[5]) - e0[rid]) <= 1e-10*abs(e0[rid]), rid
    # a sample of runs, smallest to mid-size, equal bit for bit to the same run alone
    order = np.argsort(targets)
    for k in (order[0], order[64], order[128]):
        rid, st = ids[k], states[k]
        r = gpu.make_run(cfg, st)
        r.initialize()
        cyc, t = r.advance(0.1, 100000)
        s = b.status(rid)
        assert s["cycle"] == cyc and s["t"] == t
        for f in FIELDS:
            assert np.array_equal(b.field(rid, f), r.field(f)), (rid, f)
        r.close()
    b.close()
