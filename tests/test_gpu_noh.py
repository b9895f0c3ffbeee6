"""The planar Noh golden and the Limited Q on the GPU, through the C ABI.

The engine runs the reference's Noh-planar-1d.py set-up (NBSpline(5), LimitedMonaghanGingold Q,
reflecting wall, start-up iterateIdealH, CheapSynchronousRK2 to t = 0.6) and must land inside the
reference's error-norm bounds and energy bound, as the oracle does; its steps track the oracle's.
The Limited Q's pair term is compared with the oracle's in 2-D and 3-D, with periodic and
reflected images, grad-h on and off."""
import numpy as np
import pytest

from spheral_amd import abi, problems

import noh_planar as NP

pytestmark = pytest.mark.gpu

TOL = 1e-10


def noh_pair(gpu, oracle_mod):
    cfg, st = problems.noh_planar_1d()
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    return g, o


def noh_norms(run):
    ref = NP.reference_norms()
    return NP.norms(run.field("position")[:, 0], run.field("mass density"), run.field("pressure"),
                    run.field("velocity")[:, 0], run.field("specific thermal energy"),
                    1.0 / run.field("H")[:, 0, 0], run.time()[0], 1.35 * 0.01, ref["rmin"], ref["rmax"])


def test_iterate_ideal_h_matches_oracle(gpu, oracle_mod):
    print("[case] noh iterateIdealH", flush=True)
    g, o = noh_pair(gpu, oracle_mod)
    pg = g.iterate_ideal_h(50, 1.0e-4)
    po = o.iterate_ideal_h(50, 1.0e-4)
    assert pg == po
    a, b = g.field("H"), o.field("H")
    assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b))


def test_noh_first_steps_track_oracle(gpu, oracle_mod):
    print("[case] noh 40 steps", flush=True)
    g, o = noh_pair(gpu, oracle_mod)
    g.iterate_ideal_h()
    o.iterate_ideal_h()
    for k in range(40):
        dg = g.step(1.0e40)
        do = o.step(1.0e40)
        assert abs(dg - do) <= 1e-10 * do, k
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9 * max(np.abs(b).max(), 1.0)), f


def test_noh_planar_golden_norms(gpu, oracle_mod):
    """The reference's acceptance test, run by the engine: all fifteen L1/L2/Linf norms within
    np.allclose(.., tol, tol) of LnormRef['SPH'] and |dE/E| <= 1e-13; the end state is the oracle's
    to the accumulated round-off of the two summation orders."""
    ref = NP.reference_norms()
    print("[case] noh to t=0.6", flush=True)
    g, o = noh_pair(gpu, oracle_mod)
    g.iterate_ideal_h()
    o.iterate_ideal_h()
    E0 = g.conservation()
    ng = NP.drive(g, ref["goal_time"])
    no = NP.drive(o, ref["goal_time"])
    E1 = g.conservation()
    ours = noh_norms(g)
    print("[noh] steps", ng, no, {k: v["L1"] for k, v in ours.items()}, flush=True)
    assert ng == no
    assert g.time()[0] == pytest.approx(ref["goal_time"], abs=1e-15)
    assert NP.within(ours, ref["norms"], ref["tol"]) == []
    e0, e1 = E0[4] + E0[5], E1[4] + E1[5]
    assert abs((e1 - e0) / e0) <= ref["energy_tol"]
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H"):
        a, b = g.field(f), o.field(f)
        d = np.abs(a - b).max() / max(np.abs(b).max(), 1.0)
        print(f"[noh] {f}: max |gpu - oracle| / scale = {d:.3e}", flush=True)
        assert d <= 1e-6, f


def lq_cases():
    out = []
    cfg, st = problems.kelvin_helmholtz_2d(3000)
    cfg.precision = abi.SPX_PREC_FP64
    out.append(("kelvin_helmholtz_2d_periodic", cfg, st))
    cfg, st = problems.uniform_lattice(2, 24, kernel=abi.SPX_KERNEL_BSPLINE)
    cfg.bc_lo[0] = cfg.bc_hi[0] = abi.SPX_BC_REFLECTING
    st["v"] = np.random.default_rng(9).normal(0.0, 0.5, st["x"].shape)
    out.append(("lattice_2d_reflecting_x_periodic_y", cfg, st))
    cfg, st = problems.taylor_green_3d(12)
    cfg.bc_lo[2] = cfg.bc_hi[2] = abi.SPX_BC_REFLECTING
    st["v"] = st["v"] + np.random.default_rng(4).normal(0.0, 0.05, st["v"].shape)
    out.append(("taylor_green_12^3_reflecting_z", cfg, st))
    return out


LQ = lq_cases()


@pytest.mark.parametrize("gradh", [1, 0])
@pytest.mark.parametrize("name,cfg,st", LQ, ids=[c[0] for c in LQ])
def test_limited_q_derivatives_match_oracle(gpu, oracle_mod, name, cfg, st, gradh):
    """Two evaluations with the Q's gradient copied in between: the second one limits every
    pair with nonzero gradients, images included (reflected tensors with grad-h, zero without)."""
    print(f"[case] {name} limited Q gradh={gradh}", flush=True)
    cfg = cfg.copy()
    cfg.viscosity = abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD
    cfg.gradh_correction = gradh
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for r in (g, o):
        r.evaluate_derivatives()
        r.post_state_update()
        r.evaluate_derivatives()
    S = o.field("abs delta velocity hydro")
    assert np.all(np.abs(g.field("delta velocity hydro") - o.field("delta velocity hydro")) <= TOL * S + 1e-300)
    S = o.field("abs delta specific thermal energy")
    assert np.all(np.abs(g.field("delta specific thermal energy") - o.field("delta specific thermal energy"))
                  <= TOL * S + 1e-300)
    a, b = g.field("max viscous pressure"), o.field("max viscous pressure")
    assert np.all(np.abs(a - b) <= TOL * np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300))


@pytest.mark.parametrize("gradh", [1, 0])
@pytest.mark.parametrize("name,cfg,st", LQ[1:2], ids=[LQ[1][0]])
def test_limited_q_steps_track_oracle(gpu, oracle_mod, name, cfg, st, gradh):
    """The step path keeps the Q's gradient by swapping buffers; five steps against the oracle
    (with grad-h off the midpoint's image gradients are zero, as the integrator leaves them)."""
    print(f"[case] {name} limited Q steps gradh={gradh}", flush=True)
    cfg = cfg.copy()
    cfg.viscosity = abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD
    cfg.gradh_correction = gradh
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for _ in range(5):
        dg = g.step()
        do = o.step()
        assert abs(dg - do) <= 1e-10 * do
    cs = o.field("sound speed").max()
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H"):
        a, b = g.field(f), o.field(f)
        sc = max(np.abs(b).max(), cs if f == "velocity" else 0.0)
        assert np.all(np.abs(a - b) <= 1e-9 * sc), f


def test_limited_q_gradient_after_steps_is_the_references(gpu, oracle_mod):
    """After a step the Q holds that step's DvDx (the end-of-step postStateUpdate copy): an
    evaluation through the package calls right after stepping limits with it, on both sides."""
    name, cfg, st = LQ[1]
    print(f"[case] {name} limited Q: steps then evaluate", flush=True)
    cfg = cfg.copy()
    cfg.viscosity = abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for _ in range(3):
        g.step()
        o.step()
    g.update_connectivity()
    o.update_connectivity()
    g.evaluate_derivatives()
    o.evaluate_derivatives()
    S = o.field("abs delta velocity hydro")
    assert np.all(np.abs(g.field("delta velocity hydro") - o.field("delta velocity hydro")) <= 1e-9 * S + 1e-300)
    a, b = g.field("max viscous pressure"), o.field("max viscous pressure")
    assert np.all(np.abs(a - b) <= 1e-9 * np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300))
