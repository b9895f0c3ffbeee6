"""The reference's planar Noh golden (tests/functional/Hydro/Noh/Noh-planar-1d.py) pins the oracle's
LimitedMonaghanGingold Q, NBSpline kernel, reflecting wall, start-up H iteration and the
CheapSynchronousRK2 / compatible-energy step end to end; two invariants pin the Limited Q on
their own."""
import numpy as np
import pytest

from spheral_amd import abi, problems

import noh_planar as NP


def run_noh_oracle(oracle_mod, mutate=None):
    cfg, st = problems.noh_planar_1d()
    if mutate:
        mutate(cfg)
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    passes = o.iterate_ideal_h(50, 1.0e-4)
    E0 = o.conservation()
    steps = NP.drive(o, NP.reference_norms()["goal_time"])
    return o, passes, steps, E0, o.conservation()


def norms_of(run, h0=1.35 * 0.01):
    ref = NP.reference_norms()
    t = run.time()[0]
    return NP.norms(run.field("position")[:, 0], run.field("mass density"), run.field("pressure"),
                    run.field("velocity")[:, 0], run.field("specific thermal energy"),
                    1.0 / run.field("H")[:, 0, 0], t, h0, ref["rmin"], ref["rmax"])


def test_noh_planar_matches_the_reference_norms(oracle_mod):
    ref = NP.reference_norms()
    o, passes, steps, E0, E1 = run_noh_oracle(oracle_mod)
    assert o.time()[0] == pytest.approx(ref["goal_time"], abs=1e-15)
    ours = norms_of(o)
    print("[noh] passes", passes, "steps", steps, {k: v["L1"] for k, v in ours.items()})
    assert NP.within(ours, ref["norms"], ref["tol"]) == []
    # total energy (kinetic + thermal) conserved to the script's bound
    e0, e1 = E0[4] + E0[5], E1[4] + E1[5]
    assert abs((e1 - e0) / e0) <= ref["energy_tol"]


def test_noh_norms_reject_the_plain_q(oracle_mod):
    """The golden tells the Q apart: the plain MonaghanGingold Q (same Cl, Cq) misses it."""
    ref = NP.reference_norms()
    o, *_ = run_noh_oracle(oracle_mod, lambda c: setattr(c, "viscosity", abi.SPX_Q_MONAGHAN_GINGOLD))
    assert len(NP.within(norms_of(o), ref["norms"], ref["tol"])) >= 10


def _linear_flow(A, n1d=14, nperh=1.5):
    dx = 1.0 / n1d
    g = (np.arange(n1d) + 0.5) * dx
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel()], 1)
    n = len(x)
    H = np.zeros((n, 2, 2))
    H[:, 0, 0] = H[:, 1, 1] = 1.0 / (nperh * dx)
    st = dict(x=x, v=x @ np.asarray(A).T, m=np.full(n, dx * dx), rho=np.ones(n), eps=np.full(n, 1.0), H=H)
    cfg = abi.default_config(2, kernel=abi.SPX_KERNEL_NBSPLINE, kernel_order=5, nperh=nperh)
    return cfg, st


@pytest.mark.parametrize("A", [[[-1.0, 0.3], [0.2, -0.5]], [[-0.7, 0.0], [0.0, 0.4]]])
def test_limited_q_vanishes_in_a_linear_flow(oracle_mod, A):
    """With the corrected velocity gradient, DvDx of a linear field is the field's matrix at every
    node; the limiter ratio is then one, phi = 1 and the extrapolated midpoint velocities agree,
    so the Limited Q is round-off while the plain Q of the same compressing flow is not."""
    out = {}
    for q in (abi.SPX_Q_MONAGHAN_GINGOLD, abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD):
        cfg, st = _linear_flow(A)
        cfg.viscosity = q
        o = oracle_mod.OracleRun(cfg, len(st["m"]))
        o.set_state(**st)
        o.initialize()
        o.evaluate_derivatives()
        G = o.field("velocity gradient")
        assert np.allclose(G, np.broadcast_to(np.asarray(A), G.shape), atol=1e-12)
        o.post_state_update()              # the Q's copy of DvDx
        o.evaluate_derivatives()
        out[q] = o.field("max viscous pressure")
    mg, lim = out[abi.SPX_Q_MONAGHAN_GINGOLD], out[abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD]
    assert mg.max() > 0.01
    assert lim.max() <= 1e-10 * mg.max()


def test_limited_q_is_the_plain_q_before_any_gradient(oracle_mod):
    """The Q's gradient starts at zero (ArtificialViscosityHandle::initializeProblemStartup), so
    the first evaluation's limiter is zero and the Limited Q equals the plain one bit for bit."""
    out = {}
    for q in (abi.SPX_Q_MONAGHAN_GINGOLD, abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD):
        cfg, st = _linear_flow([[-1.0, 0.3], [0.2, -0.5]])
        cfg.viscosity = q
        o = oracle_mod.OracleRun(cfg, len(st["m"]))
        o.set_state(**st)
        o.initialize()
        o.evaluate_derivatives()
        out[q] = (o.field("max viscous pressure"), o.field("delta velocity hydro"))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_iterate_ideal_h_converges_to_the_package_ideal_h(oracle_mod):
    """After SpheralController's start-up iteration the interior keeps h = nPerh dx (a uniform
    lattice already has the ideal count), the open end grows h, and one more SPHSmoothingScale
    evaluation moves no node by more than a few tolerances."""
    cfg, st = problems.noh_planar_1d()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    passes = o.iterate_ideal_h(50, 1.0e-4)
    assert 1 < passes < 50
    h = 1.0 / o.field("H")[:, 0, 0]
    x = o.field("position")[:, 0]
    interior = (x > 0.1) & (x < 0.9)
    assert np.allclose(h[interior], 0.0135, rtol=2e-4)
    assert h[-1] > 1.5 * 0.0135
    o.evaluate_derivatives()
    h1 = 1.0 / o.field("new H")[:, 0, 0]
    assert np.max(np.abs(h1 / h - 1.0)) < 1e-3
