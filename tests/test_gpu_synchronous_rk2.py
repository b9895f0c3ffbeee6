"""SynchronousRK2 with the SPH package (Integrator/SynchronousRK2.cc:40-135) on the GPU against
the oracle: start-of-step derivatives, the trial half step with the full policies (k_half_update),
the midpoint grad-h Omega and derivatives, the full step from the start state.  Same bars as the
CheapSynchronousRK2 step tests: dt within 1e-10, state within 1e-9 of the field's scale, total
energy within 1e-12."""
import numpy as np
import pytest

from spheral_amd import abi, problems

pytestmark = pytest.mark.gpu


def cases():
    out = []
    cfg, st = problems.sod_1d(400)
    out.append(("sod1d_reflecting", cfg, st, 5, False))
    cfg, st = problems.sod_1d(400)
    cfg.gradh_correction = 0
    out.append(("sod1d_gradh_off", cfg, st, 5, False))
    cfg, st = problems.kelvin_helmholtz_2d(3000)
    cfg.precision = abi.SPX_PREC_FP64
    out.append(("kelvin_helmholtz_2d_periodic", cfg, st, 5, False))
    cfg, st = problems.taylor_green_3d(16)
    out.append(("taylor_green_16^3_periodic", cfg, st, 5, False))
    cfg, st = problems.noh_planar_1d(100)
    out.append(("noh_planar_limited_q_nbspline", cfg, st, 10, True))
    cfg, st = problems.noh_planar_1d(100)
    cfg.gradh_correction = 0
    out.append(("noh_planar_limited_q_gradh_off", cfg, st, 10, True))
    cfg, st = problems.evrard_3d(3000, precision=abi.SPX_PREC_FP64)
    out.append(("evrard_3000_sph_tree_gravity", cfg, st, 5, False))
    for c in out:
        c[1].integrator = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
    return out


CASES = cases()


@pytest.mark.parametrize("name,cfg,st,steps,iterate", CASES, ids=[c[0] for c in CASES])
def test_synchronous_rk2_steps_track_oracle(gpu, oracle_mod, name, cfg, st, steps, iterate):
    print(f"[case] {name}", flush=True)
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    if iterate:
        assert g.iterate_ideal_h() == o.iterate_ideal_h()
    for _ in range(steps):
        dg = g.step()
        do = o.step()
        assert abs(dg - do) <= 1e-10*do
    cs = o.field("sound speed").max()
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H", "grad h corrections"):
        a, b = g.field(f), o.field(f)
        sc = max(np.abs(b).max(), cs if f == "velocity" else 0.0)
        assert np.all(np.abs(a - b) <= 1e-9*sc), f
    cg, co = g.conservation(), o.conservation()
    E = abs(co[4] + co[5]) + abs(co[6])
    assert abs((cg[4] + cg[5] + cg[6]) - (co[4] + co[5] + co[6])) <= 1e-12*E


def test_synchronous_rk2_differs_from_cheap(gpu):
    """The two integrators are different schemes: a Sod step under each gives different states
    (the midpoint derivatives come from a fresh evaluation, not the previous step's)."""
    out = []
    for integ in (abi.SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2, abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2):
        cfg, st = problems.sod_1d(400)
        cfg.integrator = integ
        g = gpu.make_run(cfg, st)
        g.initialize()
        for _ in range(3):
            g.step()
        out.append(g.field("specific thermal energy"))
    assert np.abs(out[0] - out[1]).max() > 1e-8
