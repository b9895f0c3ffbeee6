"""TreeGravity in 2-D (the reference's QuadTreeGravity, Gravity_PYB11.py:53) on the GPU against the
oracle: a self-gravitating disk (problems.gravity_disk_2d, the Evrard set-up in 2-D).  Same quadtree,
same acceptance decisions: accelerations, potential and the dt vote at roundoff; then steps of the
hydro + gravity problem under CheapSynchronousRK2, and of the gravity-only problem under
SynchronousRK2, track the oracle."""
import numpy as np
import pytest

from spheral_amd import abi, problems

pytestmark = pytest.mark.gpu

TOL = 1e-10


def pair(gpu, oracle_mod, cfg, st):
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    return g, o


@pytest.mark.parametrize("n,opening", [(3000, 0.5), (3000, 0.05), (20000, 0.5)])
def test_quadtree_gravity_matches_oracle(gpu, oracle_mod, n, opening):
    print(f"[case] disk 2-D n={n} opening={opening}", flush=True)
    cfg, st = problems.gravity_disk_2d(n)
    cfg.opening = opening
    g, o = pair(gpu, oracle_mod, cfg, st)
    g.evaluate_derivatives()
    o.evaluate_derivatives()
    ag, ao = g.field("delta velocity"), o.field("delta velocity")
    assert ag.shape == (n, 2)
    assert np.all(np.abs(ag - ao) <= TOL*np.abs(ao).max())
    pg, po = g.field("gravitational potential"), o.field("gravitational potential")
    assert np.all(np.abs(pg - po) <= TOL*np.abs(po).max())
    gh, gg = g.dt()
    oh, og = o.dt()
    assert abs(gg - og) <= TOL*og
    for _ in range(5):
        assert abs(g.step() - o.step()) <= TOL*o.time()[2]
    a, b = g.field("velocity"), o.field("velocity")
    assert np.all(np.abs(a - b) <= 1e-9*max(np.abs(b).max(), 1e-300))
    ca, cb = g.conservation(), o.conservation()
    assert abs(ca[6] - cb[6]) <= 1e-9*abs(cb[6])          # TreeGravity's extra energy


def test_quadtree_gravity_only_synchronous_rk2(gpu, oracle_mod):
    """Gravity alone under SynchronousRK2 (tree and walk at the start and the midpoint of each step)."""
    print("[case] disk 2-D gravity only", flush=True)
    cfg, st = problems.gravity_disk_2d(4000)
    cfg.gravity_only = 1
    cfg.integrator = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
    g, o = pair(gpu, oracle_mod, cfg, st)
    for k in range(10):
        dg, do = g.step(1.0e40), o.step(1.0e40)
        assert abs(dg - do) <= TOL*do, k
    for f in ("position", "velocity"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9*np.abs(b).max()), f
