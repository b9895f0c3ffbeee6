"""SynchronousRK2 with TreeGravity alone on the GPU, against the oracle: the reference's collapsing
sphere (4224 lattice points, opening 1) over 40 steps, and a two-body orbit."""
import numpy as np
import pytest

from spheral_amd import problems

from test_oracle_nbody import collapse_fit, two_body

pytestmark = pytest.mark.gpu


def test_collapse_tracks_oracle(gpu, oracle_mod):
    print("[case] collisionless sphere 40 steps", flush=True)
    cfg, st, _ = problems.collisionless_sphere_3d()
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for k in range(40):
        dg = g.step(1.0e40)
        do = o.step(1.0e40)
        assert abs(dg - do) <= 1e-10 * do, k
    for f in ("position", "velocity", "gravitational potential"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9 * np.abs(b).max()), f
    cg, sg = collapse_fit(g)
    co, so = collapse_fit(o)
    print("[collapse] gpu", cg, sg, "oracle", co, so, flush=True)
    assert abs(cg[0] - co[0]) < 1e-8 and abs(sg / so - 1.0) < 1e-6


def test_two_body_orbit_matches_oracle(gpu, oracle_mod):
    print("[case] two-body orbit", flush=True)
    cfg, st, period = two_body()
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, 2)
    o.set_state(**st)
    o.initialize()
    n = 0
    while o.time()[0] < period * (1.0 - 1e-15):
        o.step(period)
        g.step(period)
        n += 1
    assert g.time()[1] == n
    assert np.abs(g.field("position") - o.field("position")).max() < 1e-9


def test_coincident_points_share_a_deepest_leaf(gpu, oracle_mod):
    """Pairs of points at one position: the octree lists both in one cell at its deepest level (the
    walk's multi-point leaf loop, and a monopole pass over every level down to it), and each point
    skips the other's term by the softening guard, as the oracle's walk does."""
    print("[case] collisionless sphere with coincident pairs, 10 steps", flush=True)
    cfg, st, _ = problems.collisionless_sphere_3d()
    n = len(st["m"])
    for k in range(0, n - 1, 97):
        for key in ("x", "v"):
            st[key][k + 1] = st[key][k]
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, n)
    o.set_state(**st)
    o.initialize()
    for k in range(10):
        dg = g.step(1.0e40)
        do = o.step(1.0e40)
        assert abs(dg - do) <= 1e-10 * do, k
    for f in ("position", "velocity", "gravitational potential"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9 * np.abs(b).max()), f
    x = g.field("position")
    assert np.array_equal(x[1:n:97][:len(range(0, n - 1, 97))], x[0:n - 1:97])
