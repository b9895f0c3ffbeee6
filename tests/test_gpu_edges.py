"""Edge cases of the step against the oracle: no nodes (rejected on the host), an empty batch, a
single node (no neighbours: the ideal-H growth, Omega = 1, dt from the node's own sound speed),
isolated nodes, and a blob with far outliers (one cell grid spanning very different h)."""
import numpy as np
import pytest

from spheral_amd import abi, problems

pytestmark = pytest.mark.gpu


def _iso(n, dim, h):
    H = np.zeros((n, dim, dim))
    for a in range(dim):
        H[:, a, a] = 1.0/h
    return H


def _track(gpu, oracle_mod, cfg, st, steps=3):
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    for _ in range(steps):
        dg, do = g.step(), o.step()
        assert abs(dg - do) <= 1e-10*do
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H", "grad h corrections"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9*max(np.abs(b).max(), 1e-300)), f
    return g, o


def test_zero_nodes_rejected(gpu):
    with pytest.raises(gpu.SpxError) as e:
        gpu.Run(abi.default_config(3), 0)
    assert e.value.code == abi.SPX_ERR_CAPACITY


def test_empty_batch_steps_nothing(gpu):
    cfg, _ = problems.kelvin_helmholtz_2d(500, precision=abi.SPX_PREC_FP64)
    b = gpu.Batch(cfg, capacity=1000, max_runs=2)
    assert b.step() == 0
    assert b.usage() == (0, 0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_single_node_tracks_oracle(gpu, oracle_mod, dim):
    print(f"[case] single node {dim}-D", flush=True)
    cfg = abi.default_config(dim)
    st = dict(x=np.zeros((1, dim)), v=np.full((1, dim), 0.3), m=np.ones(1), rho=np.ones(1), eps=np.ones(1),
              H=_iso(1, dim, 0.1))
    g, o = _track(gpu, oracle_mod, cfg, st)
    assert g.field("grad h corrections")[0] == 1.0
    assert g.field("H")[0, 0, 0] < 10.0          # h grew: no neighbours inside the support


def test_isolated_nodes_track_oracle(gpu, oracle_mod):
    print("[case] 64 isolated nodes 3-D", flush=True)
    g1 = np.arange(4)*10.0
    x = np.stack([a.ravel() for a in np.meshgrid(g1, g1, g1, indexing="ij")], 1)
    n = len(x)
    rng = np.random.default_rng(3)
    st = dict(x=x, v=rng.normal(0, 0.1, (n, 3)), m=np.ones(n), rho=np.ones(n), eps=rng.uniform(0.5, 1.5, n),
              H=_iso(n, 3, 0.1))
    g, _ = _track(gpu, oracle_mod, abi.default_config(3), st)
    off, nbr, img = g.neighbors()
    assert off[-1] == 0


def test_blob_with_far_outliers_tracks_oracle(gpu, oracle_mod):
    print("[case] gaussian blob + outliers 3-D", flush=True)
    rng = np.random.default_rng(11)
    n0, no = 2000, 12
    x = np.concatenate([rng.normal(0, 0.3, (n0, 3)), rng.uniform(-5.0, 5.0, (no, 3))])
    h = np.concatenate([np.full(n0, 0.08), np.full(no, 0.5)])
    n = n0 + no
    st = dict(x=x, v=rng.normal(0, 0.1, (n, 3)), m=np.full(n, 1.0/n), rho=np.ones(n), eps=np.full(n, 0.5),
              H=np.einsum("i,ab->iab", 1.0/h, np.eye(3)))
    _track(gpu, oracle_mod, abi.default_config(3, nperh=1.51), st)
