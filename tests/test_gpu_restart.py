"""A run started from the reference's own restart file (its Sod test's cycle-157 FlatFileIO dump,
two NodeLists, 1-D) evolves on the GPU as the oracle evolves it, and its state goes back out in the
same layout."""
import io
import os

import numpy as np
import pytest

from spheral_amd import abi, restart

from test_restart import GOLDEN

pytestmark = pytest.mark.gpu


def sod_restart_problem():
    d = restart.read_flat(GOLDEN)
    parts = [restart.nodelist_state(d, nl, 1) for nl in ("nodes1", "nodes2")]
    st = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    # the dump's state fixes gamma = 1.4 (eps = 2.5 at P = rho = 1, 1.795 at P = 0.1795, rho = 0.25)
    # and nPerh = 2.01 (h = 2.01 dx); the tube is [-0.5, 0.5] with reflecting ends
    cfg = abi.default_config(1, kernel=abi.SPX_KERNEL_BSPLINE, nperh=2.01, gamma=1.4)
    cfg.bc_lo[0] = cfg.bc_hi[0] = abi.SPX_BC_REFLECTING
    cfg.xmin[0], cfg.xmax[0] = -0.5, 0.5
    g = "restart_cycle157/integrator/"
    cfg.dt_min, cfg.dt_max, cfg.dt_growth, cfg.last_dt = d[g + "dtMin"], d[g + "dtMax"], d[g + "dtGrowth"], d[g + "lastDt"]
    return cfg, st, [len(p["m"]) for p in parts]


def test_run_from_reference_restart_tracks_oracle_and_writes_back(gpu, oracle_mod):
    print("[case] sod restart cycle 157", flush=True)
    cfg, st, sizes = sod_restart_problem()
    g = gpu.make_run(cfg, st)
    g.initialize()
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    e0 = g.conservation()
    for _ in range(10):
        dg, do = g.step(), o.step()
        assert abs(dg - do) <= 1e-10 * do
    for f in ("position", "velocity", "mass density", "specific thermal energy", "H"):
        a, b = g.field(f), o.field(f)
        assert np.all(np.abs(a - b) <= 1e-9 * max(np.abs(b).max(), 1.0)), f
    e1 = g.conservation()
    assert abs((e1[4] + e1[5]) - (e0[4] + e0[5])) <= 1e-12 * abs(e0[4] + e0[5])
    state = restart.run_state(g)
    buf = io.StringIO()
    lo = 0
    for label, n in zip(("nodes1", "nodes2"), sizes):
        restart.write_nodelist(buf, label, {k: v[lo:lo + n] for k, v in state.items()}, prefix="restart_cycle167/")
        lo += n
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "spx_restart_gpu")
    with open(path, "w") as f:
        f.write(buf.getvalue())
    back = restart.read_flat(path)
    os.remove(path)
    got = {k: np.concatenate([restart.nodelist_state(back, nl, 1)[k] for nl in ("nodes1", "nodes2")]) for k in state}
    for k in state:
        assert np.array_equal(got[k], state[k]), k
