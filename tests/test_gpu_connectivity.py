"""The reference-numbered ConnectivityMap (spheral_amd.connectivity) built from the engine's
neighbour sets equals the one built from the oracle's: same kept ghosts (control, image, creation
order) and the same sorted lists, directly and through the reference-named surface after steps."""
import numpy as np
import pytest

from spheral_amd import abi, connectivity, problems
from test_mirror import sod_script

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a.labels == b.labels and a.sizes == b.sizes
    for k in range(len(a.labels)):
        for key in ("control", "code", "created", "parent"):
            assert np.array_equal(a.ghosts[k][key], b.ghosts[k][key]), key
        for l in range(len(a.labels)):
            for p, q in zip(a._csr[(k, l)], b._csr[(k, l)]):
                assert np.array_equal(p, q)


def _oracle_dump(oracle_mod, cfg, st, nodelists=None, boundaries=None):
    o = oracle_mod.OracleRun(cfg, len(st["m"]))
    o.set_state(**st)
    o.initialize()
    off, nbr, img = o.neighbors()
    return connectivity.connectivity_map(cfg, o.field("position"), o.field("H"), off, nbr, img, nodelists, boundaries)


@pytest.mark.parametrize("name", ["kelvin_helmholtz_2d", "taylor_green_16^3", "sod_two_nodelists"])
def test_engine_dump_equals_oracle_dump(gpu, oracle_mod, name):
    print(f"[case] connectivity dump {name}", flush=True)
    nls = None
    if name == "kelvin_helmholtz_2d":
        cfg, st = problems.kelvin_helmholtz_2d(3000, precision=abi.SPX_PREC_FP64)
    elif name == "taylor_green_16^3":
        cfg, st = problems.taylor_green_3d(16)
    else:
        cfg, st = problems.sod_1d(400)
        nl = int(np.sum(st["x"][:, 0] < 0.0))
        nls = [("nodes1", 0, nl), ("nodes2", nl, 400 - nl)]
    g = gpu.make_run(cfg, st)
    g.initialize()
    off, nbr, img = g.neighbors()
    cm = connectivity.connectivity_map(cfg, g.field("position"), g.field("H"), off, nbr, img, nls)
    co = _oracle_dump(oracle_mod, cfg, st, nls)
    _same(cm, co)
    # ConnectivityMap::nodePairList in hash order: the same pairs, as many as the engine counts
    pl, po = cm.nodePairList(), co.nodePairList()
    assert np.array_equal(pl, po)
    assert len(pl) == g.num_pairs()


def test_script_connectivity_map_after_steps(gpu, oracle_mod):
    print("[case] mirror sod connectivityMap after 5 steps", flush=True)
    cfg, _, integ, lists = sod_script()
    integ.initializeProblemStartup()
    for _ in range(5):
        integ.step(1.0)
    cm = integ.connectivityMap()
    r = integ._run
    st = dict(x=r.field("position"), v=r.field("velocity"), m=r.field("mass"), rho=r.field("mass density"),
              eps=r.field("specific thermal energy"), H=r.field("H"))
    nls = [(nl.name, a, b - a) for nl, a, b in integ._slices]
    _same(cm, _oracle_dump(oracle_mod, cfg, st, nls))
    i = lists[0].numInternalNodes - 1
    own, other = cm.connectivityForNode(lists[0], i)
    assert len(own) > 0 and len(other) > 0 and other[0] == 0
