"""Restart state exchange in the reference's FlatFileIO layout (spheral_amd/restart.py), checked on
the restart file the reference's own Sod test directory holds
(tests/functional/Hydro/Sod/Results/Sod-planar-2nodelists-1d-restart_cycle157, copied as data)."""
import io
import os

import numpy as np

from spheral_amd import restart

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "Sod-planar-2nodelists-1d-restart_cycle157")


def test_reads_the_reference_sod_restart():
    d = restart.read_flat(GOLDEN)
    a = restart.nodelist_state(d, "nodes1", 1)
    b = restart.nodelist_state(d, "nodes2", 1)
    assert len(a["m"]) == 400 and len(b["m"]) == 100
    assert a["x"].shape == (400, 1) and a["H"].shape == (400, 1, 1)
    assert a["x"][0, 0] == -0.499375 and a["m"][0] == 0.00125 and a["H"][0, 0, 0] == 398.01
    assert b["x"][0, 0] == 0.102009 and b["v"][0, 0] == 0.673559 and b["rho"][0] == 0.4679
    assert d["restart_cycle157/integrator/currentTime"] == 0.15
    assert d["restart_cycle157/integrator/lastDt"] == 0.000903564
    assert d["restart_cycle157/control/self/totalSteps"] == 157
    # equal-mass nodes, left state still at rest, right-moving rarefied gas on the right
    assert np.all(a["m"] == 0.00125) and np.all(b["m"] == 0.00125)
    assert np.all(np.diff(np.concatenate([a["x"][:, 0], b["x"][:, 0]])) > 0)


def test_write_read_round_trip_is_exact():
    d = restart.read_flat(GOLDEN)
    rng = np.random.default_rng(3)
    states = {"nodes1": restart.nodelist_state(d, "nodes1", 1)}
    n = 50
    H = np.zeros((n, 3, 3))
    A = rng.normal(size=(n, 3, 3))
    H[:] = A + np.transpose(A, (0, 2, 1)) + 8 * np.eye(3)
    states["fluid3d"] = dict(x=rng.normal(size=(n, 3)), v=rng.normal(size=(n, 3)), m=rng.uniform(0.1, 1, n),
                             rho=rng.uniform(0.5, 2, n), eps=rng.uniform(0.1, 3, n), H=H)
    buf = io.StringIO()
    for label, st in states.items():
        restart.write_nodelist(buf, label, st, prefix="restart_cycle3/")
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "spx_restart_roundtrip")
    with open(path, "w") as f:
        f.write(buf.getvalue())
    back = restart.read_flat(path)
    os.remove(path)
    for label, st in states.items():
        dim = st["x"].shape[1]
        got = restart.nodelist_state(back, label, dim)
        for k in ("x", "v", "m", "rho", "eps", "H"):
            assert np.array_equal(got[k], st[k]), (label, k)


def test_value_layout_is_the_reference_stream_format():
    """A Vector prints as '( x y z )' and a SymTensor as its unique elements in storage order."""
    buf = io.StringIO()
    H = np.array([[[1.0, 2.0, 3.0], [2.0, 4.0, 5.0], [3.0, 5.0, 6.0]]])
    restart.write_nodelist(buf, "n", dict(x=np.array([[1.0, 2.0, 3.0]]), v=np.zeros((1, 3)), m=[1.0], rho=[1.0],
                                          eps=[1.0], H=H))
    lines = {l.split()[0]: l for l in buf.getvalue().splitlines()}
    assert lines["n/positions"] == "n/positions 1 ( 1 2 3 )"
    assert lines["n/H"] == "n/H 1 ( 1 2 3 4 5 6 )"
    assert lines["n/numNodes"] == "n/numNodes 1"


def test_written_paths_cover_what_the_reference_restores():
    """Every path NodeList::restoreState and FluidNodeList::restoreState read (NodeList.cc:529-541,
    FluidNodeList.cc:181-185; FlatFileIO fails on a missing path) is written, and a state's work
    field is carried through."""
    buf = io.StringIO()
    st = dict(x=np.zeros((2, 2)), v=np.zeros((2, 2)), m=[1.0, 1.0], rho=[1.0, 1.0], eps=[1.0, 1.0],
              H=np.stack([np.eye(2)] * 2), work=[0.5, 0.25])
    restart.write_nodelist(buf, "fluid", st, prefix="restart_cycle0/")
    paths = {l.split()[0] for l in buf.getvalue().splitlines()}
    need = {"name", "numNodes", "mass", "positions", "velocity", "H", "work", "massDensity",
            "specificThermalEnergy"}
    assert {"restart_cycle0/fluid/" + k for k in need} <= paths
    lines = {l.split()[0]: l for l in buf.getvalue().splitlines()}
    assert lines["restart_cycle0/fluid/work"] == "restart_cycle0/fluid/work 2 0.5 0.25"
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "spx_restart_work")
    with open(path, "w") as f:
        f.write(buf.getvalue())
    back = restart.read_flat(path)
    os.remove(path)
    assert np.array_equal(restart.nodelist_state(back, "fluid", 2)["work"], [0.5, 0.25])
