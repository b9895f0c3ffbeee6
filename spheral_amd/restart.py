"""Restart state exchange in the reference's FlatFileIO ASCII layout.

A Spheral restart written with FlatFileIO (FileIO/FlatFileIO.cc) is a text file of lines
``<path> <value>`` for scalars and ``<path> <n> <v1> ... <vn>`` for Fields and vectors, one line per
path (FlatFileIO.cc:486-545).  Geometric values print as ``( a b c )`` (Geometry/GeomVectorInline.hh:
1270-1300, GeomSymmetricTensorInline.hh:2694-2725): a Vector lists its nDim components, a SymTensor its
unique elements in storage order (xx | xx xy yy | xx xy xz yy yz zz, GeomSymmetricTensorBase.hh).
FlatFileIO writes with 20 significant digits, so doubles survive the round trip.

A NodeList dumps ``<label>/name``, ``numNodes``, ``mass``, ``positions``, ``velocity``, ``H``,
``work`` (NodeList.cc:503-518; restoreState reads every one of them back, NodeList.cc:526-541) and a
FluidNodeList adds ``massDensity`` and ``specificThermalEnergy`` (FluidNodeList.cc:160-186); the restart registrar prefixes each object's label (DataOutput/RestartRegistrar.cc:
125-176), and older files put every label under a ``restart_cycle<N>/`` group.

``read_flat`` parses a file into {path: value}; ``nodelist_state`` extracts one NodeList as the
engine's state record (x, v, m, rho, eps, H with full H matrices); ``write_nodelist`` writes a state
back in the same layout, so a run can start from a reference restart and hand its state back.
"""
import numpy as np

_SYM_INDEX = {1: [(0, 0)], 2: [(0, 0), (0, 1), (1, 1)], 3: [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]}


def _parse_values(tokens):
    """Numbers, or parenthesised groups of numbers, as a flat list of floats or of tuples."""
    out, group, inside = [], [], False
    for t in tokens:
        if t == "(":
            inside, group = True, []
        elif t == ")":
            inside = False
            out.append(tuple(group))
        elif inside:
            group.append(float(t))
        else:
            out.append(float(t))
    return out


def read_flat(path):
    """{path: value} of a FlatFileIO ASCII file.  A line with one token after the path is a
    scalar (kept as text when it is not a number); otherwise the first token is the count."""
    data = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            key, rest = parts[0], parts[1:]
            if len(rest) == 1:
                try:
                    data[key] = float(rest[0])
                except ValueError:
                    data[key] = rest[0]
                continue
            if not rest:
                data[key] = ""
                continue
            try:
                n = int(rest[0])
            except ValueError:
                data[key] = " ".join(rest)
                continue
            vals = _parse_values(rest[1:])
            if len(vals) != n:
                raise ValueError(f"{key}: {n} values announced, {len(vals)} read")
            data[key] = vals
    return data


def _prefix(data, label):
    """The path prefix under which NodeList ``label`` was dumped (with or without a cycle group)."""
    for k in data:
        if k.endswith("/" + label + "/numNodes") or k == label + "/numNodes":
            return k[: -len("numNodes")]
    raise KeyError(f"no NodeList '{label}' in the restart data")


def nodelist_state(data, label, dim):
    """(x, v, m, rho, eps, H) of NodeList ``label`` as engine state arrays."""
    p = _prefix(data, label)
    n = int(data[p + "numNodes"])

    def vec(name):
        a = np.array(data[p + name], dtype=np.float64).reshape(n, -1)
        if a.shape[1] != dim:
            raise ValueError(f"{name}: {a.shape[1]} components, expected {dim}")
        return a

    Hs = np.array(data[p + "H"], dtype=np.float64).reshape(n, -1)
    idx = _SYM_INDEX[dim]
    if Hs.shape[1] != len(idx):
        raise ValueError(f"H: {Hs.shape[1]} elements, expected {len(idx)}")
    H = np.zeros((n, dim, dim))
    for c, (a, b) in enumerate(idx):
        H[:, a, b] = Hs[:, c]
        H[:, b, a] = Hs[:, c]
    st = dict(x=vec("positions"), v=vec("velocity"), m=np.array(data[p + "mass"], np.float64),
              rho=np.array(data[p + "massDensity"], np.float64),
              eps=np.array(data[p + "specificThermalEnergy"], np.float64), H=H)
    if p + "work" in data:
        st["work"] = np.array(data[p + "work"], np.float64)
    return st


def _num(x):
    # 20 significant digits as FlatFileIO's stream precision prints them
    return f"{float(x):.20g}"


def write_nodelist(f, label, state, prefix=""):
    """Append NodeList ``label`` in FlatFileIO layout to the open text file ``f``."""
    x, v = np.asarray(state["x"]), np.asarray(state["v"])
    n, dim = x.shape
    p = prefix + label + "/"
    grp = lambda a: "( " + " ".join(_num(c) for c in a) + " )"
    f.write(f"{p}name {label}\n")
    f.write(f"{p}numNodes {n}\n")
    f.write(f"{p}mass {n} " + " ".join(_num(m) for m in state["m"]) + "\n")
    f.write(f"{p}positions {n} " + " ".join(grp(r) for r in x) + "\n")
    f.write(f"{p}velocity {n} " + " ".join(grp(r) for r in v) + "\n")
    idx = _SYM_INDEX[dim]
    H = np.asarray(state["H"])
    f.write(f"{p}H {n} " + " ".join(grp([h[a, b] for a, b in idx]) for h in H) + "\n")
    # the NodeList's work field (a load-balancing measure the hydro does not evolve): carried through
    # when the state has it, zeros otherwise
    work = np.asarray(state.get("work", np.zeros(n)), dtype=np.float64)
    f.write(f"{p}work {n} " + " ".join(_num(w) for w in work) + "\n")
    f.write(f"{p}massDensity {n} " + " ".join(_num(r) for r in state["rho"]) + "\n")
    f.write(f"{p}specificThermalEnergy {n} " + " ".join(_num(e) for e in state["eps"]) + "\n")


def run_state(run):
    """The state record of an engine run (spheral_amd.engine.Run) for write_nodelist."""
    return dict(x=run.field("position"), v=run.field("velocity"), m=run.field("mass"),
                rho=run.field("mass density"), eps=run.field("specific thermal energy"), H=run.field("H"))
