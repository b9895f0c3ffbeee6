"""Pins of the oracle's connectivity (Neighbor/ConnectivityMap.cc:905-947) by brute force,
as the reference's own neighbour test does (tests/unit/Neighbor/NeighborTestBase.py:134-184:
random positions, random anisotropic H, criterion min(|Hi r|, |Hj r|) <= radius)."""
import numpy as np
import pytest

from spheral_amd import abi, problems


def brute(x, H, ext, images):
    """O(N^2) over explicit images: images is a list of (code, shift-function) pairs."""
    n, D = x.shape
    out = []
    for i in range(n):
        s = set()
        for code, fn in images:
            xj = fn(x)
            r = x[i][None, :] - xj
            ei = np.einsum("ab,nb->na", H[i], r)
            ej = np.einsum("nab,nb->na", H, r)
            ok = (np.sum(ei*ei, 1) <= ext*ext) | (np.sum(ej*ej, 1) <= ext*ext)
            for j in np.nonzero(ok)[0]:
                if code == 0 and j == i:
                    continue
                s.add((int(j), int(code)))
        out.append(sorted(s))
    return out


def flatten(lists):
    off = np.cumsum([0] + [len(l) for l in lists])
    nb = np.array([p[0] for l in lists for p in l], np.int64)
    im = np.array([p[1] for l in lists for p in l], np.int32)
    return off, nb, im


def random_sym(rng, n, D, hmin, hmax):
    H = np.zeros((n, D, D))
    for i in range(n):
        Q, _ = np.linalg.qr(rng.normal(size=(D, D)))
        ev = 1.0/rng.uniform(hmin, hmax, D)
        H[i] = Q @ np.diag(ev) @ Q.T
        H[i] = 0.5*(H[i] + H[i].T)
    return H


@pytest.mark.parametrize("D", [1, 2, 3])
def test_open_boundaries_anisotropic_H(oracle_mod, D):
    rng = np.random.default_rng(11 + D)
    n = {1: 60, 2: 150, 3: 250}[D]
    x = rng.uniform(0, 1, (n, D))
    H = random_sym(rng, n, D, 0.02, 0.12)
    cfg = abi.default_config(D)
    r = oracle_mod.OracleRun(cfg, n)
    r.set_state(x, np.zeros_like(x), np.ones(n), np.ones(n), np.ones(n), H)
    r.update_connectivity()
    off, nb, im = r.neighbors()
    boff, bnb, bim = flatten(brute(x, H, cfg.neighbor_extent, [(0, lambda y: y)]))
    assert np.array_equal(off, boff) and np.array_equal(nb, bnb) and np.array_equal(im, bim)


def _images(D, kinds, lo, hi):
    """All image combinations written as the plain geometry: periodic shift by L, mirror 2p - x."""
    import itertools
    opts = []
    for a in range(D):
        o = [(0, None)]
        if kinds[a] == "p":
            o += [(1, lambda y, a=a: y[:, a] - (hi - lo)), (2, lambda y, a=a: y[:, a] + (hi - lo))]
        elif kinds[a] == "r":
            o += [(1, lambda y, a=a: 2*lo - y[:, a]), (2, lambda y, a=a: 2*hi - y[:, a])]
        opts.append(o)
    out = []
    for combo in itertools.product(*opts):
        code = 0
        fns = []
        for a, (c, fn) in enumerate(combo):
            code |= c << (2*a)
            if fn is not None:
                fns.append((a, fn))
        def F(y, fns=fns):
            z = y.copy()
            for a, fn in fns:
                z[:, a] = fn(y)
            return z
        out.append((code, F))
    return out


@pytest.mark.parametrize("D,kinds", [(1, "r"), (2, "pp"), (2, "pr"), (3, "ppp"), (3, "rpn")])
def test_periodic_and_reflecting_images(oracle_mod, D, kinds):
    rng = np.random.default_rng(5)
    n = {1: 50, 2: 220, 3: 400}[D]
    x = rng.uniform(0, 1, (n, D))
    h = rng.uniform(0.05, 0.09, n)
    H = np.zeros((n, D, D))
    for a in range(D):
        H[:, a, a] = 1.0/h
    cfg = abi.default_config(D)
    for a, k in enumerate(kinds):
        b = {"p": abi.SPX_BC_PERIODIC, "r": abi.SPX_BC_REFLECTING, "n": abi.SPX_BC_NONE}[k]
        cfg.bc_lo[a] = cfg.bc_hi[a] = b
        cfg.xmin[a], cfg.xmax[a] = 0.0, 1.0
    r = oracle_mod.OracleRun(cfg, n)
    r.set_state(x, np.zeros_like(x), np.ones(n), np.ones(n), np.ones(n), H)
    r.update_connectivity()
    off, nb, im = r.neighbors()
    boff, bnb, bim = flatten(brute(x, H, cfg.neighbor_extent, _images(D, kinds, 0.0, 1.0)))
    assert np.array_equal(off, boff) and np.array_equal(nb, bnb) and np.array_equal(im, bim)


def test_perfect_lattice_has_32_neighbours(oracle_mod):
    """nPerh = 2, extent 1: integer points with i^2+j^2+k^2 <= 4 -> 32 neighbours (exact
    arithmetic at n = 16: every coordinate and difference is a dyadic rational)."""
    cfg, st = problems.uniform_lattice(3, 16)
    r = oracle_mod.OracleRun(cfg, len(st["m"]))
    r.set_state(**st)
    r.update_connectivity()
    off, _, _ = r.neighbors()
    assert np.all(np.diff(off) == 32)


def test_pairs_are_unique_and_symmetric(oracle_mod):
    cfg, st = problems.taylor_green_3d(10)
    r = oracle_mod.OracleRun(cfg, len(st["m"]))
    r.set_state(**st)
    r.update_connectivity()
    off, nb, im = r.neighbors()
    n = len(st["m"])
    pairs = set()
    for i in range(n):
        for k in range(off[i], off[i + 1]):
            pairs.add((i, int(nb[k]), int(im[k])))
    # internal pairs are symmetric; an image pair (i, j, c) mirrors as (j, i, c') with the opposite sides
    flip = lambda c: sum((((c >> (2*a)) & 3) and (3 - ((c >> (2*a)) & 3))) << (2*a) for a in range(3))
    for (i, j, c) in pairs:
        assert (j, i, flip(c)) in pairs
