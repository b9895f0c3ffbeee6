"""Neighbour sets in the reference's ConnectivityMap numbering, with its ghost nodes.

The engine reports each neighbour as (control node, image code) (`spx_run_neighbors`).  The
reference instead numbers a ghost by its place in the NodeList: numInternal + its rank among the
ghosts that survive culling, in creation order.  This module rebuilds that numbering on the host:

- creation (Integrator.cc:414-423; PeriodicBoundary.cc:78-82; PlanarBoundary.cc:102-117;
  findNodesTouchingThroughPlanes.cc:40-66): boundaries in the order they were appended, a periodic
  boundary's first planar part (enter = its first plane, exit = its second) before its second; each
  part scans the NodeList's nodes present at that moment (internal, then the ghosts of earlier
  parts) in index order.  hmax is the largest 1/min eig(H) of the nodes closer than
  kernelExtent*h to either plane; a node becomes a control node when 0 <= d_exit/hmax <=
  kernelExtent, d_exit its signed distance to the exit plane, and its ghost is appended;
- culling (Integrator.cc:448-527): a ghost is kept when an internal node of any NodeList has it
  as a neighbour, or when it is the control node of another ghost; kept ghosts keep their order;
- lists (ConnectivityMap.cc:642): each node's neighbours in one NodeList sorted by index.

Image codes follow the engine's: per axis 2 bits at 2a, 1 = the image beyond the low plane
(controls near the high plane for a periodic axis, near the low plane for a reflecting one), 2 =
beyond the high plane.
"""
import numpy as np

from . import abi


def default_boundaries(cfg):
    """The boundary sequence of a script that appends them axis by axis, the low plane first:
    ("periodic", a, first) for a periodic axis whose first plane is on side `first` (1 low, 2
    high), ("reflecting", a, side) per reflecting plane."""
    out = []
    for a in range(cfg.dim):
        if cfg.bc_lo[a] == abi.SPX_BC_PERIODIC:
            out.append(("periodic", a, 1))
            continue
        if cfg.bc_lo[a] == abi.SPX_BC_REFLECTING:
            out.append(("reflecting", a, 1))
        if cfg.bc_hi[a] == abi.SPX_BC_REFLECTING:
            out.append(("reflecting", a, 2))
    return out


def _image_coord(cfg, axis, code, xa):
    """Coordinate `axis` of the image with per-axis code 1 or 2, in the engine's operation order."""
    lo, hi = cfg.xmin[axis], cfg.xmax[axis]
    bc = cfg.bc_lo[axis] if code == 1 else cfg.bc_hi[axis]
    three = cfg.dim == 3
    if bc == abi.SPX_BC_PERIODIC:
        if code == 2:
            return ((xa + (hi - xa)) + (xa - lo)) if three else hi + (xa - lo)
        return ((xa + (lo - xa)) + (xa - hi)) if three else lo + (xa - hi)
    if code == 1:
        return ((xa + (lo - xa)) - (xa - lo)) if three else lo - (xa - lo)
    return ((xa + (hi - xa)) - (xa - hi)) if three else hi - (xa - hi)


def _planar_parts(cfg, boundaries):
    """(axis, code, exit_sign, exit_point, enter_point) of every planar part in creation order;
    exit_sign*(x - exit_point) is the signed distance to the exit plane."""
    parts = []
    for b in boundaries:
        if b[0] == "periodic":
            a = b[1]
            lo, hi = cfg.xmin[a], cfg.xmax[a]
            lo_first = (b[2] if len(b) > 2 else 1) == 1
            p1 = (a, 1, -1.0, hi, lo)               # enter = low plane, exit = high plane
            p2 = (a, 2, 1.0, lo, hi)                # enter = high plane, exit = low plane
            parts.extend([p1, p2] if lo_first else [p2, p1])
        elif b[0] == "reflecting":
            a, side = b[1], b[2]
            p = cfg.xmin[a] if side == 1 else cfg.xmax[a]
            parts.append((a, side, 1.0 if side == 1 else -1.0, p, p))
        else:
            raise ValueError(f"unknown boundary {b!r}")
    return parts


def ghost_nodes(cfg, x, H, boundaries=None):
    """Ghosts of one NodeList before culling, in creation order.  Returns (parent, control, code,
    position): parent = the node index (internal or ghost, in the pre-cull numbering) the ghost was
    made from, control = the internal node it images."""
    x = np.asarray(x, np.float64)
    n, dim = x.shape
    if boundaries is None:
        boundaries = default_boundaries(cfg)
    ext = cfg.neighbor_extent
    hmaxi = 1.0/np.linalg.eigvalsh(np.asarray(H, np.float64)).min(axis=1)
    pos, hm = x.copy(), hmaxi.copy()
    parent = np.arange(n)
    control = np.arange(n)
    code = np.zeros(n, np.int64)
    for a, c, sgn, pexit, penter in _planar_parts(cfg, boundaries):
        xa = pos[:, a]
        d_exit = sgn*(xa - pexit)
        near = np.minimum(np.abs(xa - pexit), np.abs(xa - penter)) < ext*hm
        hmax = hm[near].max() if near.any() else 0.0
        if not hmax > 0.0:
            continue
        dist = d_exit/hmax
        sel = np.nonzero((dist >= 0.0) & (dist <= ext))[0]
        if sel.size == 0:
            continue
        if np.any((code[sel] >> (2*a)) & 3):
            raise ValueError("a node and its image on one axis both reach a plane: the domain is "
                             "narrower than two kernel supports")
        newpos = pos[sel].copy()
        newpos[:, a] = _image_coord(cfg, a, c, pos[sel, a])
        pos = np.concatenate([pos, newpos])
        hm = np.concatenate([hm, hm[sel]])
        parent = np.concatenate([parent, sel])
        control = np.concatenate([control, control[sel]])
        code = np.concatenate([code, code[sel] | (c << (2*a))])
    return parent[n:], control[n:], code[n:], pos[n:]


def node_pair_hash(i_node, i_list, j_node, j_list):
    """NodePairIdxType::hash (Neighbor/NodePairIdxType.hh:34-53) on 64-bit size_t: the pair is
    flipped to (lower, higher) by (list, node), then (i_list << 59) + (i_node << 32) +
    (j_list << 27) + j_node.  Requires lists < 32 and nodes < 2^27, as the reference does."""
    i_node, i_list, j_node, j_list = (np.asarray(a, np.uint64) for a in (i_node, i_list, j_node, j_list))
    if np.any(i_node >= 1 << 27) or np.any(j_node >= 1 << 27) or np.any(i_list >= 32) or np.any(j_list >= 32):
        raise ValueError("NodePairIdxType needs node indices < 2^27 and NodeList indices < 32")
    flip = (i_list > j_list) | ((i_list == j_list) & (i_node > j_node))
    a_l, a_n = np.where(flip, j_list, i_list), np.where(flip, j_node, i_node)
    b_l, b_n = np.where(flip, i_list, j_list), np.where(flip, i_node, j_node)
    u = np.uint64
    return (a_l << u(59)) + (a_n << u(32)) + (b_l << u(27)) + b_n


class ConnectivityDump:
    """ConnectivityMap contents in the reference's numbering for NodeLists `labels`.

    ghosts[k]: dict(control, parent, code, position, created) of NodeList k's kept ghosts in index
    order (ghost g has index numInternal_k + g; parent is in the same numbering; created is the
    ghost's place in the creation order before culling).
    connectivityForNode(k, i) -> list over NodeLists l of the sorted neighbour indices of node i."""

    def __init__(self, labels, sizes, ghosts, csr):
        self.labels, self.sizes, self.ghosts, self._csr = labels, sizes, ghosts, csr

    def _index(self, k):
        if isinstance(k, (int, np.integer)):
            return int(k)
        return self.labels.index(getattr(k, "name", k))

    def connectivityForNode(self, k, i):
        """Neighbours of node i of NodeList k (an index, a label, or an object with .name)."""
        k = self._index(k)
        out = []
        for l in range(len(self.labels)):
            off, idx = self._csr[(k, l)]
            out.append(idx[off[i]:off[i + 1]])
        return out

    def numNeighborsForNode(self, k, i):
        return sum(len(v) for v in self.connectivityForNode(k, i))

    def nodePairList(self):
        """ConnectivityMap::nodePairList() (Neighbor/ConnectivityMap.hh:79): one NodePairIdxType per
        (internal node i of NodeList k, neighbour j of NodeList l) that calculatePairInteraction
        admits -- l > k, or l == k and j > i, or l < k and j a ghost of l
        (ConnectivityMapInline.hh:296-309) -- stored as (i_node, i_list, j_node, j_list) in
        ascending NodePairIdxType::hash order (ConnectivityMap.cc:925-927, 1028;
        NodePairIdxType.hh:34-53)."""
        parts = []
        for k in range(len(self.labels)):
            for l in range(len(self.labels)):
                off, idx = self._csr[(k, l)]
                i = np.repeat(np.arange(self.sizes[k], dtype=np.int64), np.diff(off))
                j = np.asarray(idx, np.int64)
                if l > k:
                    keep = np.ones(len(j), bool)
                elif l == k:
                    keep = j > i
                else:
                    keep = j >= self.sizes[l]
                parts.append((i[keep], np.full(keep.sum(), k, np.int64), j[keep], np.full(keep.sum(), l, np.int64)))
        pl = np.zeros(sum(len(p[0]) for p in parts),
                      dtype=[("i_node", np.int64), ("i_list", np.int64), ("j_node", np.int64), ("j_list", np.int64)])
        if len(pl) == 0:
            return pl
        for f, a in zip(pl.dtype.names, map(np.concatenate, zip(*parts))):
            pl[f] = a
        order = np.argsort(node_pair_hash(pl["i_node"], pl["i_list"], pl["j_node"], pl["j_list"]), kind="stable")
        return pl[order]

    def write(self, f):
        """Append the dump to the open text file `f`, one FlatFileIO-style line per path."""
        for k, lab in enumerate(self.labels):
            g = self.ghosts[k]
            ng = len(g["control"])
            f.write(f"{lab}/numInternalNodes {self.sizes[k]}\n")
            f.write(f"{lab}/ghostControl {ng} " + " ".join(map(str, g["control"])) + "\n")
            f.write(f"{lab}/ghostImage {ng} " + " ".join(map(str, g["code"])) + "\n")
            for l, labl in enumerate(self.labels):
                off, idx = self._csr[(k, l)]
                for i in range(self.sizes[k]):
                    row = idx[off[i]:off[i + 1]]
                    f.write(f"{lab}/{i}/{labl} {len(row)} " + " ".join(map(str, row)) + "\n")


def connectivity_map(cfg, x, H, offsets, nbr, img, nodelists=None, boundaries=None):
    """The reference-numbered ConnectivityMap of a state whose neighbour sets are (offsets, nbr,
    img) as `spx_run_neighbors` reports them.  nodelists: [(label, first, count)] ranges of the
    node order (one NodeList "nodes" when None)."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    if nodelists is None:
        nodelists = [("nodes", 0, n)]
    labels = [nl[0] for nl in nodelists]
    starts = np.array([nl[1] for nl in nodelists], np.int64)
    sizes = [int(nl[2]) for nl in nodelists]
    if sum(sizes) != n or np.any(np.diff(starts) != np.array(sizes[:-1])) or starts[0] != 0:
        raise ValueError("NodeList ranges must tile the node order")
    nl_of = np.repeat(np.arange(len(sizes)), sizes)
    local = np.arange(n) - starts[nl_of]

    offsets, nbr, img = (np.asarray(a, np.int64) for a in (offsets, nbr, img))
    src = np.repeat(np.arange(n), np.diff(offsets))
    jl = nl_of[nbr]                                   # NodeList of each neighbour
    index = local[nbr].copy()                         # internal neighbours: local index
    pre, kept_rank, ghosts = [], [], []
    for k, (lab, first, cnt) in enumerate(nodelists):
        parent, control, code, position = ghost_nodes(cfg, x[first:first + cnt], H[first:first + cnt], boundaries)
        keys = control*64 + code
        order = np.argsort(keys, kind="stable")
        pre.append((keys[order], order))
        used = np.zeros(len(control), bool)
        sel = (jl == k) & (img != 0)
        if sel.any():
            q = local[nbr[sel]]*64 + img[sel]
            at = np.searchsorted(keys[order], q)
            ok = (at < len(keys)) & (keys[order][np.minimum(at, len(keys) - 1)] == q)
            if not ok.all():
                raise ValueError(f"NodeList {lab}: a neighbour image has no ghost under the reference's rule")
            used[order[at]] = True
        # every ghost that is the control node of a ghost is kept, whether or not that ghost is
        # (Integrator.cc:470-476 flags each boundary's ghost control nodes before culling)
        par_ghost = parent[parent >= cnt] - cnt
        used[par_ghost] = True
        rank = np.cumsum(used) - 1
        kept_rank.append(rank)
        keep = np.nonzero(used)[0]
        par = parent[keep]
        par = np.where(par >= cnt, cnt + rank[np.maximum(par - cnt, 0)], par)
        ghosts.append(dict(control=control[keep], parent=par, code=code[keep], position=position[keep], created=keep))
        if sel.any():
            index[sel] = cnt + rank[order[at]]
    csr = {}
    for k in range(len(sizes)):
        for l in range(len(sizes)):
            m = (nl_of[src] == k) & (jl == l)
            s, v = local[src[m]], index[m]
            o = np.lexsort((v, s))
            s, v = s[o], v[o]
            off = np.zeros(sizes[k] + 1, np.int64)
            np.add.at(off, s + 1, 1)
            csr[(k, l)] = (np.cumsum(off), v)
    return ConnectivityDump(labels, sizes, ghosts, csr)
