"""Order-free comparison of quadtrees in the reference layout.

The device tree numbers the boxes that level restriction and re-meshing add
in parent order; the reference appends them and reorganises differently.
Boxes are therefore compared by cell -- (level, ix, iy) from the centre --
together with everything attached to a box: parent, children in child order,
colleagues in list order, centre, and values.
"""
import numpy as np


def _get(t, name):
    v = getattr(t, name)
    return np.asarray(v)


def cells(t, boxlen=1.0):
    lev = _get(t, "ilevel").astype(np.int64)
    bs = np.asarray(t.boxsize)[lev]
    c = np.asarray(t.centers)
    ix = np.rint((c[:, 0] + boxlen / 2) / bs - 0.5).astype(np.int64)
    iy = np.rint((c[:, 1] + boxlen / 2) / bs - 0.5).astype(np.int64)
    return (lev << 58) | (ix << 29) | iy


def canonical(t, boxlen=1.0):
    """dict cell -> (nchild, parent cell, children cells, colleague cells, index)."""
    key = cells(t, boxlen)
    nb = key.shape[0]
    par = _get(t, "iparent")
    nch = _get(t, "nchild")
    ich = _get(t, "ichild").reshape(nb, 4)
    ncoll = _get(t, "ncoll")
    coll = _get(t, "coll").reshape(nb, 9)
    out = {}
    for b in range(nb):
        p = int(key[par[b] - 1]) if par[b] > 0 else -1
        kids = tuple(int(key[c - 1]) if c > 0 else -1 for c in ich[b])
        cl = tuple(int(key[c - 1]) for c in coll[b, : ncoll[b]])
        out[int(key[b])] = (int(nch[b]), p, kids, cl, b)
    return out


def compare_trees(a, b, boxlen=1.0, check_colleagues=True):
    """Raise AssertionError on the first structural difference; returns the
    index maps (cell -> box index) of both."""
    ca, cb = canonical(a, boxlen), canonical(b, boxlen)
    sa, sb = set(ca), set(cb)
    if sa != sb:
        raise AssertionError(f"box sets differ: {len(sa - sb)} only in the first, {len(sb - sa)} only in the second "
                             f"({len(sa)} vs {len(sb)} boxes)")
    for k in ca:
        x, y = ca[k], cb[k]
        if x[:3] != y[:3]:
            raise AssertionError(f"box {k}: nchild/parent/children differ: {x[:3]} vs {y[:3]}")
        if check_colleagues and x[3] != y[3]:
            raise AssertionError(f"box {k}: colleagues differ: {x[3]} vs {y[3]}")
    return {k: v[4] for k, v in ca.items()}, {k: v[4] for k, v in cb.items()}


def compare_values(a, va, b, vb, boxlen=1.0, leaves_only=True):
    """Max relative difference of the values of matching boxes (scaled by the
    largest value of the first tree)."""
    ia, ib = compare_trees(a, b, boxlen, check_colleagues=False)
    nch = _get(a, "nchild")
    ks = [k for k in ia if (not leaves_only or nch[ia[k]] == 0)]
    xa = np.stack([va[ia[k]] for k in ks])
    xb = np.stack([vb[ib[k]] for k in ks])
    scale = max(np.abs(xa).max(), 1e-300)
    return float(np.abs(xa - xb).max() / scale)


def level_restricted(t, iperiod, boxlen=1.0):
    """True when touching leaves (edges and corners, periodic wrap) differ by at most one level."""
    key = cells(t, boxlen)
    lev = (key >> 58).astype(np.int64)
    ix = ((key >> 29) & ((1 << 29) - 1)).astype(np.int64)
    iy = (key & ((1 << 29) - 1)).astype(np.int64)
    nch = _get(t, "nchild")
    leaves = np.nonzero(nch == 0)[0]
    L = int(lev.max())
    # paint the finest grid with the level of the leaf covering each cell
    g = np.full((1 << L, 1 << L), -1, dtype=np.int64)
    for b in leaves:
        s = 1 << (L - lev[b])
        g[ix[b] * s : (ix[b] + 1) * s, iy[b] * s : (iy[b] + 1) * s] = lev[b]
    if (g < 0).any():
        return False
    pads = g
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            if iperiod:
                h = np.roll(np.roll(pads, dx, axis=0), dy, axis=1)
                if np.abs(h - pads).max() > 1:
                    return False
            else:
                n = pads.shape[0]
                a = pads[max(dx, 0) : n + min(dx, 0), max(dy, 0) : n + min(dy, 0)]
                bb = pads[max(-dx, 0) : n + min(-dx, 0), max(-dy, 0) : n + min(-dy, 0)]
                if a.size and np.abs(a - bb).max() > 1:
                    return False
    return True


def same_tree_arrays(a, b):
    """Box-for-box comparison in the given numbering: nlevels, laddr, levels,
    parents, children, colleague lists (the first ncoll entries) and centres
    (bit for bit).  Returns a list of differences, empty when equal."""
    msgs = []
    if a.nlevels != b.nlevels:
        return [f"nlevels {a.nlevels} != {b.nlevels}"]
    nb = a.nboxes
    if nb != b.nboxes:
        return [f"box counts {nb} != {b.nboxes}"]
    if not np.array_equal(a.laddr, b.laddr):
        msgs.append("laddr differs")
    for name in ("ilevel", "iparent", "nchild", "ichild", "ncoll"):
        x, y = np.asarray(getattr(a, name)), np.asarray(getattr(b, name))
        if not np.array_equal(x, y):
            bad = np.nonzero((x != y).reshape(nb, -1).any(axis=1))[0]
            msgs.append(f"{name} differs at boxes {bad[:8].tolist()}")
    if not msgs:
        m = np.arange(9)[None, :] < np.asarray(a.ncoll)[:, None]
        if not np.array_equal(np.where(m, a.coll, 0), np.where(m, b.coll, 0)):
            msgs.append("colleague lists differ")
    if not np.array_equal(np.asarray(a.centers)[:nb], np.asarray(b.centers)[:nb]):
        msgs.append("centres differ")
    return msgs
