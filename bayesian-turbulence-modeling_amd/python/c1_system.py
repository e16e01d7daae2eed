"""The large pressure-type system of the backward-facing step (a 250k-cell
mixed triangle/quad mesh) for the stand-alone linear-solve path.

Domain: step height h = 1 with expansion ratio 1.2 (upstream channel
y in [1, 6], downstream y in [0, 6]), 10 h upstream and 30 h downstream of the
step.  Mesh: wall-clustered quad layers along the top and bottom walls of both
blocks (a wall-graded tensor grid); the core of each block is a seeded
Delaunay triangulation (core="delaunay", the default).  Its points are the
core's boundary nodes (shared with the quad layers and, at x = 0, with the
other block) plus one seeded point per interior grid node, displaced uniformly
within +-0.3 of a grid cell.  The triangulation (scipy.spatial.Delaunay, i.e.
Qhull) is computed in the grid's index coordinates, where the point spacing is
isotropic, and mapped node by node through the monotone wall grading to
physical space; mirrored ghost points outside the core make every boundary
segment a Gabriel edge, so the triangles conform to the quad layers.
core="split" gives the older variant: jittered core quads split into two
triangles along alternating diagonals.  Faces are numbered internal first; the
owner is the lower cell index.

System (the matrix class of assemblePressureCorrection,
incompressible/src/SIMPLESolver.cpp:322-450): per-cell coefficient nu_t
log-normal(0, sigma = 1) (seeded), face coefficient
a_f = (w nu_o + (1 - w) nu_n) A_f / delta_f, upper = lower = -a_f,
diag = sum of the cell's a_f plus the outlet Dirichlet term nu_o A_b / delta_b
(SIMPLESolver.cpp:432-439); an SPD M-matrix.  Right-hand side b ~ N(0, 1)
(seeded) with its mean removed.

make_system(...) returns a dict of numpy arrays (own, nbr, diag, upper,
lower, b) plus sizes and geometry summaries.
"""
import numpy as np


def _clustered(n, length, ratio_lo, ratio_hi=None):
    """n cells over [0, length], geometric growth away from both ends
    (ratio_hi None: only from the low end)."""
    if ratio_hi is None:
        g = ratio_lo ** np.arange(n)
    else:
        half = n // 2
        a = ratio_lo ** np.arange(half)
        b = (ratio_hi ** np.arange(n - half))[::-1]
        g = np.concatenate([a, b])
    x = np.concatenate([[0.0], np.cumsum(g)])
    return x / x[-1] * length


def _cells_and_nodes(nx_up, nx_dn, ny_up, ny_low, jitter, seed, keep_rows=None):
    """Structured node grid of the two blocks (conformal at x = 0) and their
    quads.  Interior nodes are jittered (seeded); keep_rows(block, ny) names
    node rows left straight (the Delaunay core's boundary rows)."""
    rng = np.random.default_rng(seed)
    # streamwise nodes: clustered towards the step on both sides
    xu = -_clustered(nx_up, 10.0, 1.01)[::-1]                     # [-10, 0], fine at x = 0
    xd = _clustered(nx_dn, 30.0, 1.004)                           # [0, 30], fine at x = 0
    # wall-normal nodes
    yu = 1.0 + _clustered(ny_up, 5.0, 1.03, 1.03)                 # [1, 6]
    yl = _clustered(ny_low, 1.0, 1.05, 1.05)                      # [0, 1]
    yd = np.concatenate([yl, yu[1:]])                             # [0, 6]
    blocks = []
    base = 0
    node_xy = []
    for xs, ys in ((xu, yu), (xd, yd)):
        X, Y = np.meshgrid(xs, ys, indexing="ij")                 # [nx+1, ny+1]
        ids = base + np.arange(X.size).reshape(X.shape)
        node_xy.append(np.stack([X.ravel(), Y.ravel()], axis=1))
        blocks.append(ids)
        base += X.size
    xy = np.concatenate(node_xy)
    # merge the shared interface column (upstream x = 0 nodes == downstream x = 0, y >= 1)
    up, dn = blocks
    shared_dn = dn[0, len(yl) - 1:]                               # downstream nodes at x = 0, y in [1, 6]
    remap = np.arange(len(xy))
    remap[shared_dn] = up[-1, :]
    # interior jitter (not on walls / inlet / outlet / interface)
    for b, (ids, xs, ys) in enumerate(((up, xu, yu), (dn, xd, yd))):
        inner = ids[1:-1, 1:-1]
        hx = np.diff(xs)[:-1][:, None] * np.ones((1, len(ys) - 2))
        hy = np.ones((len(xs) - 2, 1)) * np.diff(ys)[:-1][None, :]
        dx = rng.uniform(-jitter, jitter, inner.shape) * np.minimum(hx, np.diff(xs)[1:][:, None])
        dy = rng.uniform(-jitter, jitter, inner.shape) * np.minimum(hy, np.diff(ys)[1:][None, :])
        if keep_rows is not None:
            fixed = np.zeros(len(ys), bool)
            fixed[list(keep_rows(b, len(ys) - 1))] = True
            dx[:, fixed[1:-1]] = 0.0
            dy[:, fixed[1:-1]] = 0.0
        xy[inner.ravel(), 0] += dx.ravel()
        xy[inner.ravel(), 1] += dy.ravel()
    quads = []
    for ids in (up, dn):
        q = np.stack([ids[:-1, :-1], ids[1:, :-1], ids[1:, 1:], ids[:-1, 1:]], axis=-1)   # ccw
        quads.append(q)
    return xy, remap, quads, [(up, xu, yu), (dn, xd, yd)]


def _delaunay_core(ids, xs, ys, wl, jit, rng):
    """Seeded Delaunay triangulation of one block's core (node rows wl ..
    ny - wl, all columns).  Returns the new interior points (physical
    coordinates) and the triangles as rows of (kind, index): kind 0 a
    structured node id of the block's boundary ring, kind 1 a new point."""
    from scipy.spatial import Delaunay
    nx, ny = len(xs) - 1, len(ys) - 1
    j0, j1 = wl, ny - wl
    # boundary ring of the core box in index space (integer lattice points)
    ring = [(i, j0) for i in range(nx + 1)] + [(i, j1) for i in range(nx + 1)] + \
           [(0, j) for j in range(j0 + 1, j1)] + [(nx, j) for j in range(j0 + 1, j1)]
    ring = np.array(ring, dtype=float)
    # one seeded point per interior grid node, within +-jit of it
    ii, jj = np.meshgrid(np.arange(1, nx), np.arange(j0 + 1, j1), indexing="ij")
    pts = np.stack([ii.ravel() + rng.uniform(-jit, jit, ii.size),
                    jj.ravel() + rng.uniform(-jit, jit, ii.size)], axis=1)
    # ghosts: the first interior row / column mirrored across each side, and the corners
    g = [pts[np.abs(pts[:, 1] - (j0 + 1)) <= jit] * [1, -1] + [0, 2 * j0],
         pts[np.abs(pts[:, 1] - (j1 - 1)) <= jit] * [1, -1] + [0, 2 * j1],
         pts[np.abs(pts[:, 0] - 1) <= jit] * [-1, 1],
         pts[np.abs(pts[:, 0] - (nx - 1)) <= jit] * [-1, 1] + [2 * nx, 0],
         np.array([[-1.0, j0 - 1], [nx + 1, j0 - 1], [-1.0, j1 + 1], [nx + 1, j1 + 1]])]
    allp = np.concatenate([ring, pts] + g)
    nreal = len(ring) + len(pts)
    simp = Delaunay(allp).simplices
    simp = simp[(simp < nreal).all(axis=1)]
    # physical coordinates through the monotone wall grading
    X = lambda a: np.interp(a, np.arange(nx + 1), xs)
    Y = lambda a: np.interp(a, np.arange(ny + 1), ys)
    phys = np.stack([X(pts[:, 0]), Y(pts[:, 1])], axis=1)
    ring_ids = ids[ring[:, 0].astype(int), ring[:, 1].astype(int)]
    # orientation: counter-clockwise in index space, and the same in physical space
    P = allp[simp]
    cross = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])
    simp[cross < 0] = simp[cross < 0][:, [0, 2, 1]]
    return phys, simp, ring_ids, len(ring)


def make_system(nx_up=150, nx_dn=450, ny_up=200, ny_low=40, wall_layers=12, jitter=0.15, seed=0,
                core="delaunay"):
    """The C1 system; the defaults give ~250k cells."""
    if core not in ("delaunay", "split"):
        raise ValueError("core must be 'delaunay' or 'split'")
    keep = (lambda b, ny: (wall_layers, ny - wall_layers)) if core == "delaunay" else None
    xy, remap, quads, blocks = _cells_and_nodes(nx_up, nx_dn, ny_up, ny_low, jitter, seed, keep)
    polys = []                                   # list of node-index arrays (ccw)
    drng = np.random.default_rng(seed + 2)
    for q, (ids, xs, ys) in zip(quads, blocks):
        nxq, nyq = q.shape[:2]
        ii, jj = np.meshgrid(np.arange(nxq), np.arange(nyq), indexing="ij")
        near_wall = (jj < wall_layers) | (jj >= nyq - wall_layers)
        qq = remap[q]
        polys.append(("quad", qq[near_wall]))
        if core == "delaunay":
            phys, simp, ring_ids, nring = _delaunay_core(ids, xs, ys, wall_layers, 0.3, drng)
            new_ids = len(xy) + np.arange(len(phys))
            xy = np.concatenate([xy, phys])
            remap = np.concatenate([remap, new_ids])
            node_of = np.concatenate([remap[ring_ids], new_ids])
            polys.append(("tri", node_of[simp]))
            continue
        for flip in (0, 1):
            sel = (~near_wall) & (((ii + jj) % 2) == flip)
            c = qq[sel]
            if flip == 0:      # diagonal 0-2
                polys.append(("tri", np.concatenate([c[:, [0, 1, 2]], c[:, [0, 2, 3]]])))
            else:              # diagonal 1-3
                polys.append(("tri", np.concatenate([c[:, [0, 1, 3]], c[:, [1, 2, 3]]])))
    tri = np.concatenate([p for k, p in polys if k == "tri"])
    quad = np.concatenate([p for k, p in polys if k == "quad"])
    n_tri, n_quad = len(tri), len(quad)
    # cells: quads first, then triangles; geometry
    cell_nodes = [quad, tri]
    centroids, areas = [], []
    for cn in cell_nodes:
        P = xy[cn]                                           # [m, k, 2]
        Pn = np.roll(P, -1, axis=1)
        cross = P[..., 0] * Pn[..., 1] - Pn[..., 0] * P[..., 1]
        A = 0.5 * cross.sum(axis=1)
        cx = ((P[..., 0] + Pn[..., 0]) * cross).sum(axis=1) / (6.0 * A)
        cy = ((P[..., 1] + Pn[..., 1]) * cross).sum(axis=1) / (6.0 * A)
        centroids.append(np.stack([cx, cy], axis=1))
        areas.append(A)
    C = np.concatenate(centroids)
    V = np.concatenate(areas)
    assert np.all(V > 0), "inverted cell"
    n = len(C)
    # edges (a, b) with the cell on their left (ccw)
    ea, eb, ec = [], [], []
    off = 0
    for cn in cell_nodes:
        k = cn.shape[1]
        for e in range(k):
            ea.append(cn[:, e]); eb.append(cn[:, (e + 1) % k]); ec.append(off + np.arange(len(cn)))
        off += len(cn)
    ea = np.concatenate(ea); eb = np.concatenate(eb); ec = np.concatenate(ec)
    key = np.minimum(ea, eb).astype(np.int64) * (len(xy) + 1) + np.maximum(ea, eb)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.r_[True, ks[1:] != ks[:-1]]
    grp = np.cumsum(first) - 1
    counts = np.bincount(grp)
    assert counts.max() <= 2, "non-manifold edge"
    # internal faces: pairs
    pair_start = np.flatnonzero(first)[counts == 2]
    i0, i1 = order[pair_start], order[pair_start + 1]
    c0, c1 = ec[i0], ec[i1]
    own = np.minimum(c0, c1); nbr = np.maximum(c0, c1)
    # face geometry from the owner's edge orientation
    io = np.where(c0 == own, i0, i1)
    pa, pb = xy[ea[io]], xy[eb[io]]
    t = pb - pa
    Af = np.hypot(t[:, 0], t[:, 1])
    nrm = np.stack([t[:, 1], -t[:, 0]], axis=1) / Af[:, None]     # outward of the owner (left cell)
    xf = 0.5 * (pa + pb)
    # internal faces ordered by owner, then neighbour (the reference's face order)
    fo = np.lexsort((nbr, own))
    own, nbr, Af, nrm, xf = own[fo], nbr[fo], Af[fo], nrm[fo], xf[fo]
    d = C[nbr] - C[own]
    delta = np.abs((d * nrm).sum(axis=1))
    dn_ = np.abs(((C[nbr] - xf) * nrm).sum(axis=1))
    w = dn_ / np.maximum(delta, 1e-300)
    # boundary faces: the outlet (x = 30) gets the Dirichlet term
    single = np.flatnonzero(first)[counts == 1]
    ib = order[single]
    cb = ec[ib]
    pa, pb = xy[ea[ib]], xy[eb[ib]]
    xfb = 0.5 * (pa + pb)
    tb = pb - pa
    Ab = np.hypot(tb[:, 0], tb[:, 1])
    nb_ = np.stack([tb[:, 1], -tb[:, 0]], axis=1) / Ab[:, None]
    outlet = np.abs(xfb[:, 0] - 30.0) < 1e-9
    deltab = np.abs(((xfb - C[cb]) * nb_).sum(axis=1))
    # coefficients
    rng = np.random.default_rng(seed + 1)
    nut = np.exp(rng.normal(0.0, 1.0, n))
    nut_f = w * nut[own] + (1.0 - w) * nut[nbr]
    a = nut_f * Af / delta
    diag = np.zeros(n)
    np.add.at(diag, own, a)
    np.add.at(diag, nbr, a)
    np.add.at(diag, cb[outlet], nut[cb[outlet]] * Ab[outlet] / deltab[outlet])
    b = rng.normal(0.0, 1.0, n)
    b -= b.mean()
    return dict(n=n, nif=len(own), n_tri=n_tri, n_quad=n_quad, n_outlet=int(outlet.sum()),
                own=own.astype(np.int32), nbr=nbr.astype(np.int32), diag=diag, upper=-a, lower=-a.copy(),
                b=b, volume=V, centroid=C)


def csr_matrix(sys_):
    """The system as a scipy CSR matrix (tests and host-side checks)."""
    import scipy.sparse as sp
    n = sys_["n"]
    own, nbr = sys_["own"], sys_["nbr"]
    rows = np.concatenate([np.arange(n), own, nbr])
    cols = np.concatenate([np.arange(n), nbr, own])
    vals = np.concatenate([sys_["diag"], sys_["upper"], sys_["lower"]])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
