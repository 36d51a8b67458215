"""Seeded synthetic geometries in place of the reference's shapely + gmsh chain.

The reference turns polygons into a gmsh ``.geo`` file and meshes it in a
subprocess (openbte/geometry.py:442-554, openbte/mesh.py:354-478); gmsh is not
available here, so these generators produce the same ``Mesh`` contract
directly:

* ``pore_cell_2d``  -- periodic square cell with circular / square pores,
  unstructured triangles from a Delaunay triangulation of jittered lattice
  points.  Opposite edges carry identical node sets, as gmsh's ``Periodic Line``
  does, so every ``Periodic_x_a`` side has a partner in ``Periodic_x_b``.
* ``quad_cell_2d``  -- pore-free uniform quadrilateral periodic cell (the
  reference reads quads too, openbte/mesh.py:91,168-172); used for closed-form
  checks.
* ``structured_tri_cell_2d`` -- nx x ny pixels, 4 triangles per pixel (the
  differentiation workload's pixel grid).

Region names follow the reference's periodic convention: ``Periodic_x_a``
(x = -L/2), ``Periodic_x_b`` (x = +L/2, the kept side), likewise for y, and
``Boundary`` for pore walls.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np
from scipy.spatial import Delaunay

from .mesh import build_mesh
from .objects import Mesh

_REGIONS = ['Periodic_x_a', 'Periodic_x_b', 'Periodic_y_a', 'Periodic_y_b', 'Boundary']


def _edge_region(Lx: float, Ly: float):
    def classify(side_nodes_xy, centroids):
        tol = 1e-9 * max(Lx, Ly)
        c = centroids
        tag = np.full(len(c), 4, np.int64)
        tag[np.abs(c[:, 0] + Lx / 2) < tol] = 0
        tag[np.abs(c[:, 0] - Lx / 2) < tol] = 1
        tag[np.abs(c[:, 1] + Ly / 2) < tol] = 2
        tag[np.abs(c[:, 1] - Ly / 2) < tol] = 3
        return tag
    return classify


def _pore_nodes(kind: str, center, size: float, h: float) -> np.ndarray:
    cx, cy = center
    if kind == 'circle':
        r = size
        n = max(8, int(round(2 * np.pi * r / h)))
        th = 2 * np.pi * np.arange(n) / n
        return np.stack([cx + r * np.cos(th), cy + r * np.sin(th)], axis=1)
    if kind == 'square':
        a = size  # half side
        n = max(2, int(round(2 * a / h)))
        t = np.linspace(-a, a, n + 1)[:-1]
        bottom = np.stack([cx + t, np.full(n, cy - a)], axis=1)
        right = np.stack([np.full(n, cx + a), cy + t], axis=1)
        top = np.stack([cx - t, np.full(n, cy + a)], axis=1)
        left = np.stack([np.full(n, cx - a), cy - t], axis=1)
        return np.concatenate([bottom, right, top, left])
    raise ValueError(kind)


def _inside(kind: str, center, size: float, pts: np.ndarray, pad: float = 0.0) -> np.ndarray:
    d = pts - np.asarray(center)[None, :]
    if kind == 'circle':
        return np.hypot(d[:, 0], d[:, 1]) < size + pad
    return (np.abs(d[:, 0]) < size + pad) & (np.abs(d[:, 1]) < size + pad)


def pore_area(kind: str, size: float) -> float:
    return np.pi * size ** 2 if kind == 'circle' else (2 * size) ** 2


def pore_cell_2d(L: float = 10.0,
                 porosity: float = 0.25,
                 n_cells: int = 2000,
                 seed: int = 0,
                 pores: Optional[Sequence[Tuple[str, Tuple[float, float], float]]] = None,
                 jitter: float = 0.3) -> Mesh:
    """Periodic square cell of side ``L`` (nm) with pores, ~``n_cells`` triangles.

    ``pores`` is a list of ``(kind, (cx, cy), size)`` with kind ``'circle'``
    (size = radius) or ``'square'`` (size = half side); by default one centred
    circle of the given porosity (r = L sqrt(porosity/pi)).  Lattice points are
    jittered by U(-jitter h, +jitter h) with ``numpy.random.default_rng(seed)``.
    """
    rng = np.random.default_rng(seed)
    if pores is None:
        pores = [] if porosity <= 0 else [('circle', (0.0, 0.0), L * np.sqrt(porosity / np.pi))]
    solid = L * L - sum(pore_area(k, s) for k, _, s in pores)
    h_target = np.sqrt(2.0 * solid / n_cells)
    N = max(2, int(round(L / h_target)))
    h = L / N
    xs = -L / 2 + h * np.arange(N + 1)
    # boundary nodes: identical node sets on opposite edges
    edge = [np.stack([xs, np.full(N + 1, -L / 2)], 1), np.stack([xs, np.full(N + 1, L / 2)], 1),
            np.stack([np.full(N - 1, -L / 2), xs[1:-1]], 1), np.stack([np.full(N - 1, L / 2), xs[1:-1]], 1)]
    bnodes = np.concatenate(edge)
    g = xs[1:-1]
    X, Y = np.meshgrid(g, g, indexing='ij')
    inner = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-jitter * h, jitter * h, (len(g) ** 2, 2))
    keep = np.ones(len(inner), bool)
    pnodes = []
    for kind, c, s in pores:
        keep &= ~_inside(kind, c, s, inner, pad=0.6 * h)
        pnodes.append(_pore_nodes(kind, c, s, h))
        lo = np.array(c) - s
        hi = np.array(c) + s
        if np.any(lo < -L / 2 + 1.5 * h) or np.any(hi > L / 2 - 1.5 * h):
            raise ValueError("pore too close to the cell edge")
    pts = np.concatenate([bnodes, inner[keep]] + pnodes)
    tri = Delaunay(pts)
    if len(tri.coplanar):
        raise RuntimeError("Delaunay dropped points")
    simp = tri.simplices.astype(np.int64)
    cen = pts[simp].mean(axis=1)
    ok = np.ones(len(simp), bool)
    for kind, c, s in pores:
        ok &= ~_inside(kind, c, s, cen)
    simp = simp[ok]
    # counter-clockwise orientation and no degenerate triangles
    a = pts[simp]
    det = (a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) - (a[:, 1, 1] - a[:, 0, 1]) * (a[:, 2, 0] - a[:, 0, 0])
    if np.any(np.abs(det) < 1e-10 * h * h):
        raise RuntimeError("degenerate triangle")
    flip = det < 0
    simp[flip] = simp[flip][:, [0, 2, 1]]
    used = np.unique(simp)
    remap = -np.ones(len(pts), np.int64)
    remap[used] = np.arange(len(used))
    pts = pts[used]
    simp = remap[simp]
    mesh = build_mesh(pts, simp, _edge_region(L, L), _REGIONS,
                      periodic={'Periodic_x': np.array([L, 0.0]), 'Periodic_y': np.array([0.0, L])})
    # every wall side must lie on a pore perimeter
    if 'Boundary' in mesh.side_physical_regions:
        wc = mesh.side_centroids[mesh.side_physical_regions['Boundary']]
        near = np.zeros(len(wc), bool)
        for kind, c, s in pores:
            near |= _inside(kind, c, s, wc, pad=0.1 * h) & ~_inside(kind, c, s, wc, pad=-0.2 * h)
        if not np.all(near):
            raise RuntimeError("wall side away from a pore perimeter")
    return mesh


def quad_cell_2d(L: float = 10.0, N: int = 20, Ly: Optional[float] = None, Ny: Optional[int] = None) -> Mesh:
    """Uniform N x Ny quadrilateral periodic cell, no pores (closed-form checks)."""
    Ly = L if Ly is None else Ly
    Ny = N if Ny is None else Ny
    xs = np.linspace(-L / 2, L / 2, N + 1)
    ys = np.linspace(-Ly / 2, Ly / 2, Ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.arange((N + 1) * (Ny + 1)).reshape(N + 1, Ny + 1)
    q = np.stack([idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]], axis=-1).reshape(-1, 4)
    return build_mesh(pts, q, _edge_region(L, Ly), _REGIONS,
                      periodic={'Periodic_x': np.array([L, 0.0]), 'Periodic_y': np.array([0.0, Ly])})


def structured_tri_cell_2d(L: float = 10.0, N: int = 40) -> Mesh:
    """N x N pixels of side L/N, each split into 4 triangles by its centre.

    Cell ``4*p + k`` belongs to pixel ``p = i*N + j`` (x index i, y index j).
    """
    h = L / N
    xs = -L / 2 + h * np.arange(N + 1)
    X, Y = np.meshgrid(xs, xs, indexing='ij')
    corners = np.stack([X.ravel(), Y.ravel()], 1)
    ci = np.arange(N)
    CX, CY = np.meshgrid(-L / 2 + h * (ci + 0.5), -L / 2 + h * (ci + 0.5), indexing='ij')
    centres = np.stack([CX.ravel(), CY.ravel()], 1)
    pts = np.concatenate([corners, centres])
    idx = np.arange((N + 1) ** 2).reshape(N + 1, N + 1)
    c = (N + 1) ** 2 + np.arange(N * N).reshape(N, N)
    a, b, cc, d = idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]
    tris = np.stack([np.stack([a, b, c], -1), np.stack([b, cc, c], -1),
                     np.stack([cc, d, c], -1), np.stack([d, a, c], -1)], axis=2).reshape(-1, 3)
    return build_mesh(pts, tris, _edge_region(L, L), _REGIONS,
                      periodic={'Periodic_x': np.array([L, 0.0]), 'Periodic_y': np.array([0.0, L])})


def random_pore_cell_2d(seed: int, n_cells: int, h_ref: float = 0.274, max_pores: int = 6,
                        porosity_range=(0.05, 0.40)) -> Mesh:
    """Periodic cell with 1..max_pores non-overlapping circular / square pores.

    Total porosity ~ U(porosity_range), split among the pores by a Dirichlet
    draw; the cell side grows with the cell count so the mesh spacing stays
    near ``h_ref`` nm.  Placement is rejection-sampled with a clearance of 3
    mesh spacings between pores and from the cell edges.
    """
    rng = np.random.default_rng(seed)
    for _attempt in range(200):
        k = int(rng.integers(1, max_pores + 1))
        phi = float(rng.uniform(*porosity_range))
        L = float(np.sqrt(n_cells * h_ref ** 2 / (2.0 * (1.0 - phi))))
        h = np.sqrt(2.0 * L * L * (1 - phi) / n_cells)
        shares = rng.dirichlet(np.ones(k)) * phi * L * L
        kinds = ['circle' if rng.random() < 0.5 else 'square' for _ in range(k)]
        sizes = [np.sqrt(a / np.pi) if kd == 'circle' else 0.5 * np.sqrt(a) for kd, a in zip(kinds, shares)]
        radius = [s if kd == 'circle' else s * np.sqrt(2.0) for kd, s in zip(kinds, sizes)]
        centres = []
        ok = True
        for r in radius:
            lo = -L / 2 + r + 3 * h
            hi = L / 2 - r - 3 * h
            if lo >= hi:
                ok = False
                break
            for _ in range(500):
                c = rng.uniform(lo, hi, 2)
                if all(np.hypot(*(c - c2)) > r + r2 + 3 * h for c2, r2 in zip(centres, radius)):
                    centres.append(c)
                    break
            else:
                ok = False
                break
        if not ok:
            continue
        pores = [(kd, (float(c[0]), float(c[1])), float(s)) for kd, c, s in zip(kinds, centres, sizes)]
        try:
            return pore_cell_2d(L, phi, n_cells, seed=seed, pores=pores)
        except (RuntimeError, ValueError):
            continue
    raise RuntimeError(f'could not place pores for seed {seed}')


def _periodic_canonical(nodes: np.ndarray, L: float) -> np.ndarray:
    """Index of the node each node maps to when x = +L/2 and y = +L/2 are
    wrapped onto -L/2 (periodic partners share one canonical node)."""
    tol = 1e-9 * L
    wrapped = nodes.copy()
    wrapped[np.abs(wrapped[:, 0] - L / 2) < tol, 0] -= L
    wrapped[np.abs(wrapped[:, 1] - L / 2) < tol, 1] -= L
    key = np.round(wrapped / (1e-7 * L)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    return first[inv.reshape(-1)]


def _prism_tets(v: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Split prisms (bottom v0 v1 v2, top v3 v4 v5 with v[i+3] above v[i]) into
    3 tetrahedra each.  Every quad face is cut along the diagonal through its
    smallest-key vertex, so neighbouring prisms (and periodic partners, whose
    keys agree) cut shared faces the same way (Dompierre et al., "How to
    subdivide pyramids, prisms and hexahedra into tetrahedra", 1999)."""
    # the 6 rotations / reflections that bring each vertex to position 0
    perms = np.array([[0, 1, 2, 3, 4, 5], [1, 2, 0, 4, 5, 3], [2, 0, 1, 5, 3, 4],
                      [3, 5, 4, 0, 2, 1], [4, 3, 5, 1, 0, 2], [5, 4, 3, 2, 1, 0]])
    kv = key[v]
    w = v[np.arange(len(v))[:, None], perms[np.argmin(kv, axis=1)]]
    kw = key[w]
    a = np.minimum(kw[:, 1], kw[:, 5]) < np.minimum(kw[:, 2], kw[:, 4])
    t = np.empty((len(v), 3, 4), np.int64)
    c0, c1, c2, c3, c4, c5 = (w[:, i] for i in range(6))
    t[a, 0] = np.stack([c0, c1, c2, c5], 1)[a]
    t[a, 1] = np.stack([c0, c1, c5, c4], 1)[a]
    t[~a, 0] = np.stack([c0, c1, c2, c4], 1)[~a]
    t[~a, 1] = np.stack([c0, c4, c2, c5], 1)[~a]
    t[:, 2] = np.stack([c0, c4, c5, c3], 1)
    return t.reshape(-1, 4)


def pore_slab_3d(L: float = 10.0, H: float = 5.0, porosity: float = 0.25, n_cells: int = 20000,
                 seed: int = 0, nz: Optional[int] = None) -> Mesh:
    """Periodic porous slab: the ``pore_cell_2d`` cross-section (side ``L``
    nm, centred cylindrical pore of the given porosity) extruded over a
    thickness ``H`` in ``nz`` layers, each triangular prism split into 3
    tetrahedra.  x and y are periodic (``Periodic_x``, ``Periodic_y``); the
    pore wall and the two slab faces z = 0, z = H form the diffuse
    ``Boundary``.  ``n_cells`` is the target tetrahedron count; by default the
    layer thickness follows the in-plane mesh spacing."""
    solid = L * L * (1.0 - porosity)
    if nz is None:
        # 3 n2 nz = n_cells with n2 = 2 solid / h^2 and nz = H / h
        h = (6.0 * solid * H / n_cells) ** (1.0 / 3.0)
        nz = max(1, int(round(H / h)))
    n2 = max(16, int(round(n_cells / (3 * nz))))
    base = pore_cell_2d(L, porosity, n2, seed=seed)
    p2, t2 = base.nodes, base.elems
    m = len(p2)
    canon = _periodic_canonical(p2, L)
    zs = np.linspace(0.0, H, nz + 1)
    nodes = np.concatenate([np.concatenate([p2, np.full((m, 1), z)], 1) for z in zs])
    key = (np.tile(canon, nz + 1) * (nz + 1) + np.repeat(np.arange(nz + 1), m)).astype(np.int64)
    k = np.arange(nz)[:, None, None]
    bottom = t2[None, :, :] + k * m
    prisms = np.concatenate([bottom, bottom + m], axis=2).reshape(-1, 6)
    tets = _prism_tets(prisms, key)

    def classify(side_nodes, centroids):
        tol = 1e-9 * max(L, H)
        c = centroids
        tag = np.full(len(c), 4, np.int64)
        tag[np.abs(c[:, 0] + L / 2) < tol] = 0
        tag[np.abs(c[:, 0] - L / 2) < tol] = 1
        tag[np.abs(c[:, 1] + L / 2) < tol] = 2
        tag[np.abs(c[:, 1] - L / 2) < tol] = 3
        return tag

    return build_mesh(nodes, tets, classify, _REGIONS,
                      periodic={'Periodic_x': np.array([L, 0.0, 0.0]), 'Periodic_y': np.array([0.0, L, 0.0])})
