"""Unstructured triangle meshes from seeded point sets (Delaunay, scipy/Qhull).

The reference meshes with the external Triangle program (constrained Delaunay,
input_files/mesh/create_mesh_file.m:5), which is not available here.  These generators place points on
the material interfaces at the local spacing and remove the background points that would crowd them,
then triangulate the point set; with interface points spaced at the background spacing the Delaunay
triangulation contains the interface segments almost everywhere, and `interface_conformity` reports the
fraction that is present.  Elements take their material at the centroid.  All triangles are returned
counter-clockwise in the reference's array conventions (see mesh.py).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import Delaunay

from .mesh import Mesh, compute_neighbors, rotate_first_vertex_outside


def _segments_points(p0, p1, h):
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    n = max(1, int(np.ceil(np.linalg.norm(p1 - p0) / h)))
    t = np.linspace(0.0, 1.0, n + 1)
    return p0[None, :] + t[:, None] * (p1 - p0)[None, :]


def _dist_to_polyline(pts, line):
    """Distance from points [N,2] to a polyline [M,2] (vectorised over segments in chunks)."""
    d = np.full(len(pts), np.inf)
    a, b = line[:-1], line[1:]
    ab = b - a
    L2 = np.maximum((ab ** 2).sum(1), 1e-30)
    for k in range(0, len(a), 256):
        aa, vv, ll = a[k:k + 256], ab[k:k + 256], L2[k:k + 256]
        ap = pts[:, None, :] - aa[None, :, :]
        t = np.clip((ap * vv[None]).sum(-1) / ll[None], 0.0, 1.0)
        q = aa[None] + t[..., None] * vv[None]
        d = np.minimum(d, np.sqrt(((pts[:, None, :] - q) ** 2).sum(-1)).min(1))
    return d


def triangulate(points: np.ndarray, *, sponge_box=None, attribute=None, min_area_frac: float = 1e-6) -> Mesh:
    pts = np.ascontiguousarray(points, np.float64)
    tri = Delaunay(pts, qhull_options="Qbb Qc").simplices.astype(np.int64)
    v = pts[tri]
    a2 = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1])
    flip = a2 < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    a2 = np.abs(a2)
    keep = a2 > min_area_frac * np.median(a2)  # drop Qhull slivers on collinear boundary points
    tri = tri[keep] + 1
    if sponge_box is not None:
        tri = rotate_first_vertex_outside(pts, tri, sponge_box)
    nb = compute_neighbors(tri)
    cen = pts[tri - 1].mean(axis=1)
    attr = np.ones(len(tri), np.int64) if attribute is None else np.asarray(attribute(cen), np.int64)
    return Mesh(pts, tri, nb, attr)


def interface_conformity(mesh: Mesh, polyline: np.ndarray, tol: float = 1e-6) -> float:
    """Fraction of the polyline's segments that are mesh edges (vertices matched within tol)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(mesh.point)
    d, idx = tree.query(polyline)
    ok = d < tol
    edges = set()
    t = mesh.triangle - 1
    for a, b in ((0, 1), (1, 2), (2, 0)):
        lo, hi = np.minimum(t[:, a], t[:, b]), np.maximum(t[:, a], t[:, b])
        edges.update(zip(lo.tolist(), hi.tolist()))
    hits = 0
    for k in range(len(polyline) - 1):
        if ok[k] and ok[k + 1]:
            i, j = sorted((int(idx[k]), int(idx[k + 1])))
            hits += (i, j) in edges
    return hits / max(1, len(polyline) - 1)


def layered_point_set(width, height, h, interfaces, *, seed=1, jitter=0.25, clearance=0.45):
    """Jittered background grid of spacing h, minus points within clearance*h of any interface, plus
    points every h along the interfaces and the domain boundary."""
    rng = np.random.default_rng(seed)
    nx, nz = int(round(width / h)), int(round(height / h))
    hx, hz = width / nx, height / nz
    X, Z = np.meshgrid(np.arange(1, nx) * hx, np.arange(1, nz) * hz)
    P = np.stack([X.ravel(), Z.ravel()], 1)
    P += rng.uniform(-jitter, jitter, P.shape) * np.array([hx, hz])
    lines = []
    for f in interfaces:
        xs = np.linspace(0.0, width, int(np.ceil(width / h)) + 1)
        line = np.stack([xs, f(xs)], 1)
        lines.append(line)
        near = np.abs(P[:, 1] - f(P[:, 0])) < clearance * h * 1.5
        if np.any(near):
            d = _dist_to_polyline(P[near], line)
            drop = np.zeros(len(P), bool)
            drop[np.nonzero(near)[0][d < clearance * h]] = True
            P = P[~drop]
    bx = np.linspace(0.0, width, nx + 1)
    bz = np.linspace(0.0, height, nz + 1)
    border = np.concatenate([np.stack([bx, np.zeros_like(bx)], 1), np.stack([bx, np.full_like(bx, height)], 1),
                             np.stack([np.zeros_like(bz[1:-1]), bz[1:-1]], 1),
                             np.stack([np.full_like(bz[1:-1], width), bz[1:-1]], 1)])
    iface = [ln[1:-1] for ln in lines]
    # interface end points lie on the side boundaries: add them there too
    ends = np.concatenate([ln[[0, -1]] for ln in lines]) if lines else np.zeros((0, 2))
    pts = np.concatenate([P, border, ends] + iface)
    pts = np.unique(np.round(pts, 9), axis=0)
    return pts, lines
