"""Triangle meshes in the reference's in-memory layout (Mesh_Input, fun_DG.c:90-119).

The reference meshes with the external Triangle program (input_files/mesh/create_mesh_file.m:5), which
is not available here; these generators produce the same arrays:

* ``point``     [Point_Num, 2] float64 vertex (x, z), z pointing up;
* ``triangle``  [Tri, 3] int64 1-based vertex ids, counter-clockwise (fun_DG.c:447 needs J > 0);
* ``neighbor``  [Tri, 3] int64 1-based element opposite vertex 0, 1, 2 (Triangle's .neigh), -1 = none;
* ``attribute`` [Tri] int64 region attribute (.ele column 5).

Every generator rotates each element's vertex list (keeping it counter-clockwise) so that its first
vertex lies outside the sponge box whenever any vertex does; the reference's absorbing-element list is
built from the first vertex only (fun_DG.c:202-281) and is only complete for such meshes.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class Mesh:
    point: np.ndarray
    triangle: np.ndarray
    neighbor: np.ndarray
    attribute: np.ndarray

    @property
    def tri(self) -> int:
        return int(self.triangle.shape[0])

    @property
    def point_num(self) -> int:
        return int(self.point.shape[0])

    def vertices(self) -> np.ndarray:
        """[Tri, 3, 2] vertex coordinates."""
        return self.point[self.triangle - 1]

    def centroids(self) -> np.ndarray:
        return self.vertices().mean(axis=1)

    def area2(self) -> np.ndarray:
        v = self.vertices()
        return (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 2, 0] - v[:, 0, 0]) * (
            v[:, 1, 1] - v[:, 0, 1])


def compute_neighbors(triangle: np.ndarray) -> np.ndarray:
    """Triangle's .neigh convention: column c = element sharing the edge opposite vertex c (1-based)."""
    t = np.asarray(triangle, dtype=np.int64)
    n = t.shape[0]
    a = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])  # edge opposite v0 is (v1,v2), opposite v1 is (v2,v0) ...
    b = np.concatenate([t[:, 2], t[:, 0], t[:, 1]])
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(t.max()) + 1) + hi
    owner = np.concatenate([np.arange(n)] * 3)
    col = np.repeat(np.arange(3), n)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    nb = np.full((n, 3), -1, dtype=np.int64)
    same = ks[1:] == ks[:-1]
    i0 = order[:-1][same]
    i1 = order[1:][same]
    if len(same) > 1 and np.any(same[1:] & same[:-1]):
        raise ValueError("an edge is shared by more than two elements")
    nb[owner[i0], col[i0]] = owner[i1] + 1
    nb[owner[i1], col[i1]] = owner[i0] + 1
    return nb


def rotate_first_vertex_outside(point: np.ndarray, triangle: np.ndarray, box: tuple[float, float, float, float]):
    """Rotate each vertex list (v0,v1,v2) -> (v1,v2,v0) / (v2,v0,v1) so v0 is outside ``box``
    (xmin, xmax, zmin, zmax, strict) whenever any vertex is; orientation is preserved."""
    x0, x1, z0, z1 = box
    p = point[triangle - 1]
    out = (p[..., 0] > x1) | (p[..., 0] < x0) | (p[..., 1] > z1) | (p[..., 1] < z0)
    shift = np.where(out[:, 0], 0, np.where(out[:, 1], 1, np.where(out[:, 2], 2, 0)))
    idx = (np.arange(3)[None, :] + shift[:, None]) % 3
    return np.take_along_axis(triangle, idx, axis=1)


def structured(nx: int, nz: int, h: float, *, x0: float = 0.0, z0: float = 0.0, diagonal: str | np.ndarray = "same",
               jitter: float = 0.0, seed: int = 1, sponge_box=None, attribute=None) -> Mesh:
    """``nx`` x ``nz`` squares of side ``h``, each split into two counter-clockwise triangles.

    diagonal: "same" (every square split along its (i,j)-(i+1,j+1) diagonal) or "random" (seeded choice
    per square) or a bool array [nz, nx] (True = the other diagonal).  jitter: interior vertices move by
    U(-jitter, +jitter) in x and z; vertices on the outer boundary do not move.
    """
    rng = np.random.default_rng(seed)
    xs = x0 + h * np.arange(nx + 1, dtype=np.float64)
    zs = z0 + h * np.arange(nz + 1, dtype=np.float64)
    X, Z = np.meshgrid(xs, zs)  # [nz+1, nx+1], row j = z index
    if jitter > 0:
        dx = rng.uniform(-jitter, jitter, size=X.shape)
        dz = rng.uniform(-jitter, jitter, size=X.shape)
        interior = np.zeros(X.shape, dtype=bool)
        interior[1:-1, 1:-1] = True
        dx[~interior] = 0.0
        dz[~interior] = 0.0
        X = X + dx
        Z = Z + dz
    point = np.stack([X.ravel(), Z.ravel()], axis=1)
    vid = np.arange((nx + 1) * (nz + 1), dtype=np.int64).reshape(nz + 1, nx + 1) + 1
    a = vid[:-1, :-1]
    b = vid[:-1, 1:]
    c = vid[1:, 1:]
    d = vid[1:, :-1]
    if isinstance(diagonal, str):
        if diagonal == "same":
            flip = np.zeros((nz, nx), dtype=bool)
        elif diagonal == "random":
            flip = rng.random((nz, nx)) < 0.5
        else:
            raise ValueError(diagonal)
    else:
        flip = np.asarray(diagonal, dtype=bool)
    t1 = np.where(flip[..., None], np.stack([a, b, d], -1), np.stack([a, b, c], -1))
    t2 = np.where(flip[..., None], np.stack([b, c, d], -1), np.stack([a, c, d], -1))
    triangle = np.stack([t1, t2], axis=2).reshape(-1, 3)
    if sponge_box is not None:
        triangle = rotate_first_vertex_outside(point, triangle, sponge_box)
    neighbor = compute_neighbors(triangle)
    if attribute is None:
        attr = np.ones(triangle.shape[0], dtype=np.int64)
    elif callable(attribute):
        attr = np.asarray(attribute(point[triangle - 1].mean(axis=1)), dtype=np.int64)
    else:
        attr = np.full(triangle.shape[0], int(attribute), dtype=np.int64)
    m = Mesh(point, triangle.astype(np.int64), neighbor, attr)
    if np.any(m.area2() <= 0):
        raise ValueError("generated a non counter-clockwise element")
    return m


def sponge_box_of(width: float, height: float, bound_x: float, bound_z: float):
    return (bound_x, width - bound_x, bound_z, height - bound_z)
