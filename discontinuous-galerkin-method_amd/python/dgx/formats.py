"""Text formats of the reference's inputs: MATLAB ``save -ascii`` matrices (16-byte fields
"%16.7e", one row per line), as written by input_files/mesh/create_mesh_file.m:21,34,49 and
precompute_matrices.m:35-44, and the Triangle-derived mesh files read by Mesh_Input (fun_DG.c:90-119):

* ``.node``  header (Point_Num, 2, 0, 1), then rows (id, x, z, marker);
* ``.ele``   header (Tri, 3, 1, 0, 0),   then rows (id, v1, v2, v3, attribute);
* ``.neigh`` header (Tri, 3, 0, 0),      then rows (id, n1, n2, n3), -1 = no neighbour.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import Mesh


def write_ascii(path: os.PathLike | str, a: np.ndarray) -> None:
    a = np.atleast_2d(np.asarray(a, np.float64))
    np.savetxt(path, a, fmt="%16.7e", delimiter="")


def write_mesh(prefix: os.PathLike | str, mesh: Mesh) -> tuple[str, str, str]:
    """Write <prefix>.node/.ele/.neigh; returns the three paths."""
    prefix = str(prefix)
    T, P = mesh.tri, mesh.point_num
    node = np.zeros((P + 1, 4))
    node[0] = [P, 2, 0, 1]
    node[1:, 0] = np.arange(1, P + 1)
    node[1:, 1:3] = mesh.point
    ele = np.zeros((T + 1, 5))
    ele[0] = [T, 3, 1, 0, 0]
    ele[1:, 0] = np.arange(1, T + 1)
    ele[1:, 1:4] = mesh.triangle
    ele[1:, 4] = mesh.attribute
    neigh = np.zeros((T + 1, 4))
    neigh[0, :2] = [T, 3]
    neigh[1:, 0] = np.arange(1, T + 1)
    neigh[1:, 1:4] = mesh.neighbor
    paths = (prefix + ".ele", prefix + ".node", prefix + ".neigh")
    write_ascii(paths[0], ele)
    write_ascii(paths[1], node)
    write_ascii(paths[2], neigh)
    return paths


def read_binary(path: os.PathLike | str) -> np.ndarray:
    return np.fromfile(path, dtype="<f8")


MESH_MAGIC = b"DGXMESH1"


def write_mesh_binary(path: os.PathLike | str, mesh: Mesh) -> str:
    """Binary mesh (dgx_write_mesh_binary's format): magic, int64 tri and point count, float64 points,
    int64 triangles, .neigh columns and attributes, little-endian."""
    with open(path, "wb") as f:
        f.write(MESH_MAGIC)
        np.array([mesh.tri, mesh.point_num], "<i8").tofile(f)
        np.ascontiguousarray(mesh.point, "<f8").tofile(f)
        np.ascontiguousarray(mesh.triangle, "<i8").tofile(f)
        np.ascontiguousarray(mesh.neighbor, "<i8").tofile(f)
        np.ascontiguousarray(mesh.attribute, "<i8").tofile(f)
    return str(path)


def read_mesh_binary(path: os.PathLike | str) -> Mesh:
    with open(path, "rb") as f:
        if f.read(8) != MESH_MAGIC:
            raise ValueError(f"{path} is not a binary mesh file")
        T, P = (int(v) for v in np.fromfile(f, "<i8", 2))
        point = np.fromfile(f, "<f8", 2 * P).reshape(P, 2)
        tri = np.fromfile(f, "<i8", 3 * T).reshape(T, 3)
        nb = np.fromfile(f, "<i8", 3 * T).reshape(T, 3)
        attr = np.fromfile(f, "<i8", T)
    if len(attr) != T:
        raise ValueError(f"{path} is truncated")
    return Mesh(point, tri, nb, attr)
