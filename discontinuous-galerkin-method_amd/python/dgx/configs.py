"""The workloads the solver is measured on, built from seeded generators.

Coordinates: x to the right, z up (as in the reference's slope_model.poly); depth = H - z.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import lib as L
from .mesh import Mesh, structured
from .problem import Model, SourceSpec, Sponge


@dataclasses.dataclass
class Workload:
    name: str
    mesh: Mesh
    model: Model
    sponge: Sponge
    sources: list
    receivers: np.ndarray
    nsteps: int
    snapshot_every: int = 0


def homogeneous_2km(nsteps: int = 1500, *, n: int = 100, h: float = 20.0) -> Workload:
    """Homogeneous elastic 2 km x 2 km: 100x100 squares of 20 m split along the same diagonal (20,000
    triangles); Vp 3000, Vs 1732, rho 2000; explosive 15 Hz Ricker at x=1000 m, depth 500 m; 101
    receivers at 10 m depth every 20 m; sponge 100 m on all sides; snapshots every 250 steps."""
    W = H = n * h
    sp = Sponge.for_domain(W, H, 100.0)
    mesh = structured(n, n, h, diagonal="same", sponge_box=sp.box(), attribute=3)
    model = Model.homogeneous(mesh.tri, 3000.0, 1732.0, 2000.0)
    src = [SourceSpec(1000.0, H - 500.0, L.DGX_SRC_EXPLOSIVE, 0, 15.0)]
    rx = np.stack([np.linspace(0.0, W, 101), np.full(101, H - 10.0)], axis=1)
    return Workload("homogeneous_2km_20k_tri", mesh, model, sp, src, rx, nsteps, 250)


def gradient_8x4km(nsteps: int = 4000, *, nx: int = 800, nz: int = 400, h: float = 10.0, seed: int = 1,
                   source_x: float = 4000.0, source_depth: float = 500.0) -> Workload:
    """8 km x 4 km, 800x400 squares of 10 m, seeded random diagonal per square, interior vertices
    jittered by U(-2.5, 2.5) m (640,000 triangles).  Vp = 2000 + 2500 depth/4000 (+800 below 1.5 km),
    Vs = Vp/sqrt(3), rho = 310 Vp^0.25 at element centroids; 20 Hz explosive Ricker at (4000 m, 500 m
    depth); 401 surface receivers every 20 m."""
    W, H = nx * h, nz * h
    sp = Sponge.for_domain(W, H, 100.0)
    mesh = structured(nx, nz, h, diagonal="random", jitter=0.25 * h, seed=seed, sponge_box=sp.box(), attribute=3)
    depth = H - mesh.centroids()[:, 1]
    vp = 2000.0 + 2500.0 * depth / 4000.0 + np.where(depth >= 1500.0, 800.0, 0.0)
    model = Model.from_velocities(vp, vp / np.sqrt(3.0), 310.0 * vp ** 0.25)
    src = [SourceSpec(source_x, H - source_depth, L.DGX_SRC_EXPLOSIVE, 0, 20.0)]
    rx = np.stack([np.linspace(0.0, W, 401), np.full(401, H)], axis=1)
    return Workload("gradient_8x4km_640k_tri", mesh, model, sp, src, rx, nsteps, 0)
