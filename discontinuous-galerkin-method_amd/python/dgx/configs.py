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


def slope_6x3km(nsteps: int = 5000, *, h: float = 5.48, nshots: int = 64, nrec: int = 2001, seed: int = 1) -> Workload:
    """6 km x 3 km, flat sea surface at z = 3 km over a 20-degree planar sea floor (z = 2800 - x tan20),
    water (1500, 0, 1000) above three dipping elastic layers (Vp 2500/3500/4500, Vs = Vp/1.8, rho
    2100/2300/2500) with a coupled acoustic-elastic interface.  Delaunay mesh of a seeded point set with
    the interfaces inserted at the background spacing h (h = 5.48 m gives ~1.2 M triangles).  nshots
    explosive 25 Hz Ricker shots 10 m below the surface, evenly spaced; nrec shared surface receivers."""
    from .unstructured import layered_point_set, triangulate
    W, H = 6000.0, 3000.0
    t20 = np.tan(np.radians(20.0))

    def seafloor(x):
        return 2800.0 - t20 * x

    def i1(x):
        return 2400.0 - 0.30 * x

    def i2(x):
        return 1600.0 - 0.20 * x

    sp = Sponge.for_domain(W, H, 100.0)
    pts, _ = layered_point_set(W, H, h, [seafloor, i1, i2], seed=seed)

    def attr(c):
        x, z = c[:, 0], c[:, 1]
        return np.where(z > seafloor(x), 1, np.where(z > i1(x), 3, np.where(z > i2(x), 4, 5)))
    mesh = triangulate(pts, sponge_box=sp.box(), attribute=attr)
    a = mesh.attribute
    vp = np.select([a == 1, a == 3, a == 4], [1500.0, 2500.0, 3500.0], 4500.0)
    vs = np.where(a == 1, 0.0, vp / 1.8)
    rho = np.select([a == 1, a == 3, a == 4], [1000.0, 2100.0, 2300.0], 2500.0)
    model = Model.from_velocities(vp, vs, rho)
    xs = np.linspace(150.0, W - 150.0, nshots)
    src = [SourceSpec(float(x), H - 10.0, L.DGX_SRC_EXPLOSIVE, k, 25.0) for k, x in enumerate(xs)]
    rx = np.stack([np.linspace(0.0, W, nrec), np.full(nrec, H - 10.0)], axis=1)
    return Workload(f"slope_6x3km_{mesh.tri // 1000}k_tri_{nshots}_shots", mesh, model, sp, src, rx, nsteps, 0)


SERVING_MESH_TRIANGLES = (20_000, 50_000, 120_000, 250_000, 500_000, 800_000, 1_200_000, 2_000_000)


def layered_random_model(mesh: Mesh, height: float, rng) -> Model:
    """3-6 horizontal layers at random depths, Vp in 1800-5000 m/s increasing with depth,
    Vs = Vp/sqrt(3), Gardner density 310 Vp^0.25."""
    n = int(rng.integers(3, 7))
    tops = np.sort(rng.uniform(0.0, height, n - 1))
    vps = np.sort(rng.uniform(1800.0, 5000.0, n))
    depth = height - mesh.centroids()[:, 1]
    layer = np.searchsorted(tops, depth)
    vp = vps[layer]
    return Model.from_velocities(vp, vp / np.sqrt(3.0), 310.0 * vp ** 0.25)


def serving_meshes(sizes=SERVING_MESH_TRIANGLES, *, h: float = 10.0, seed: int = 1):
    """The resident mesh set: jittered structured meshes (random diagonal per square, interior vertices
    moved by +-25 % of h) of the given triangle counts, near-square domains, each with its own random
    3-6 layer model.  Returns [(mesh, model, sponge)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        half = n / 2.0
        nx = int(round(np.sqrt(half * 1.25)))
        nz = int(round(half / nx))
        W, H = nx * h, nz * h
        sp = Sponge.for_domain(W, H, 100.0)
        mesh = structured(nx, nz, h, diagonal="random", jitter=0.25 * h, seed=seed + k, sponge_box=sp.box(),
                          attribute=3)
        out.append((mesh, layered_random_model(mesh, H, rng), sp))
    return out


def serving_requests(meshes, n: int = 2000, *, rate: float = 20.0, seed: int = 3, steps=(1000, 6000),
                     shots=(1, 16), f0=(10.0, 30.0), nrec: int = 51):
    """Seeded open-loop request stream: Poisson arrivals at `rate` per second; each request picks a
    resident mesh uniformly, 1-16 explosive shots at random interior positions (100-600 m deep),
    one Ricker frequency in 10-30 Hz, 1000-6000 steps, and `nrec` receivers along the surface.
    Returns (arrival times [n], list of request specs)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    specs = []
    for _ in range(n):
        mi = int(rng.integers(0, len(meshes)))
        mesh = meshes[mi][0]
        W, H = mesh.point[:, 0].max(), mesh.point[:, 1].max()
        ns = int(rng.integers(shots[0], shots[1] + 1))
        xs = rng.uniform(150.0, W - 150.0, ns)
        zs = H - rng.uniform(100.0, min(600.0, H - 200.0), ns)
        specs.append(dict(mesh=mi, xz=np.stack([xs, zs], 1), f0=float(rng.uniform(*f0)),
                          nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                          rec=np.stack([np.linspace(120.0, W - 120.0, nrec), np.full(nrec, H - 5.0)], 1)))
    return arrivals, specs


def carbonate_4x3km(nsteps: int = 8000, *, nvugs: int = 300, nfrac: int = 50, h_min: float = 1.0,
                    h_max: float = 40.0, growth: float = 0.1, nshots: int = 256, seed: int = 7) -> Workload:
    """Fractured-vuggy carbonate, 4 km x 3 km: background (5500, 3000, 2650); nvugs fluid-filled circular
    vugs (r 5-40 m; 1500, 0, 1000); nfrac thin fractures (50-300 m long, 2 m aperture; 3000, 1200, 2300).
    Mesh: Delaunay of a graded point set, spacing h_min at the feature boundaries growing by `growth`
    per metre of distance up to h_max (1:40), with points on every vug circle and fracture wall.
    nshots 30 Hz explosive shots along a line 50 m deep; dt from the smallest element."""
    from scipy import ndimage

    from .unstructured import triangulate
    rng = np.random.default_rng(seed)
    W, H = 4000.0, 3000.0
    sp = Sponge.for_domain(W, H, 100.0)
    vugs = []
    while len(vugs) < nvugs:
        r = rng.uniform(5.0, 40.0)
        c = rng.uniform([300.0 + r, 300.0 + r], [W - 300.0 - r, H - 300.0 - r])
        if all(np.hypot(*(c - v[:2])) > r + v[2] + 5.0 for v in vugs):
            vugs.append((c[0], c[1], r))
    vugs = np.array(vugs)
    fr = []
    for _ in range(nfrac):
        L_ = rng.uniform(50.0, 300.0)
        a = rng.uniform(0.0, np.pi)
        c = rng.uniform([400.0, 400.0], [W - 400.0, H - 400.0])
        d = 0.5 * L_ * np.array([np.cos(a), np.sin(a)])
        fr.append((c - d, c + d))
    # distance to the nearest feature boundary on a 1 m raster (0.5 m cells)
    cell = 0.5
    nx, nz = int(W / cell), int(H / cell)
    X = (np.arange(nx) + 0.5) * cell
    Z = (np.arange(nz) + 0.5) * cell
    mask = np.zeros((nz, nx), bool)
    for (cx, cz, r) in vugs:
        i0, i1 = int(max(0, (cx - r - 2) / cell)), int(min(nx, (cx + r + 2) / cell))
        j0, j1 = int(max(0, (cz - r - 2) / cell)), int(min(nz, (cz + r + 2) / cell))
        xx, zz = np.meshgrid(X[i0:i1], Z[j0:j1])
        mask[j0:j1, i0:i1] |= np.abs(np.hypot(xx - cx, zz - cz) - r) < cell
    for p0, p1 in fr:
        n = int(np.hypot(*(p1 - p0)) / (cell / 2)) + 1
        for off in (-1.0, 1.0):
            nrm = np.array([-(p1 - p0)[1], (p1 - p0)[0]]) / np.hypot(*(p1 - p0))
            pts = p0[None] + np.linspace(0, 1, n)[:, None] * (p1 - p0)[None] + off * nrm[None]
            ii = np.clip((pts[:, 0] / cell).astype(int), 0, nx - 1)
            jj = np.clip((pts[:, 1] / cell).astype(int), 0, nz - 1)
            mask[jj, ii] = True
    dist = ndimage.distance_transform_edt(~mask) * cell
    hfield = np.clip(h_min + growth * dist, h_min, h_max)
    # graded point set: level k places a jittered grid of spacing s_k where s_k <= h < 2 s_k
    pts = []
    s = h_min
    while s <= h_max:
        gx = np.arange(s / 2, W, s)
        gz = np.arange(s / 2, H, s)
        GX, GZ = np.meshgrid(gx, gz)
        P = np.stack([GX.ravel(), GZ.ravel()], 1) + rng.uniform(-0.15, 0.15, (GX.size, 2)) * s
        hi = hfield[np.clip((P[:, 1] / cell).astype(int), 0, nz - 1), np.clip((P[:, 0] / cell).astype(int), 0, nx - 1)]
        top = 2 * s if 2 * s <= h_max else np.inf
        keep = (hi >= s) & (hi < top) & (hi > 1.5 * h_min if s == h_min else True)
        pts.append(P[keep])
        s *= 2
    # feature boundaries at spacing h_min, background points cleared around them
    bnd = []
    for (cx, cz, r) in vugs:
        n = max(8, int(np.ceil(2 * np.pi * r / h_min)))
        t = np.linspace(0, 2 * np.pi, n, endpoint=False)
        bnd.append(np.stack([cx + r * np.cos(t), cz + r * np.sin(t)], 1))
    for p0, p1 in fr:
        Lf = np.hypot(*(p1 - p0))
        n = max(2, int(np.ceil(Lf / h_min)))
        nrm = np.array([-(p1 - p0)[1], (p1 - p0)[0]]) / Lf
        line = p0[None] + np.linspace(0, 1, n + 1)[:, None] * (p1 - p0)[None]
        bnd += [line + nrm[None], line - nrm[None]]
    from scipy.spatial import cKDTree
    Bd = np.concatenate(bnd)
    # fracture walls inside a vug are not interfaces; where features touch or cross, keep one of any
    # two boundary points closer than 0.6 h_min
    tree_v = cKDTree(vugs[:, :2])
    dv, iv = tree_v.query(Bd, k=1)
    on_circle = np.abs(dv - vugs[iv, 2]) < 1e-6
    Bd = Bd[on_circle | (dv > vugs[iv, 2] + 0.5 * h_min)]
    drop = np.zeros(len(Bd), bool)
    for i, j in cKDTree(Bd).query_pairs(0.6 * h_min):
        if not drop[i]:
            drop[j] = True
    Bd = Bd[~drop]
    # where two levels meet, drop coarse points closer than 0.75 of the finer spacing to finer points
    kept = [pts[0]]
    s = h_min
    for lvl in pts[1:]:
        finer = np.concatenate(kept + [Bd])
        d, _ = cKDTree(finer).query(lvl, distance_upper_bound=0.75 * 2 * s)
        kept.append(lvl[~np.isfinite(d) | (d > 0.75 * 2 * s)])
        s *= 2
    P = np.concatenate(kept)
    d, _ = cKDTree(Bd).query(P, distance_upper_bound=0.6 * h_min)
    P = P[~np.isfinite(d) | (d > 0.55 * h_min)]
    edge = np.arange(0.0, W + 1e-9, h_max)
    edgez = np.arange(h_max, H, h_max)
    border = np.concatenate([np.stack([edge, 0 * edge], 1), np.stack([edge, 0 * edge + H], 1),
                             np.stack([0 * edgez, edgez], 1), np.stack([0 * edgez + W, edgez], 1)])
    allp = np.unique(np.round(np.concatenate([P, Bd, border]), 6), axis=0)
    allp = allp[(allp[:, 0] >= 0) & (allp[:, 0] <= W) & (allp[:, 1] >= 0) & (allp[:, 1] <= H)]

    def attribute(c):
        a = np.full(len(c), 3)
        tree = cKDTree(vugs[:, :2])
        dd, ii = tree.query(c, k=4)
        inside = np.zeros(len(c), bool)
        for j in range(4):
            inside |= dd[:, j] < vugs[ii[:, j], 2]
        for p0, p1 in fr:
            v = p1 - p0
            t = np.clip(((c - p0) @ v) / (v @ v), 0, 1)
            dist_ = np.hypot(*(c - (p0 + t[:, None] * v)).T)
            a = np.where(dist_ < 1.0, 4, a)
        return np.where(inside, 1, a)
    mesh = triangulate(allp, sponge_box=sp.box(), attribute=attribute)
    a = mesh.attribute
    vp = np.select([a == 1, a == 4], [1500.0, 3000.0], 5500.0)
    vs = np.select([a == 1, a == 4], [0.0, 1200.0], 3000.0)
    rho = np.select([a == 1, a == 4], [1000.0, 2300.0], 2650.0)
    model = Model.from_velocities(vp, vs, rho)
    xs = np.linspace(200.0, W - 200.0, nshots)
    src = [SourceSpec(float(x), H - 50.0, L.DGX_SRC_EXPLOSIVE, k, 30.0) for k, x in enumerate(xs)]
    rx = np.stack([np.linspace(0.0, W, 401), np.full(401, H - 10.0)], axis=1)
    return Workload(f"carbonate_4x3km_{mesh.tri // 1000}k_tri", mesh, model, sp, src, rx, nsteps, 0)
