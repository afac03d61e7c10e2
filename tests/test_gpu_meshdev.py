"""Mesh set-up and point location on the device against the host code, bit for bit.

The device build (element order, face codes, vertex renumbering, materials and sponge factors in the kernels'
layout) must equal dgx_build_mesh_host array by array: the reference builds the same data in serial host loops
(Neighbor_Edge fun_DG.c:557-739, NeighborNodes fun_DG.c:1195-1276, Elastic_A_B fun_DG.c:1071-1088), and the
host build is the one every parity test has run on.  The device locator must equal the host locator (the
reference's vertex rules of Get_Source / Get_Receiver, fun_DG.c:284-418, and the first containing element
elsewhere) on vertices, near-vertex points, edge points, interior points and points outside the mesh.  An
inconsistent mesh is refused by dgx_create (which builds on the device) with the host build's error.
"""
import ctypes as C

import numpy as np
import pytest

import dgx
from dgx import lib as L
from dgx.mesh import Mesh, compute_neighbors

pytestmark = pytest.mark.gpu


def workloads():
    yield "homogeneous_2km_40", dgx.configs.homogeneous_2km(10, n=40, h=50.0)
    yield "gradient_jittered_120x60", dgx.configs.gradient_8x4km(10, nx=120, nz=60, h=66.0)
    yield "slope_delaunay_h60", dgx.configs.slope_6x3km(10, h=60.0, nshots=2, nrec=41)


def shuffled(mesh, seed):
    """The same mesh with its elements listed in a random order and each element's vertex list rotated (still
    counter-clockwise)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(mesh.tri)
    tri = mesh.triangle[order]
    rot = rng.integers(0, 3, mesh.tri)
    tri = np.stack([np.roll(t, -r) for t, r in zip(tri, rot)])
    return Mesh(mesh.point.copy(), tri, compute_neighbors(tri), mesh.attribute[order])


def problem(mesh, model, absorb):
    keep = [np.ascontiguousarray(mesh.point, np.float64), np.ascontiguousarray(mesh.triangle, np.int64),
            np.ascontiguousarray(mesh.neighbor, np.int64), np.ascontiguousarray(model.lam, np.float64),
            np.ascontiguousarray(model.miu, np.float64), np.ascontiguousarray(model.den, np.float64)]
    ab = L.dptr()
    if absorb is not None:
        keep.append(np.ascontiguousarray(absorb, np.float64))
        ab = L.dp(keep[-1])
    p = L.ProblemC(mesh.tri, mesh.point_num, L.dp(keep[0]), L.lp(keep[1]), L.lp(keep[2]), L.dp(keep[3]),
                   L.dp(keep[4]), L.dp(keep[5]), ab, 1.0, L.DGX_FP64, 1, 1, 0, L.dptr(), 0, None, 0, None,
                   L.DGX_LAYOUT_ELEMENT, 0, 0)
    return p, keep


def compare(p, shot_inner, tri_pad, explosive):
    f = L.load().dgx_mesh_build_compare
    f.restype = C.c_int
    f.argtypes = [C.POINTER(L.ProblemC), C.c_int, C.c_long, C.c_char_p, C.POINTER(C.c_long)]
    bad = C.c_long(-1)
    rc = f(C.byref(p), shot_inner, tri_pad, explosive, C.byref(bad))
    return rc, bad.value


@pytest.mark.parametrize("which", range(3))
def test_device_mesh_build_equals_host(which):
    name, w = list(workloads())[which]
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, 10, sponge=w.sponge)
    for variant, mesh in (("as built", w.mesh), ("shuffled", shuffled(w.mesh, 11 + which))):
        T = mesh.tri
        absorb = st.absorb_fun if variant == "as built" else None
        model = w.model if variant == "as built" else dgx.Model.by_attribute(mesh.attribute)
        p, keep = problem(mesh, model, absorb)
        rng = np.random.default_rng(which)
        explosive = (rng.random(T) < 0.02).astype(np.uint8).tobytes()
        for shot_inner, tri_pad in ((0, (T + 2047) // 2048 * 2048), (1, (T + 15) // 16 * 16)):
            for ex in (None, explosive):
                print(f"case device mesh build {name} {variant}: {T} elements, shot_inner {shot_inner}, "
                      f"explosive {ex is not None}", flush=True)
                rc, bad = compare(p, shot_inner, tri_pad, ex)
                assert rc == 0, L.load().dgx_last_error()
                assert bad == 0, L.load().dgx_last_error()


def host_locate(mesh, xz, ops=None, receivers=True):
    lib = L.load()
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    loc = C.c_void_p()
    L.check(lib.dgx_locator_create(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), C.byref(loc)), "locator")
    try:
        if receivers:
            out = (L.Receiver * len(xz))()
            rc = lib.dgx_locate_receivers(loc, len(xz), L.dp(xz), out)
            return rc, out
        out = (L.Source * len(xz))()
        for k, (x, z) in enumerate(xz):
            rc = lib.dgx_locate_source(loc, C.byref(ops), x, z, L.DGX_SRC_EXPLOSIVE, 0, k, C.byref(out[k]))
            if rc:
                return rc, out
        return 0, out
    finally:
        lib.dgx_locator_destroy(loc)


def device_locate(mesh, xz, ops=None, receivers=True):
    lib = L.load()
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    n = len(xz)
    if receivers:
        out = (L.Receiver * max(1, n))()
        rc = lib.dgx_locate_receivers_device(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), n, L.dp(xz), out)
        return rc, out
    out = (L.Source * max(1, n))()
    kind = np.full(n, L.DGX_SRC_EXPLOSIVE, np.int32)
    shot = np.zeros(n, np.int32)
    wav = np.arange(n, dtype=np.int64)
    rc = lib.dgx_locate_sources_device(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), C.byref(ops), n, L.dp(xz),
                                       kind.ctypes.data_as(C.POINTER(C.c_int)),
                                       shot.ctypes.data_as(C.POINTER(C.c_int)), L.lp(wav), out)
    return rc, out


def queries(mesh, rng, n=400):
    """Vertices, points within 0.8e-3 and 6e-3 of vertices (inside the source and receiver vertex tolerances
    1e-3 and 1e-2, or only the receivers'), edge midpoints, random interior points."""
    v = mesh.point[rng.choice(mesh.point_num, n // 4, replace=False)]
    near = v + rng.uniform(-0.8e-3, 0.8e-3, v.shape)
    near2 = v + rng.uniform(-6e-3, 6e-3, v.shape)
    t = rng.choice(mesh.tri, n // 4, replace=False)
    a, b = mesh.point[mesh.triangle[t, 0] - 1], mesh.point[mesh.triangle[t, 1] - 1]
    c = mesh.point[mesh.triangle[t, 2] - 1]
    mid = (a + b) / 2
    u = rng.random((len(t), 2))
    u = np.where(u.sum(1, keepdims=True) > 1, 1 - u, u)
    inner = a + u[:, :1] * (b - a) + u[:, 1:] * (c - a)
    lo, hi = mesh.point.min(0), mesh.point.max(0)  # the workloads' domains are rectangles
    near, near2 = np.clip(near, lo, hi), np.clip(near2, lo, hi)
    return np.ascontiguousarray(np.concatenate([v, near, near2, mid, inner]), np.float64)


@pytest.mark.parametrize("which", range(3))
def test_device_locator_equals_host(which):
    name, w = list(workloads())[which]
    mesh = w.mesh
    rng = np.random.default_rng(100 + which)
    xz = queries(mesh, rng)
    ops = L.read_operators()
    print(f"case device locator {name}: {mesh.tri} elements, {len(xz)} points", flush=True)
    for receivers in (True, False):
        rh, oh = host_locate(mesh, xz, ops, receivers)
        rd, od = device_locate(mesh, xz, ops, receivers)
        assert rh == 0 and rd == 0, L.load().dgx_last_error()
        for k in range(len(xz)):
            assert oh[k].elem == od[k].elem, (receivers, k, xz[k])
            if receivers:
                assert bytes(oh[k].weight) == bytes(od[k].weight), k
            else:
                assert bytes(oh[k].fun) == bytes(od[k].fun), k
    # a point outside the mesh is refused by both, with the same message
    W, H = mesh.point[:, 0].max(), mesh.point[:, 1].max()
    out_xz = np.ascontiguousarray([[W / 2, H / 2], [W + 50.0, H / 2]], np.float64)
    rh, _ = host_locate(mesh, out_xz)
    mh = L.load().dgx_last_error()
    rd, _ = device_locate(mesh, out_xz)
    md = L.load().dgx_last_error()
    assert rh == rd == L.DGX_EARG and mh == md and b"outside the mesh" in md


def test_create_refuses_inconsistent_meshes_like_the_host_build():
    """Clockwise element, a neighbour that does not share the face, a vertex id out of range, a density <= 0:
    dgx_create (device build) returns the host build's code and message."""
    lib = L.load()
    mesh = dgx.structured(6, 6, 10.0, diagonal="random", seed=2)
    ops = L.read_operators()
    cases = []
    bad = mesh.triangle.copy()
    bad[7] = bad[7, ::-1]  # element 8 clockwise
    cases.append(("clockwise", dgx.Mesh(mesh.point, bad, compute_neighbors(bad), mesh.attribute), None))
    wrong = mesh.neighbor.copy()
    k = int(np.nonzero(wrong[:, 0] > 0)[0][3])
    wrong[k, 0] = 1 if wrong[k, 0] != 1 else 2  # a neighbour that does not share the face
    cases.append(("neighbour", dgx.Mesh(mesh.point, mesh.triangle, wrong, mesh.attribute), None))
    out = mesh.triangle.copy()
    out[5, 1] = mesh.point_num + 3
    cases.append(("vertex id", dgx.Mesh(mesh.point, out, mesh.neighbor, mesh.attribute), None))
    den = np.ones(mesh.tri)
    den[9] = 0.0
    cases.append(("density", mesh, den))
    for what, m, d in cases:
        print(f"case inconsistent mesh: {what}", flush=True)
        model = dgx.Model(np.ones(m.tri), np.ones(m.tri), np.ones(m.tri) if d is None else d, np.ones(m.tri))
        p, keep = problem(m, model, None)
        rc_host, _ = compare(p, 0, (m.tri + 2047) // 2048 * 2048, None)  # the host build runs first and fails
        msg_host = lib.dgx_last_error()
        h = C.c_void_p()
        rc_dev = lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h))
        msg_dev = lib.dgx_last_error()
        assert rc_host < 0 and rc_dev == rc_host and msg_dev == msg_host, (what, rc_host, rc_dev, msg_host, msg_dev)
