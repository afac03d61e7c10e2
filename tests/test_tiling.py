"""Full-step kernel tiling (CPU): every index the kernel reads or writes stays in its arrays, and the tile
plus ring of every workgroup is exactly what its stress phase needs.

The full-step kernel gives each workgroup a run of consecutive elements (own) and the elements outside the
run that share a face with it (ring), at most one per thread; the new velocities of ring elements go to
the tile's private rows vrow0 + ring entry, and the stress phase finds every face neighbour's new
velocity through nb2.  These checks run on the host-side tiling the solver uploads."""
import ctypes as C

import numpy as np
import pytest

import dgx
from dgx import lib as L
from dgx import unstructured as U

BLOCK = 256


def tiling(mesh):
    lib = L.load()
    f = lib.dgx_full_step_tiling
    f.restype = C.c_int
    f.argtypes = [C.POINTER(L.ProblemC)] + [C.POINTER(C.c_long)] * 4 + [C.c_void_p] * 4
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    nb = np.ascontiguousarray(mesh.neighbor, np.int64)
    T = mesh.tri
    ones = np.ones(T)
    p = L.ProblemC(T, mesh.point_num, L.dp(pt), L.lp(tr), L.lp(nb), L.dp(ones), L.dp(ones), L.dp(ones),
                   L.dptr(), 1.0, L.DGX_FP64, 1, 1, 0, L.dptr(), 0, None, 0, None, L.DGX_LAYOUT_ELEMENT, 0,
                   L.DGX_STEP_FULL)
    nt, nr, v0, vr = C.c_long(), C.c_long(), C.c_long(), C.c_long()
    assert f(C.byref(p), C.byref(nt), C.byref(nr), C.byref(v0), C.byref(vr), None, None, None, None) == 0
    tiles = np.zeros((nt.value, 4), np.int32)
    ring = np.zeros(max(1, nr.value), np.int32)
    nb2 = np.zeros((T, 4), np.int32)
    nbr = np.zeros((T, 4), np.int32)
    assert f(C.byref(p), C.byref(nt), C.byref(nr), C.byref(v0), C.byref(vr), tiles.ctypes.data,
             ring.ctypes.data, nb2.ctypes.data, nbr.ctypes.data) == 0
    return tiles, ring[:nr.value], nb2, nbr, v0.value, vr.value


def check_tiling(mesh):
    tiles, ring, nb2, nbr, vrow0, vrows = tiling(mesh)
    T = mesh.tri
    # tiles cover 0..T-1 in order, each at most one element per thread
    assert tiles[0, 0] == 0 and np.all(tiles[1:, 0] == tiles[:-1, 0] + tiles[:-1, 1])
    assert tiles[-1, 0] + tiles[-1, 1] == T
    assert np.all(tiles[:, 1] >= 1) and np.all(tiles[:, 1] + tiles[:, 3] <= BLOCK)
    # ring entries: consecutive ranges, in bounds, rows inside the velocity arrays
    assert np.all(tiles[1:, 2] == tiles[:-1, 2] + tiles[:-1, 3]) and tiles[-1, 2] + tiles[-1, 3] == len(ring)
    assert np.all((ring >= 0) & (ring < T))
    assert vrow0 >= T and vrow0 + len(ring) <= vrows and vrow0 % 64 == 0 and vrows % 64 == 0
    for t in range(len(tiles)):
        first, own, rb, nr = tiles[t]
        members = set(range(first, first + own))
        rng = ring[rb:rb + nr]
        assert len(set(rng.tolist())) == nr and not (set(rng.tolist()) & members)
        outside = set()
        for q in range(first, first + own):
            for j in range(3):
                e = nbr[q, j]
                if e < 0:
                    assert nb2[q, j] == -1
                elif e in members:
                    assert nb2[q, j] == e
                else:
                    outside.add(e)
                    k = nb2[q, j] - vrow0
                    assert rb <= k < rb + nr and ring[k] == e
            assert nb2[q, 3] == nbr[q, 3]
        assert outside == set(rng.tolist())  # the ring is exactly the outside face neighbours
    return len(ring) / T


def test_tiling_structured_reference_mesh():
    w = dgx.configs.homogeneous_2km(10)
    overhead = check_tiling(w.mesh)
    assert overhead < 0.35


def test_tiling_jittered_random_diagonals():
    w = dgx.configs.gradient_8x4km(10, nx=120, nz=60)
    overhead = check_tiling(w.mesh)
    assert overhead < 0.35


def test_tiling_unstructured_delaunay():
    pts, _ = U.layered_point_set(1200.0, 600.0, 12.0, [lambda x: 300.0 + 0.2 * x], seed=4)
    mesh = U.triangulate(pts)
    check_tiling(mesh)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_tiling_tiny_meshes(n):
    mesh = dgx.mesh.structured(n, 1, 10.0)
    tiles, ring, nb2, nbr, vrow0, vrows = tiling(mesh)
    assert len(tiles) == 1 and tiles[0, 1] == mesh.tri and len(ring) == 0


# ------------------------------------------------------------------------------------------- blocked kernel
def block_tiling(mesh, K, cap):
    lib = L.load()
    f = lib.dgx_block_tiling
    f.restype = C.c_int
    f.argtypes = [C.POINTER(L.ProblemC), C.c_int, C.c_int] + [C.POINTER(C.c_long)] * 2 + [C.c_void_p] * 5
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    nb = np.ascontiguousarray(mesh.neighbor, np.int64)
    T = mesh.tri
    ones = np.ones(T)
    p = L.ProblemC(T, mesh.point_num, L.dp(pt), L.lp(tr), L.lp(nb), L.dp(ones), L.dp(ones), L.dp(ones),
                   L.dptr(), 1.0, L.DGX_FP32, 1, 1, 0, L.dptr(), 0, None, 0, None, L.DGX_LAYOUT_ELEMENT, 0,
                   L.DGX_STEP_BLOCKED, 0, K)
    nt, nl = C.c_long(), C.c_long()
    assert f(C.byref(p), K, cap, C.byref(nt), C.byref(nl), None, None, None, None, None) == 0
    tiles = np.zeros((nt.value, 4), np.int32)
    levels = np.zeros((nt.value, 8), np.int32)
    gid = np.zeros(nl.value, np.int32)
    lnb = np.zeros((nl.value, 4), np.int32)
    nbr = np.zeros((T, 4), np.int32)
    assert f(C.byref(p), K, cap, C.byref(nt), C.byref(nl), tiles.ctypes.data, levels.ctypes.data, gid.ctypes.data,
             lnb.ctypes.data, nbr.ctypes.data) == 0
    return tiles, levels, gid, lnb, nbr


def check_block_tiling(mesh, K, cap):
    """The blocked kernel's tiles: own runs partition the elements in order; level d holds exactly the elements
    d face hops from the run (breadth-first distance, d <= 2K, level 2K the ghosts); levels below 2K fit `cap`
    threads, one each, and the ghosts fit the cap/4 - 1 LDS slots past them (the last of the cap/4 is the
    kernel's zero slot); every face neighbour of an element below level 2K is an element of the tile at level
    d-1, d or d+1, named by its LDS slot (its record index below 2K, cap + its index among the ghosts at 2K);
    the code word is the element's.  Returns records/own - 1 (recomputed work)."""
    tiles, levels, gid, lnb, nbr = block_tiling(mesh, K, cap)
    T = mesh.tri
    assert tiles[0, 0] == 0 and np.all(tiles[1:, 0] == tiles[:-1, 0] + tiles[:-1, 1])
    assert tiles[-1, 0] + tiles[-1, 1] == T
    lanes, ghosts = levels[:, 2 * K - 1], tiles[:, 3] - levels[:, 2 * K - 1]
    assert np.all(lanes <= cap) and np.all(ghosts <= cap // 4 - 1) and np.all(tiles[:, 1] >= 1)
    assert np.all(tiles[1:, 2] == tiles[:-1, 2] + tiles[:-1, 3]) and tiles[-1, 2] + tiles[-1, 3] == len(gid)
    for t in range(len(tiles)):
        first, own, base, n = tiles[t]
        le = levels[t]
        assert le[0] == own and np.all(np.diff(le) >= 0) and le[2 * K] == n and np.all(le[2 * K:] == n)
        g = gid[base:base + n]
        assert np.array_equal(g[:own], np.arange(first, first + own))
        assert len(set(g.tolist())) == n
        # breadth-first distance from the run
        dist = {int(e): 0 for e in g[:own]}
        frontier = list(g[:own])
        for d in range(1, 2 * K + 1):
            nxt = []
            for e in frontier:
                for x in nbr[e, :3]:
                    if x >= 0 and int(x) not in dist:
                        dist[int(x)] = d
                        nxt.append(int(x))
            frontier = nxt
        lev = np.searchsorted(le[:2 * K + 1], np.arange(n), side="right")
        assert sorted(dist) == sorted(g.tolist())
        for l in range(n):
            e = int(g[l])
            assert dist[e] == lev[l], (t, l)
            assert lnb[base + l, 3] == nbr[e, 3]
            if lev[l] == 2 * K:
                continue
            nl = le[2 * K - 1]
            for j in range(3):
                x, s = nbr[e, j], lnb[base + l, j]
                if x < 0:
                    assert s == -1
                else:
                    r = s if s < nl else nl + (s - cap)  # the neighbour's record from its LDS slot
                    assert (0 <= s < nl or cap <= s < cap + n - nl) and g[r] == x and abs(lev[r] - lev[l]) <= 1
    return (len(gid) - T) / T


@pytest.mark.parametrize("K,cap", [(1, 1024), (2, 1024), (3, 1024), (2, 512)])
def test_block_tiling_jittered(K, cap):
    w = dgx.configs.gradient_8x4km(10, nx=90, nz=50)
    check_block_tiling(w.mesh, K, cap)


def test_block_tiling_unstructured_and_tiny():
    pts, _ = U.layered_point_set(900.0, 500.0, 12.0, [lambda x: 250.0 + 0.2 * x], seed=4)
    check_block_tiling(U.triangulate(pts), 2, 1024)
    for n in (1, 2, 5):
        tiles, levels, gid, lnb, nbr = block_tiling(dgx.mesh.structured(n, 1, 10.0), 3, 1024)
        assert len(tiles) == 1 and tiles[0, 1] == 2 * n


def test_block_tiling_refuses_what_does_not_fit():
    """An element whose 2K - 1 levels need more lanes than the workgroup has threads, or whose ghost level needs
    more than the cap/4 - 1 ghost slots, cannot be tiled: the builder reports DGX_EMESH (the solver then falls
    back to the half-step kernels, or refuses an explicit blocked request), and never a partial tiling."""
    mesh = dgx.mesh.structured(30, 30, 10.0, diagonal="random", seed=2)
    lib = L.load()
    f = lib.dgx_block_tiling
    f.restype = C.c_int
    f.argtypes = [C.POINTER(L.ProblemC), C.c_int, C.c_int] + [C.POINTER(C.c_long)] * 2 + [C.c_void_p] * 5
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    nb = np.ascontiguousarray(mesh.neighbor, np.int64)
    ones = np.ones(mesh.tri)
    p = L.ProblemC(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), L.lp(nb), L.dp(ones), L.dp(ones), L.dp(ones),
                   L.dptr(), 1.0, L.DGX_FP32, 1, 1, 0, L.dptr(), 0, None, 0, None, L.DGX_LAYOUT_ELEMENT, 0,
                   L.DGX_STEP_BLOCKED, 0, 3)
    for cap in (32, 64):  # 3 steps: 5 levels of lanes and a ghost level around one element exceed these
        nt, nl = C.c_long(-7), C.c_long(-7)
        assert f(C.byref(p), 3, cap, C.byref(nt), C.byref(nl), None, None, None, None, None) == L.DGX_EMESH
        assert nt.value == -7 and nl.value == -7
    check_block_tiling(mesh, 3, 128)


def test_block_tiling_recompute_fraction_640k_order():
    """Work recomputed on the rings, for a mesh of the 640k-element kind at a quarter of its size: 2 steps per
    launch at 1024 threads recompute well under the own work."""
    w = dgx.configs.gradient_8x4km(10, nx=400, nz=200)
    tiles, levels, gid, lnb, nbr = block_tiling(w.mesh, 2, 1024)
    frac = (len(gid) - w.mesh.tri) / w.mesh.tri
    print(f"2 steps, 1024 threads: lanes/own - 1 = {frac:.3f}")
    assert frac < 0.6
