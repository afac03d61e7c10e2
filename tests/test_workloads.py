"""Workload generators (CPU): sizes, orientation, neighbour symmetry, interface conformity, material
regions and the sponge first-vertex rule the reference's absorbing-element list needs."""
import numpy as np
import pytest

import dgx
from dgx import unstructured as U


def check_mesh(mesh, sponge_box=None):
    assert np.all(mesh.area2() > 0)
    nb = mesh.neighbor
    T = mesh.tri
    # symmetric neighbour relation
    for c in range(3):
        e = np.nonzero(nb[:, c] > 0)[0]
        back = nb[nb[e, c] - 1]
        assert np.all((back == (e + 1)[:, None]).any(axis=1))
    if sponge_box is not None:
        x0, x1, z0, z1 = sponge_box
        p = mesh.point[mesh.triangle - 1]
        out = (p[..., 0] > x1) | (p[..., 0] < x0) | (p[..., 1] > z1) | (p[..., 1] < z0)
        assert np.all(out[:, 0] == out.any(axis=1))
    return T


def test_homogeneous_2km_is_the_reference_sized_case():
    w = dgx.configs.homogeneous_2km(10)
    assert check_mesh(w.mesh, w.sponge.box()) == 20000 and w.mesh.point_num == 101 * 101
    assert len(w.receivers) == 101


def test_gradient_8x4km_model():
    w = dgx.configs.gradient_8x4km(10, nx=80, nz=40)
    check_mesh(w.mesh, w.sponge.box())
    depth = 400.0 - w.mesh.centroids()[:, 1]
    vp = w.model.vp
    assert np.allclose(vp, 2000 + 2500 * depth / 4000 + np.where(depth >= 1500, 800, 0))
    assert np.allclose(w.model.den, 310 * vp ** 0.25)


def test_slope_model_layers_and_conformity():
    W, H = 6000.0, 3000.0
    t20 = np.tan(np.radians(20.0))
    fs = [lambda x: 2800 - t20 * x, lambda x: 2400 - 0.3 * x, lambda x: 1600 - 0.2 * x]
    pts, lines = U.layered_point_set(W, H, 40.0, fs, seed=2)
    m = U.triangulate(pts, sponge_box=(100, W - 100, 100, H - 100))
    check_mesh(m, (100, W - 100, 100, H - 100))
    for ln in lines:
        assert U.interface_conformity(m, ln) > 0.97
    w = dgx.configs.slope_6x3km(10, h=40.0, nshots=4, nrec=11)
    a = w.mesh.attribute
    assert set(np.unique(a)) == {1, 3, 4, 5}
    assert np.all(w.model.miu[a == 1] == 0)  # water is acoustic
    c = w.mesh.centroids()
    assert np.all(c[a == 1, 1] > fs[0](c[a == 1, 0]))


def test_serving_mesh_sizes_and_requests():
    ms = dgx.configs.serving_meshes((20_000, 50_000))
    for (m, md, sp), n in zip(ms, (20_000, 50_000)):
        assert abs(m.tri - n) / n < 0.01
        check_mesh(m, sp.box())
        assert 1800 <= md.vp.min() and md.vp.max() <= 5000
    arr, specs = dgx.configs.serving_requests(ms, n=500, rate=20.0)
    assert np.all(np.diff(arr) > 0)
    assert abs(500 / arr[-1] - 20.0) < 3.0
    assert all(1 <= len(r["xz"]) <= 16 and 1000 <= r["nsteps"] <= 6000 for r in specs)


@pytest.mark.slow
def test_carbonate_small():
    w = dgx.configs.carbonate_4x3km(10, nvugs=20, nfrac=5, growth=0.4, nshots=2)
    check_mesh(w.mesh, w.sponge.box())
    a = w.mesh.attribute
    assert {1, 3, 4} <= set(np.unique(a))
