"""libdgx host setup (include/dgx_setup.h) against the oracle and the reference's formats (CPU only)."""
import ctypes as C

import numpy as np
import pytest

import dgx
import oracle_py as O
from dgx import formats as F
from dgx import lib as L
from helpers import OPDIR, OracleCase, product_setup


def test_reader_matches_oracle_reader_bitwise():
    ops = L.read_operators()
    orc = O.read_operators(OPDIR)
    for name in ("V", "inverV", "Vr", "Vs", "S1", "S2", "S3", "FeketeNew"):
        assert np.array_equal(ops.array(name), orc[name]), name


def test_mesh_files_roundtrip(tmp_path):
    w = dgx.configs.gradient_8x4km(10, nx=40, nz=20)
    ele, node, neigh = dgx.formats.write_mesh(tmp_path / "m.1", w.mesh)
    lib = L.load()
    T, P = lib.dgx_mesh_count(ele.encode()), lib.dgx_mesh_count(node.encode())
    assert (T, P) == (w.mesh.tri, w.mesh.point_num)
    pt, tr, nb, at = np.zeros(P * 2), np.zeros(T * 3, np.int64), np.zeros(T * 3, np.int64), np.zeros(T, np.int64)
    L.check(lib.dgx_read_mesh(ele.encode(), node.encode(), neigh.encode(), T, P, L.dp(pt), L.lp(tr), L.lp(nb),
                              L.lp(at)), "dgx_read_mesh")
    assert np.array_equal(tr.reshape(-1, 3), w.mesh.triangle)
    assert np.array_equal(nb.reshape(-1, 3), w.mesh.neighbor)
    assert np.array_equal(at, w.mesh.attribute)
    assert np.allclose(pt.reshape(-1, 2), w.mesh.point, rtol=6e-8, atol=0)  # 8 significant digits
    # the oracle's restatement of ReadDouble gives the same doubles
    orc = O.parse_double_text(open(node, "rb").read(), P + 1, 4)[1:, 1:3]
    assert np.array_equal(orc, pt.reshape(-1, 2))


def test_save_data_format(tmp_path):
    m = np.array([[1.5, -2.25e-9], [3.0e4, 0.0]])
    p = tmp_path / "Vx100"
    L.check(L.load().dgx_save_data(str(p).encode(), L.dp(m), 2, 2))
    assert p.read_text() == "1.50000000000000 -0.00000000225000 \n30000.00000000000000 0.00000000000000 \n"


@pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")
def test_save_data_matches_reference_writer(tmp_path):
    m = np.random.default_rng(0).standard_normal((7, 2)) * 1e3
    a, b = tmp_path / "a", tmp_path / "b"
    L.check(L.load().dgx_save_data(str(a).encode(), L.dp(m), 7, 2))
    O.reference().Save_Data(str(b).encode(), O.dp(m), C.c_long(7), C.c_long(2))
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("maker", [lambda: dgx.configs.homogeneous_2km(50),
                                   lambda: dgx.configs.gradient_8x4km(50, nx=120, nz=60, source_x=600.0, source_depth=300.0)])
def test_setup_matches_oracle(maker):
    w = maker()
    oc = OracleCase(w)
    st = product_setup(w)
    assert st.dt == oc.dt
    assert st.t0 == oc.t0 and np.array_equal(st.wavelets, oc.wav)
    assert np.array_equal(st.absorb_fun, oc.lay.absorb_fun)
    for s, (e, kind, fun, shot, wi) in zip(st.sources, oc.src):
        assert s.elem == e and s.kind == kind
        assert np.allclose(np.array(list(s.fun)), fun, rtol=1e-14, atol=0)
    for r, (e, wgt) in zip(st.receivers, oc.rec):
        assert r.elem == e
        assert np.allclose(np.array(list(r.weight)), wgt, rtol=0, atol=1e-14)


def test_sponge_rejects_mesh_without_first_vertex_rule():
    mesh = dgx.structured(20, 20, 100.0, diagonal="same")  # no rotation: some elements violate the rule
    lib = L.load()
    ab = np.zeros(mesh.tri * 6)
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0).params()
    rc = lib.dgx_sponge(mesh.tri, L.dp(np.ascontiguousarray(mesh.point)), L.lp(np.ascontiguousarray(mesh.triangle)),
                        L.dp(np.ascontiguousarray(L.read_operators().array("FeketeNew"))), C.byref(sp), L.dp(ab))
    assert rc == -2 and b"first vertex" in lib.dgx_last_error()


def test_vertex_source_and_receiver_follow_reference_rules():
    """Source: first element holding the vertex; receiver: last element holding it (fun_DG.c:343-350,390-414)."""
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(20, 20, 100.0, diagonal="random", seed=5, sponge_box=sp.box())
    model = dgx.Model.by_attribute(np.full(mesh.tri, 3))
    st = dgx.build_setup(mesh, model, [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z)], [[1600.0, 1600.0]], 10,
                         sponge=sp)
    holders = np.nonzero((mesh.triangle == 1 + np.argmin(np.hypot(*(mesh.point - [1200, 1200]).T))).any(1))[0]
    assert st.sources[0].elem == holders.min() + 1
    holders = np.nonzero((mesh.triangle == 1 + np.argmin(np.hypot(*(mesh.point - [1600, 1600]).T))).any(1))[0]
    assert st.receivers[0].elem == holders.max() + 1


def test_binary_mesh_round_trips_and_matches_the_text_reader(tmp_path):
    """dgx_write_mesh_binary / dgx_read_mesh_binary against the numpy writer/reader of the same format and
    against the text reader (Mesh_Input) on the same mesh."""
    lib = L.load()
    mesh = dgx.mesh.structured(7, 5, 30.0, diagonal="random", seed=5, attribute=lambda c: np.where(c[:, 1] > 60, 1, 3))
    T, P = mesh.tri, mesh.point_num
    pt = np.ascontiguousarray(mesh.point, np.float64)
    tr = np.ascontiguousarray(mesh.triangle, np.int64)
    nb = np.ascontiguousarray(mesh.neighbor, np.int64)
    at = np.ascontiguousarray(mesh.attribute, np.int64)
    c_path = str(tmp_path / "c.dgxm")
    assert lib.dgx_write_mesh_binary(c_path.encode(), T, P, L.dp(pt), L.lp(tr), L.lp(nb), L.lp(at)) == 0
    m = F.read_mesh_binary(c_path)
    assert np.array_equal(m.point, pt) and np.array_equal(m.triangle, tr)
    assert np.array_equal(m.neighbor, nb) and np.array_equal(m.attribute, at)

    py_path = str(tmp_path / "py.dgxm")
    F.write_mesh_binary(py_path, mesh)
    assert open(py_path, "rb").read() == open(c_path, "rb").read()
    t, p = C.c_long(), C.c_long()
    assert lib.dgx_mesh_binary_counts(py_path.encode(), C.byref(t), C.byref(p)) == 0 and (t.value, p.value) == (T, P)
    pt2, tr2, nb2, at2 = np.zeros_like(pt), np.zeros_like(tr), np.zeros_like(nb), np.zeros_like(at)
    assert lib.dgx_read_mesh_binary(py_path.encode(), T, P, L.dp(pt2), L.lp(tr2), L.lp(nb2), L.lp(at2)) == 0
    assert np.array_equal(pt2, pt) and np.array_equal(tr2, tr) and np.array_equal(nb2, nb) and np.array_equal(at2, at)

    # the text files of the same mesh read to the same arrays, except for the %16.7e rounding of coordinates
    ele, node, neigh = F.write_mesh(tmp_path / "m", mesh)
    pt3, tr3, nb3, at3 = np.zeros_like(pt), np.zeros_like(tr), np.zeros_like(nb), np.zeros_like(at)
    assert lib.dgx_read_mesh(ele.encode(), node.encode(), neigh.encode(), T, P, L.dp(pt3), L.lp(tr3), L.lp(nb3),
                             L.lp(at3)) == 0
    assert np.array_equal(tr3, tr) and np.array_equal(nb3, nb) and np.array_equal(at3, at)
    assert np.allclose(pt3, pt, rtol=1e-7, atol=0)

    # wrong counts, a foreign file and a truncated file are errors, not reads past the end
    assert lib.dgx_read_mesh_binary(py_path.encode(), T + 1, P, L.dp(pt2), L.lp(tr2), L.lp(nb2), L.lp(at2)) == L.DGX_EARG
    assert lib.dgx_mesh_binary_counts(ele.encode(), C.byref(t), C.byref(p)) == L.DGX_EIO
    trunc = str(tmp_path / "t.dgxm")
    open(trunc, "wb").write(open(py_path, "rb").read()[:-9])
    assert lib.dgx_read_mesh_binary(trunc.encode(), T, P, L.dp(pt2), L.lp(tr2), L.lp(nb2), L.lp(at2)) == L.DGX_EIO


@pytest.mark.parametrize("maker", [lambda: dgx.configs.homogeneous_2km(10),
                                   lambda: dgx.configs.gradient_8x4km(10, nx=120, nz=60),
                                   lambda: dgx.configs.slope_6x3km(10, h=40.0, nshots=1, nrec=5)])
def test_reference_operators_equal_oracle_bitwise(maker):
    """dgx_reference_operators forms Dx, Dz, Q, A, B with the reference's expressions in its order: the arrays
    equal the oracle's (itself equal to the reference build's on the fixtures) bit for bit."""
    w = maker()
    m, T = w.mesh, w.mesh.tri
    ops = O.read_operators(OPDIR)
    lay = O.setup_reference_layout(m.point, m.triangle, m.neighbor, w.model.lam, w.model.miu, w.model.den, ops, None)
    Dx, Dz, Q, A, B = np.zeros(T * 36), np.zeros(T * 36), np.zeros(T * 30), np.zeros(T * 5), np.zeros(T * 5)
    c = lambda a, t=np.float64: np.ascontiguousarray(a, t)
    L.check(L.load().dgx_reference_operators(T, L.dp(c(m.point)), L.lp(c(m.triangle, np.int64)), L.dp(c(w.model.lam)),
                                             L.dp(c(w.model.miu)), L.dp(c(w.model.den)), C.byref(L.read_operators()),
                                             L.dp(Dx), L.dp(Dz), L.dp(Q), L.dp(A), L.dp(B)), "dgx_reference_operators")
    for name, mine, orc in (("Dx", Dx, lay.Dx), ("Dz", Dz, lay.Dz), ("Q", Q, lay.Q), ("A", A, lay.A), ("B", B, lay.B)):
        assert np.array_equal(mine, orc), (name, np.abs(mine - orc).max())
