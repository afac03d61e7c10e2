"""The C-ABI library loads and exports every function include/*.h declares; host-side error paths."""
import ctypes as C
import pathlib
import re

import numpy as np

import dgx
from dgx import lib as L

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(dgx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = L.load()
    decl = declared_functions()
    assert len(decl) > 30
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing


def _problem(mesh, **over):
    T = mesh.tri
    keep = dict(pt=np.ascontiguousarray(mesh.point), tr=np.ascontiguousarray(mesh.triangle),
                nb=np.ascontiguousarray(mesh.neighbor), m=np.ones(T))
    p = L.ProblemC(T, mesh.point_num, L.dp(keep["pt"]), L.lp(keep["tr"]), L.lp(keep["nb"]), L.dp(keep["m"]),
                   L.dp(keep["m"]), L.dp(keep["m"]), L.dptr(), 1e-3, L.DGX_FP64, 1, 10, 0, L.dptr(), 0,
                   C.POINTER(L.Source)(), 0, C.POINTER(L.Receiver)())
    for k, v in over.items():
        setattr(p, k, v)
    return p, keep


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    lib = L.load()
    mesh = dgx.structured(4, 4, 10.0)
    ops = L.read_operators()
    h = C.c_void_p()
    p, keep = _problem(mesh, nshots=0)
    assert lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h)) == -1
    p, keep = _problem(mesh, dt=0.0)
    assert lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h)) == -1
    bad = mesh.triangle[:, ::-1].copy()  # clockwise
    m2 = dgx.Mesh(mesh.point, bad, dgx.mesh.compute_neighbors(bad), mesh.attribute)
    p, keep = _problem(m2)
    assert lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h)) == -2
    assert b"counter-clockwise" in lib.dgx_last_error()
    wrong = mesh.neighbor.copy()
    wrong[0, 0] = 5  # a neighbour that does not share the face
    p, keep = _problem(dgx.Mesh(mesh.point, mesh.triangle, wrong, mesh.attribute))
    assert lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h)) == -2


def test_step_calls_reject_null_handle():
    lib = L.load()
    assert lib.dgx_step(None, 1) == -1
    assert lib.dgx_lu_velocity(None) == -1
    assert lib.dgx_get_fields(None, 0, None, None, None, None, None) == -1
