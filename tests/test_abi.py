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
    # an inconsistent mesh is found by the mesh build, which runs on the GPU (tests/test_gpu_meshdev.py)


def _with_source(mesh, elem=1, shot=0, wavelet=0, kind=None, **over):
    src = (L.Source * 1)()
    src[0].elem, src[0].shot, src[0].wavelet = elem, shot, wavelet
    src[0].kind = L.DGX_SRC_EXPLOSIVE if kind is None else kind
    wav = np.zeros(10)
    fields = dict(nsrc=1, src=C.cast(src, C.POINTER(L.Source)), wavelet=L.dp(wav), nwavelets=1)
    fields.update(over)
    p, keep = _problem(mesh, **fields)
    keep.update(src=src, wav=wav)
    return p, keep


def _with_receiver(mesh, elem):
    rec = (L.Receiver * 1)()
    rec[0].elem = elem
    rec[0].weight[0] = 1.0
    p, keep = _problem(mesh, nrec=1, rec=C.cast(rec, C.POINTER(L.Receiver)))
    keep.update(rec=rec)
    return p, keep


def test_create_rejects_invalid_configurations():
    """Every rejection below happens on the host, with DGX_EARG and a message, before any device work:
    an empty mesh, a negative step count, an unknown precision, layout or step kernel, the full-step
    kernel with the shot-innermost layout, a shot tile that is not a power of two from 4 to 64 or that
    is given for the element layout, a workgroup size the kernels are not built for (or given for the
    element layout), the reference-arithmetic precision with another layout or kernel than its own, and sources
    or receivers that name a missing element, shot, wavelet or kind."""
    lib = L.load()
    mesh = dgx.structured(4, 4, 10.0)
    ops = L.read_operators()
    T = mesh.tri
    cases = {
        "empty mesh": _problem(mesh, tri=0),
        "negative steps": _problem(mesh, nsteps=-1),
        "precision": _problem(mesh, precision=7),
        "layout": _problem(mesh, layout=9),
        "step kernel": _problem(mesh, step_kernel=5),
        "full step, shot-innermost": _problem(mesh, nshots=8, layout=L.DGX_LAYOUT_SHOT_INNER,
                                              step_kernel=L.DGX_STEP_FULL),
        "shot tile 12": _problem(mesh, nshots=8, layout=L.DGX_LAYOUT_SHOT_INNER, shot_tile=12),
        "shot tile, element layout": _problem(mesh, layout=L.DGX_LAYOUT_ELEMENT, shot_tile=16),
        "wavefronts per workgroup 3": _problem(mesh, nshots=16, layout=L.DGX_LAYOUT_SHOT_INNER, shot_tile=16,
                                               wg_waves=3),
        "2 wavefronts with shot tile 8": _problem(mesh, nshots=8, layout=L.DGX_LAYOUT_SHOT_INNER, shot_tile=8,
                                                  wg_waves=2),
        "wavefronts per workgroup, element layout": _problem(mesh, layout=L.DGX_LAYOUT_ELEMENT, wg_waves=4),
        "reference arithmetic, shot-innermost": _problem(mesh, nshots=8, precision=L.DGX_FP64_REFERENCE,
                                                         layout=L.DGX_LAYOUT_SHOT_INNER),
        "reference arithmetic, blocked kernel": _problem(mesh, precision=L.DGX_FP64_REFERENCE,
                                                         step_kernel=L.DGX_STEP_BLOCKED),
        "reference arithmetic, block steps": _problem(mesh, precision=L.DGX_FP64_REFERENCE, block_steps=2),
        "source element 0": _with_source(mesh, elem=0),
        "source element past the mesh": _with_source(mesh, elem=T + 1),
        "source shot": _with_source(mesh, shot=1),
        "source wavelet": _with_source(mesh, wavelet=1),
        "source kind": _with_source(mesh, kind=9),
        "sources without wavelets": _with_source(mesh, nwavelets=0),
        "receiver element": _with_receiver(mesh, T + 1),
    }
    for name, (p, keep) in cases.items():
        h = C.c_void_p()
        assert lib.dgx_create(C.byref(p), C.byref(ops), C.byref(h)) == L.DGX_EARG, name
        assert not h.value, name
        assert lib.dgx_last_error(), name


def test_step_calls_reject_null_handle():
    lib = L.load()
    assert lib.dgx_step(None, 1) == -1
    assert lib.dgx_lu_velocity(None) == -1
    assert lib.dgx_get_fields(None, 0, None, None, None, None, None) == -1
