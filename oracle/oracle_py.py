"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/build/liboracle.so) and of the
reference's own functions (oracle/_ref/libref.so, built from the reference sources by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SO = HERE / "build" / "liboracle.so"
REF_SO = HERE / "_ref" / "libref.so"

dptr = C.POINTER(C.c_double)
lptr = C.POINTER(C.c_long)


def dp(a):
    return a.ctypes.data_as(dptr)


def lp(a):
    return a.ctypes.data_as(lptr)


class SpongeParams(C.Structure):
    _fields_ = [("bound_x", C.c_double), ("bound_z", C.c_double), ("model_w", C.c_double),
                ("model_h", C.c_double), ("beta", C.c_double), ("pi", C.c_double)]


class OrcSource(C.Structure):
    _fields_ = [("elem", C.c_long), ("kind", C.c_int), ("fun", C.c_double * 6)]


class OrcReceiver(C.Structure):
    _fields_ = [("elem", C.c_long), ("weight", C.c_double * 6)]


class OrcProblem(C.Structure):
    _fields_ = [("tri", C.c_long), ("dt", C.c_double), ("A", dptr), ("B", dptr), ("Dx", dptr), ("Dz", dptr),
                ("Q", dptr), ("S", dptr), ("absorb_fun", dptr), ("ibe", lptr), ("node_neighbor", lptr),
                ("neighbor", lptr), ("edge_flag", lptr), ("nsrc", C.c_long), ("src", C.POINTER(OrcSource)),
                ("wavelet", dptr), ("nrec", C.c_long), ("rec", C.POINTER(OrcReceiver))]


_orc = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def oracle() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            build()
        _orc = C.CDLL(str(ORACLE_SO))
        _orc.orc_parse_double.restype = C.c_long
        _orc.orc_parse_long.restype = C.c_long
        for n in ("orc_indiameter_min", "orc_time_step", "orc_ricker", "orc_ricker_delay"):
            getattr(_orc, n).restype = C.c_double
        _orc.orc_time_step.argtypes = [C.c_long, dptr, lptr, dptr, C.c_int]
        _orc.orc_ricker.argtypes = [C.c_long, C.c_double, C.c_double, C.c_double, C.c_double]
        _orc.orc_ricker_delay.argtypes = [C.c_double, C.c_double, C.c_double]
        _orc.orc_absorb_count.restype = C.c_long
        _orc.orc_absorbing_fun.restype = C.c_long
        _orc.orc_locate.restype = C.c_long
        _orc.orc_source_vector.argtypes = [dptr, C.c_long, C.c_double, dptr]
        _orc.orc_source_vector_at.argtypes = [dptr, dptr, C.c_double, dptr]
        _orc.orc_p2_basis.argtypes = [C.c_double, C.c_double, dptr]
        _orc.orc_locate.argtypes = [C.c_long, dptr, lptr, C.c_double, C.c_double, dptr, dptr]
        _orc.orc_get_source.argtypes = [C.c_long, C.c_long, dptr, lptr, C.c_double, C.c_double, lptr, lptr]
        _orc.orc_get_receiver.argtypes = [C.c_long, C.c_long, dptr, lptr, C.c_double, C.c_double, lptr, lptr]
        _orc.orc_lu_velocity.argtypes = [C.c_long, C.c_double] + [dptr] * 11 + [lptr] * 4
        _orc.orc_lu_stress.argtypes = [C.c_long, C.c_double] + [dptr] * 11 + [lptr] * 4
        _orc.orc_source_kernel.argtypes = [dptr, C.c_long, dptr, C.c_double]
        _orc.orc_num_threads.restype = C.c_int
    return _orc


def _ref_path(variant: str = "") -> pathlib.Path:
    """oracle/_ref/libref.so (-O2 -fopenmp), or libref_O0.so for variant "O0" (the README's flag-less build)."""
    return REF_SO if not variant else REF_SO.with_name(f"libref_{variant}.so")


def reference_available(variant: str = "") -> bool:
    return _ref_path(variant).exists()


_refs = {}


def reference(variant: str = "") -> C.CDLL:
    if variant not in _refs:
        path = _ref_path(variant)
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (needs the reference checkout; see oracle/Makefile)")
        r = C.CDLL(str(path))
        r.ref_pi.restype = C.c_double
        r.ref_macro.restype = C.c_long
        r.ref_case.restype = C.c_int
        r.ref_time_steps.restype = C.c_double
        for n in ("Calculate_Time_Step", "InDiameter_Min", "Ricker_Source", "Ricker_Delay_Time"):
            getattr(r, n).restype = C.c_double
        r.Ricker_Source.argtypes = [C.c_long, C.c_double, C.c_double]
        r.Ricker_Delay_Time.argtypes = [C.c_double]
        _refs[variant] = r
    return _refs[variant]


# ------------------------------------------------------------------------------------------------
# Operator files, parsed by the oracle's own restatement of the reference reader.
# ------------------------------------------------------------------------------------------------
OP_SHAPES = {"V": (6, 6), "inverV": (6, 6), "Vr": (6, 6), "Vs": (6, 6), "S1": (6, 3), "S2": (6, 3), "S3": (6, 3),
             "FeketeNew": (6, 2), "Vb": (3, 3), "inverVb": (3, 3)}


def parse_double_text(buf: bytes, rows: int, cols: int) -> np.ndarray:
    out = np.zeros(rows * cols, np.float64)
    n = oracle().orc_parse_double(C.c_char_p(buf), C.c_long(len(buf)), dp(out), C.c_long(rows), C.c_long(cols))
    assert n == rows * cols, (n, rows, cols)
    return out.reshape(rows, cols)


def parse_long_text(buf: bytes, rows: int, cols: int) -> np.ndarray:
    out = np.zeros(rows * cols, np.int64)
    n = oracle().orc_parse_long(C.c_char_p(buf), C.c_long(len(buf)), lp(out), C.c_long(rows), C.c_long(cols))
    assert n == rows * cols, (n, rows, cols)
    return out.reshape(rows, cols)


def read_operators(directory) -> dict[str, np.ndarray]:
    d = pathlib.Path(directory)
    return {k: parse_double_text((d / f"{k}.txt").read_bytes(), *s) for k, s in OP_SHAPES.items()}


# ------------------------------------------------------------------------------------------------
# The reference's setup, restated (fun_DG.c; DG_leap_Frog_OpenMP.c:190-343 order).
# ------------------------------------------------------------------------------------------------
@dataclasses.dataclass
class RefLayout:
    tri: int
    Dx: np.ndarray
    Dz: np.ndarray
    Q: np.ndarray
    A: np.ndarray
    B: np.ndarray
    S: np.ndarray
    ibe: np.ndarray
    neighbor: np.ndarray
    edge: np.ndarray
    edge_flag: np.ndarray
    node_neighbor: np.ndarray
    jacobi: np.ndarray
    jacobi_edge: np.ndarray
    normals: np.ndarray  # [3, Tri, 2]
    absorb_fun: np.ndarray | None


def setup_reference_layout(point, triangle, neigh_raw, lam, miu, den, ops, sponge: SpongeParams | None) -> RefLayout:
    o = oracle()
    point = np.ascontiguousarray(point, np.float64)
    triangle = np.ascontiguousarray(triangle, np.int64)
    neigh_raw = np.ascontiguousarray(neigh_raw, np.int64)
    T = triangle.shape[0]
    jac, je = np.zeros(T), np.zeros(T * 3)
    o.orc_jacobian(C.c_long(T), dp(point), lp(triangle), dp(je), dp(jac))
    n1, n2, n3 = np.zeros(T * 2), np.zeros(T * 2), np.zeros(T * 2)
    o.orc_normals(C.c_long(T), dp(point), lp(triangle), dp(n1), dp(n2), dp(n3))
    nb, edge, flag = np.zeros(T * 3, np.int64), np.zeros(T * 3, np.int64), np.zeros(T * 3, np.int64)
    o.orc_neighbor_edge(C.c_long(T), lp(neigh_raw), lp(triangle), lp(nb), lp(edge), lp(flag))
    ibe = np.zeros(9, np.int64)
    o.orc_ibe(lp(ibe))
    nn = np.zeros(T * 9, np.int64)
    o.orc_neighbor_nodes(C.c_long(T), lp(triangle), lp(edge), lp(nb), lp(ibe), lp(nn))
    A, B = np.zeros(T * 5), np.zeros(T * 5)
    o.orc_elastic_ab(C.c_long(T), dp(np.ascontiguousarray(lam, np.float64)), dp(np.ascontiguousarray(miu, np.float64)),
                     dp(np.ascontiguousarray(den, np.float64)), dp(A), dp(B))
    Q = np.zeros(T * 30)
    o.orc_central_flux(C.c_long(T), dp(jac), dp(je), dp(A), dp(B), dp(n1), dp(n2), dp(n3), dp(Q))
    Dx, Dz = np.zeros(T * 36), np.zeros(T * 36)
    o.orc_dx_dz(C.c_long(T), dp(point), lp(triangle), dp(np.ascontiguousarray(ops["Vr"])),
                dp(np.ascontiguousarray(ops["Vs"])), dp(np.ascontiguousarray(ops["inverV"])), dp(Dx), dp(Dz))
    S = np.concatenate([ops["S1"].ravel(), ops["S2"].ravel(), ops["S3"].ravel()]).copy()
    absorb = None
    if sponge is not None:
        bad = o.orc_absorb_count(C.c_long(T), dp(point), lp(triangle), C.byref(sponge))
        assert bad == 0, "mesh violates the first-vertex sponge rule"
        absorb = np.zeros(T * 6)
        o.orc_absorbing_fun(C.c_long(T), dp(point), lp(triangle), dp(np.ascontiguousarray(ops["FeketeNew"])),
                            C.byref(sponge), dp(absorb))
    return RefLayout(T, Dx, Dz, Q, A, B, S, ibe, nb, edge, flag, nn, jac, je, np.stack([n1, n2, n3]).reshape(3, T, 2),
                     absorb)


def time_step(point, triangle, vp) -> float:
    T = len(triangle)
    return oracle().orc_time_step(C.c_long(T), dp(np.ascontiguousarray(point, np.float64)),
                                  lp(np.ascontiguousarray(triangle, np.int64)), dp(np.ascontiguousarray(vp, np.float64)),
                                  2)


def ricker_delay(dt, f0, pi=3.1415) -> float:
    return oracle().orc_ricker_delay(dt, f0, pi)


def wavelet(nsteps, dt, t0, f0, pi=3.1415) -> np.ndarray:
    o = oracle()
    return np.array([o.orc_ricker(n, dt, t0, f0, pi) for n in range(nsteps)])


def locate_source(point, triangle, V, x, z) -> tuple[int, np.ndarray]:
    """Reference rule on a vertex (first vertex within 1e-3, first element); else first containing element."""
    o = oracle()
    point = np.ascontiguousarray(point, np.float64)
    triangle = np.ascontiguousarray(triangle, np.int64)
    T, P = len(triangle), len(point)
    e, node = C.c_long(), C.c_long()
    jac, je = np.zeros(T), np.zeros(T * 3)
    o.orc_jacobian(C.c_long(T), dp(point), lp(triangle), dp(je), dp(jac))
    V = np.ascontiguousarray(V, np.float64)
    fun = np.zeros(6)
    if o.orc_get_source(C.c_long(T), C.c_long(P), dp(point), lp(triangle), x, z, C.byref(e), C.byref(node)) == 0:
        o.orc_source_vector(dp(V), node, jac[e.value - 1], dp(fun))
        return e.value, fun
    r, s = C.c_double(), C.c_double()
    el = o.orc_locate(C.c_long(T), dp(point), lp(triangle), x, z, C.byref(r), C.byref(s))
    assert el > 0, "source outside the mesh"
    phi = np.zeros(6)
    o.orc_p2_basis(r.value, s.value, dp(phi))
    o.orc_source_vector_at(dp(V), dp(phi), jac[el - 1], dp(fun))
    return int(el), fun


def locate_receiver(point, triangle, x, z) -> tuple[int, np.ndarray]:
    o = oracle()
    point = np.ascontiguousarray(point, np.float64)
    triangle = np.ascontiguousarray(triangle, np.int64)
    T, P = len(triangle), len(point)
    e, node = C.c_long(), C.c_long()
    w = np.zeros(6)
    if o.orc_get_receiver(C.c_long(T), C.c_long(P), dp(point), lp(triangle), x, z, C.byref(e), C.byref(node)) == 0:
        w[node.value - 1] = 1.0
        return e.value, w
    r, s = C.c_double(), C.c_double()
    el = o.orc_locate(C.c_long(T), dp(point), lp(triangle), x, z, C.byref(r), C.byref(s))
    assert el > 0, "receiver outside the mesh"
    o.orc_p2_basis(r.value, s.value, dp(w))
    return int(el), w


class Run:
    """Advance the oracle state step by step (DG_leap_Frog_OpenMP.c:405-425 order)."""

    def __init__(self, lay: RefLayout, dt: float, sources, wavelet_rows, receivers):
        self.lay, self.dt = lay, dt
        T = lay.tri
        self.f = np.zeros((5, T * 6))
        self._src = (OrcSource * max(1, len(sources)))()
        for k, (elem, kind, fun) in enumerate(sources):
            self._src[k].elem, self._src[k].kind = elem, kind
            for i in range(6):
                self._src[k].fun[i] = fun[i]
        self._wav = np.ascontiguousarray(wavelet_rows, np.float64)
        self._rec = (OrcReceiver * max(1, len(receivers)))()
        for k, (elem, w) in enumerate(receivers):
            self._rec[k].elem = elem
            for i in range(6):
                self._rec[k].weight[i] = w[i]
        self.nrec = len(receivers)
        self.p = OrcProblem(T, dt, dp(lay.A), dp(lay.B), dp(lay.Dx), dp(lay.Dz), dp(lay.Q), dp(lay.S),
                            dp(lay.absorb_fun) if lay.absorb_fun is not None else dptr(), lp(lay.ibe),
                            lp(lay.node_neighbor), lp(lay.neighbor), lp(lay.edge_flag), len(sources), self._src,
                            dp(self._wav), len(receivers), self._rec)
        self.step_index = 0

    def record(self, f: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """The receivers' values (Record, OpenMP_Function.c:61-72, with interpolation weights) of fields f."""
        rx, rz = np.zeros(max(1, self.nrec)), np.zeros(max(1, self.nrec))
        f = np.ascontiguousarray(f, np.float64)
        oracle().orc_record(C.byref(self.p), dp(f[0]), dp(f[1]), dp(rx), dp(rz))
        return rx[:self.nrec], rz[:self.nrec]

    def run(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        rx = np.zeros((n, max(1, self.nrec)))
        rz = np.zeros((n, max(1, self.nrec)))
        f = self.f
        oracle().orc_run(C.byref(self.p), C.c_long(self.step_index), C.c_long(n), dp(f[0]), dp(f[1]), dp(f[2]),
                         dp(f[3]), dp(f[4]), dp(rx), dp(rz))
        self.step_index += n
        return rx[:, :self.nrec], rz[:, :self.nrec]


def lu_velocity(lay: RefLayout, dt: float, f: np.ndarray):
    oracle().orc_lu_velocity(C.c_long(lay.tri), C.c_double(dt), dp(f[0]), dp(f[1]), dp(f[2]), dp(f[3]), dp(f[4]),
                             dp(lay.A), dp(lay.B), dp(lay.Dx), dp(lay.Dz), dp(lay.Q), dp(lay.S), lp(lay.ibe),
                             lp(lay.node_neighbor), lp(lay.neighbor), lp(lay.edge_flag))


def lu_stress(lay: RefLayout, dt: float, f: np.ndarray):
    oracle().orc_lu_stress(C.c_long(lay.tri), C.c_double(dt), dp(f[2]), dp(f[3]), dp(f[4]), dp(f[0]), dp(f[1]),
                           dp(lay.A), dp(lay.B), dp(lay.Dx), dp(lay.Dz), dp(lay.Q), dp(lay.S), lp(lay.ibe),
                           lp(lay.node_neighbor), lp(lay.neighbor), lp(lay.edge_flag))


def absorb(lay: RefLayout, f: np.ndarray):
    oracle().orc_absorb(C.c_long(lay.tri), dp(f[0]), dp(f[1]), dp(f[2]), dp(f[3]), dp(f[4]), dp(lay.absorb_fun))


# ------------------------------------------------------------------------------------------------
# The reference's own build (oracle/_ref/libref.so).
# ------------------------------------------------------------------------------------------------
class RefSetupOut(C.Structure):
    _fields_ = [(n, dptr) for n in ("Dx", "Dz", "Q", "A", "B", "absorb_fun", "jacobi", "jacobi_edge", "n1", "n2",
                                     "n3", "fun")] + \
               [(n, lptr) for n in ("neighbor", "edge_flag", "node_neighbor", "edge")] + \
               [("dt", C.c_double), ("t0", C.c_double), ("source", C.c_long * 2), ("receiver", C.c_long * 2),
                ("absorb_num", C.c_long), ("inside_num", C.c_long)]


def ref_case(opdir, mesh, epart, src_xz, rec_xz, nsteps, snap_every=0, model=None):
    """Run the reference's setup + time loop on a mesh; returns (setup dict, rec_x, rec_z, snaps)."""
    r = reference()
    T, P = mesh.tri, mesh.point_num
    arrays = {n: np.zeros(T * k) for n, k in (("Dx", 36), ("Dz", 36), ("Q", 30), ("A", 5), ("B", 5),
                                               ("absorb_fun", 6), ("jacobi", 1), ("jacobi_edge", 3), ("n1", 2),
                                               ("n2", 2), ("n3", 2))}
    arrays["fun"] = np.zeros(6)
    iarr = {n: np.zeros(T * k, np.int64) for n, k in (("neighbor", 3), ("edge_flag", 3), ("node_neighbor", 9),
                                                       ("edge", 3))}
    out = RefSetupOut(*([dp(arrays[n]) for n in ("Dx", "Dz", "Q", "A", "B", "absorb_fun", "jacobi", "jacobi_edge",
                                                   "n1", "n2", "n3", "fun")] +
                        [lp(iarr[n]) for n in ("neighbor", "edge_flag", "node_neighbor", "edge")]))
    rx, rz = np.zeros(nsteps), np.zeros(nsteps)
    nsnap = nsteps // snap_every if snap_every else 0
    snaps = np.zeros((max(nsnap, 1), 5, T * 6))
    point = np.ascontiguousarray(mesh.point, np.float64)
    tri = np.ascontiguousarray(mesh.triangle, np.int64)
    nb = np.ascontiguousarray(mesh.neighbor, np.int64)
    attr = np.ascontiguousarray(mesh.attribute, np.int64)
    ep = np.ascontiguousarray(epart, np.int64)
    if model is not None:
        lam, miu, den, vp = [np.ascontiguousarray(a, np.float64) for a in (model.lam, model.miu, model.den, model.vp)]
        mats = [dp(lam), dp(miu), dp(den), dp(vp)]
    else:
        mats = [dptr(), dptr(), dptr(), dptr()]
    rc = r.ref_case(str(opdir).encode(), C.c_long(T), C.c_long(P), dp(point), lp(tri), lp(nb), lp(attr), lp(ep),
                    *mats, C.c_double(src_xz[0]), C.c_double(src_xz[1]), C.c_double(rec_xz[0]),
                    C.c_double(rec_xz[1]), C.c_long(nsteps), C.c_long(snap_every), C.byref(out), dp(rx), dp(rz),
                    dp(snaps))
    assert rc == 0, f"ref_case failed ({rc})"
    res = dict(arrays)
    res.update(iarr)
    res.update(dt=out.dt, t0=out.t0, source=tuple(out.source), receiver=tuple(out.receiver),
               absorb_num=out.absorb_num, inside_num=out.inside_num)
    return res, rx, rz, snaps[:nsnap]
