"""Problem assembly and the solver handle, the Python mirror of the reference driver's setup
(DG_leap_Frog_OpenMP.c:190-367) and time loop (DG_leap_Frog_OpenMP.c:397-446) on top of libdgx.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Sequence

import numpy as np

from . import lib as L
from .mesh import Mesh

PI_REFERENCE = 3.1415  # dg.h:6


@dataclasses.dataclass
class Model:
    """Per-element material (the reference sets it per element by attribute, DG_leap_Frog_OpenMP.c:248-265)."""
    lam: np.ndarray
    miu: np.ndarray
    den: np.ndarray
    vp: np.ndarray

    @staticmethod
    def from_velocities(vp, vs, rho) -> "Model":
        vp = np.asarray(vp, np.float64)
        vs = np.asarray(vs, np.float64)
        rho = np.asarray(rho, np.float64)
        miu = rho * vs * vs
        lam = rho * vp * vp - 2 * miu  # DG_leap_Frog_OpenMP.c:239-246
        return Model(lam, miu, rho.copy(), vp.copy())

    @staticmethod
    def homogeneous(tri: int, vp: float, vs: float, rho: float) -> "Model":
        return Model.from_velocities(np.full(tri, vp), np.full(tri, vs), np.full(tri, rho))

    @staticmethod
    def by_attribute(attribute: np.ndarray) -> "Model":
        """The reference's two media: attributes 1,2 water (1500, 0, 1000), 3,4 rock (3000, 1800, 2200)."""
        a = np.asarray(attribute)
        water = (a == 1) | (a == 2)
        vp = np.where(water, 1500.0, 3000.0)
        vs = np.where(water, 0.0, 1800.0)
        rho = np.where(water, 1000.0, 2200.0)
        return Model.from_velocities(vp, vs, rho)


@dataclasses.dataclass
class SourceSpec:
    x: float
    z: float
    kind: int = L.DGX_SRC_EXPLOSIVE
    shot: int = 0
    f0: float = 15.0


@dataclasses.dataclass
class Sponge:
    bound_x: float
    bound_z: float
    model_w: float
    model_h: float
    beta: float = 0.1  # DG_leap_Frog_OpenMP.c:287
    pi: float = PI_REFERENCE

    @staticmethod
    def for_domain(width: float, height: float, bound: float = 100.0, **kw) -> "Sponge":
        return Sponge(bound, bound, width - 2 * bound, height - 2 * bound, **kw)

    def params(self) -> L.SpongeParams:
        return L.SpongeParams(self.bound_x, self.bound_z, self.model_w, self.model_h, self.beta, self.pi)

    def box(self):
        return (self.bound_x, self.bound_x + self.model_w, self.bound_z, self.bound_z + self.model_h)


@dataclasses.dataclass
class Setup:
    """Everything dgx_create needs, in the reference's layout and numbering."""
    mesh: Mesh
    model: Model
    ops: L.Operators
    absorb_fun: np.ndarray | None
    dt: float
    nsteps: int
    nshots: int
    wavelets: np.ndarray      # [nwavelets, nsteps]
    t0: list[float]
    f0: list[float]
    sources: list[L.Source]
    receivers: list[L.Receiver]
    receiver_xz: np.ndarray
    pi: float = PI_REFERENCE

    @property
    def nrec(self) -> int:
        return len(self.receivers)


def build_setup(mesh: Mesh, model: Model, sources: Sequence[SourceSpec], receiver_xz, nsteps: int, *,
                sponge: Sponge | None, ops: L.Operators | None = None, dt: float | None = None,
                pi: float = PI_REFERENCE, nshots: int | None = None, locate: str = "host") -> Setup:
    """Host-side set-up of a problem (dgx_setup.h).  locate: "host" searches sources and receivers with the
    host grid locator, "device" with the same search on the GPU (dgx_locate_*_device); both give the same
    elements and weights bit for bit."""
    if locate not in ("host", "device"):
        raise ValueError(f"locate must be 'host' or 'device', not {locate!r}")
    lib = L.load()
    ops = ops or L.read_operators()
    point = np.ascontiguousarray(mesh.point, np.float64)
    tri = np.ascontiguousarray(mesh.triangle, np.int64)
    T = mesh.tri
    absorb = None
    if sponge is not None:
        absorb = np.ones(T * 6, np.float64)
        fek = np.ascontiguousarray(ops.array("FeketeNew"))
        sp = sponge.params()
        L.check(lib.dgx_sponge(T, L.dp(point), L.lp(tri), L.dp(fek), C.byref(sp), L.dp(absorb)), "dgx_sponge")
    if dt is None:
        dt = lib.dgx_cfl_dt(T, L.dp(point), L.lp(tri), L.dp(np.ascontiguousarray(model.vp, np.float64)))
    freqs: list[float] = []
    for s in sources:
        if s.f0 not in freqs:
            freqs.append(s.f0)
    rxz = np.ascontiguousarray(np.asarray(receiver_xz, np.float64).reshape(-1, 2))
    recs = (L.Receiver * max(1, len(rxz)))()
    if locate == "device":
        n = len(sources)
        out = (L.Source * max(1, n))()
        if n:
            sxz = np.ascontiguousarray([[s.x, s.z] for s in sources], np.float64)
            kind = np.ascontiguousarray([s.kind for s in sources], np.int32)
            shot = np.ascontiguousarray([s.shot for s in sources], np.int32)
            wrow = np.ascontiguousarray([freqs.index(s.f0) for s in sources], np.int64)
            L.check(lib.dgx_locate_sources_device(T, mesh.point_num, L.dp(point), L.lp(tri), C.byref(ops), n,
                                                  L.dp(sxz), kind.ctypes.data_as(C.POINTER(C.c_int)),
                                                  shot.ctypes.data_as(C.POINTER(C.c_int)), L.lp(wrow), out),
                    "dgx_locate_sources_device")
        srcs = [out[i] for i in range(n)]
        if len(rxz):
            L.check(lib.dgx_locate_receivers_device(T, mesh.point_num, L.dp(point), L.lp(tri), len(rxz), L.dp(rxz),
                                                    recs), "dgx_locate_receivers_device")
    else:
        loc = C.c_void_p()
        L.check(lib.dgx_locator_create(T, mesh.point_num, L.dp(point), L.lp(tri), C.byref(loc)), "locator")
        try:
            srcs = []
            for s in sources:
                out1 = L.Source()
                L.check(lib.dgx_locate_source(loc, C.byref(ops), s.x, s.z, s.kind, s.shot, freqs.index(s.f0),
                                              C.byref(out1)), "dgx_locate_source")
                srcs.append(out1)
            if len(rxz):
                L.check(lib.dgx_locate_receivers(loc, len(rxz), L.dp(rxz), recs), "dgx_locate_receivers")
        finally:
            lib.dgx_locator_destroy(loc)
    recs = [recs[i] for i in range(len(rxz))]
    t0s, wav = [], np.zeros((max(1, len(freqs)), nsteps), np.float64)
    for k, f in enumerate(freqs):
        t0 = lib.dgx_ricker_delay(dt, f, pi)
        t0s.append(t0)
        row = np.zeros(nsteps, np.float64)
        L.check(lib.dgx_wavelet(nsteps, dt, t0, f, pi, L.dp(row)), "dgx_wavelet")
        wav[k] = row
    ns = nshots if nshots is not None else (max([s.shot for s in sources]) + 1 if sources else 1)
    return Setup(mesh, model, ops, absorb, float(dt), int(nsteps), int(ns), wav, t0s, freqs, srcs, recs, rxz, pi)


class Solver:
    """Owns one dgx_solver handle (device-resident fields for ``nshots`` shots on one mesh)."""

    def __init__(self, setup: Setup, precision: int = L.DGX_FP64, layout: int = L.DGX_LAYOUT_AUTO,
                 shot_tile: int = 0, step_kernel: int = L.DGX_STEP_AUTO, wg_waves: int = 0, block_steps: int = 0):
        self.lib = L.load()
        self.setup = setup
        m = setup.mesh
        self._keep = [np.ascontiguousarray(m.point, np.float64), np.ascontiguousarray(m.triangle, np.int64),
                      np.ascontiguousarray(m.neighbor, np.int64), np.ascontiguousarray(setup.model.lam, np.float64),
                      np.ascontiguousarray(setup.model.miu, np.float64),
                      np.ascontiguousarray(setup.model.den, np.float64),
                      np.ascontiguousarray(setup.wavelets, np.float64)]
        pt, tr, nb, la, mi, de, wv = self._keep
        ab = None
        if setup.absorb_fun is not None:
            ab = np.ascontiguousarray(setup.absorb_fun, np.float64)
            self._keep.append(ab)
        srcs = (L.Source * max(1, len(setup.sources)))(*setup.sources)
        recs = (L.Receiver * max(1, len(setup.receivers)))(*setup.receivers)
        self._keep += [srcs, recs]
        p = L.ProblemC(m.tri, m.point_num, L.dp(pt), L.lp(tr), L.lp(nb), L.dp(la), L.dp(mi), L.dp(de),
                       L.dp(ab) if ab is not None else L.dptr(), setup.dt, precision, setup.nshots, setup.nsteps,
                       wv.shape[0], L.dp(wv), len(setup.sources), srcs, len(setup.receivers), recs, layout,
                       shot_tile, step_kernel, wg_waves, block_steps)
        self.handle = C.c_void_p()
        L.check(self.lib.dgx_create(C.byref(p), C.byref(setup.ops), C.byref(self.handle)), "dgx_create")
        self.precision = precision

    def close(self):
        if self.handle:
            self.lib.dgx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        L.check(self.lib.dgx_step(self.handle, n), "dgx_step")

    def absorb_boundary(self):
        L.check(self.lib.dgx_absorb_boundary(self.handle), "dgx_absorb_boundary")

    def source_kernel(self, w: float):
        L.check(self.lib.dgx_source_kernel(self.handle, w), "dgx_source_kernel")

    def lu_velocity(self):
        L.check(self.lib.dgx_lu_velocity(self.handle), "dgx_lu_velocity")

    def lu_stress_elastic(self):
        L.check(self.lib.dgx_lu_stress_elastic(self.handle), "dgx_lu_stress_elastic")

    def record(self, t: int):
        L.check(self.lib.dgx_record(self.handle, t), "dgx_record")

    def synchronize(self):
        L.check(self.lib.dgx_synchronize(self.handle), "dgx_synchronize")

    @property
    def current_step(self) -> int:
        return int(self.lib.dgx_current_step(self.handle))

    def fields(self, shot: int = 0) -> np.ndarray:
        """[5, Tri*6] Vx, Vz, Txx, Tzz, Txz in the caller's element order."""
        T = self.setup.mesh.tri
        out = np.zeros((5, T * 6), np.float64)
        L.check(self.lib.dgx_get_fields(self.handle, shot, *[L.dp(out[c]) for c in range(5)]), "dgx_get_fields")
        return out

    def snapshot_begin(self, shot: int = 0, mask: int = 31):
        """Start copying the fields in `mask` (bit c: Vx, Vz, Txx, Tzz, Txz) of one shot to the host; the
        next steps may be enqueued before snapshot_wait collects them."""
        L.check(self.lib.dgx_snapshot_begin(self.handle, shot, mask), "dgx_snapshot_begin")
        self._snap_mask = mask

    def snapshot_wait(self, out: np.ndarray | None = None) -> np.ndarray:
        """[5, Tri*6] fields of the snapshot begun last, into `out` if given (rows outside the snapshot's
        mask are left as they are: zero in a new array)."""
        T = self.setup.mesh.tri
        if out is None:
            out = np.zeros((5, T * 6), np.float64)
        assert out.shape == (5, T * 6) and out.dtype == np.float64 and out.flags.c_contiguous
        L.check(self.lib.dgx_snapshot_wait(self.handle, *[L.dp(out[c]) for c in range(5)]), "dgx_snapshot_wait")
        return out

    def set_fields(self, f: np.ndarray, shot: int = 0):
        f = np.ascontiguousarray(f, np.float64)
        L.check(self.lib.dgx_set_fields(self.handle, shot, *[L.dp(f[c]) for c in range(5)]), "dgx_set_fields")

    def traces(self, shot: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """([nsteps, nrec], [nsteps, nrec]) recorded Vx, Vz."""
        n, r = self.setup.nsteps, self.setup.nrec
        x = np.zeros((n, max(r, 1)), np.float64)
        z = np.zeros((n, max(r, 1)), np.float64)
        L.check(self.lib.dgx_get_traces(self.handle, shot, L.dp(x), L.dp(z)), "dgx_get_traces")
        return x[:, :r], z[:, :r]

    def kernel_bench(self, which: int, n: int) -> float:
        """Average ms of one fused kernel (0 velocity, 1 stress, 2 full step, 3 blocked: one launch of
        block_steps steps) over n back-to-back launches.  Leaves the fields invalid."""
        ms = C.c_double()
        L.check(self.lib.dgx_kernel_bench(self.handle, which, n, C.byref(ms)), "dgx_kernel_bench")
        return ms.value

    def layout_info(self):
        """(layout, shot_tile, {tile: best ms per step over the workgroup sizes} from the autotuner)."""
        lay, tile = C.c_int(), C.c_int()
        tiles = (C.c_int * 32)()
        ms = (C.c_double * 32)()
        n = L.check(self.lib.dgx_layout_info(self.handle, C.byref(lay), C.byref(tile), tiles, ms, 32), "layout_info")
        best = {}
        for k in range(min(n, 32)):
            best[tiles[k]] = min(best.get(tiles[k], float("inf")), ms[k])
        return lay.value, tile.value, best

    def tile_info(self):
        """(shot_tile, wavefronts per workgroup, {(tile, waves): ms per step} from the autotuner)."""
        tile, waves = C.c_int(), C.c_int()
        tl, wv, ms = (C.c_int * 32)(), (C.c_int * 32)(), (C.c_double * 32)()
        n = L.check(self.lib.dgx_tile_info(self.handle, C.byref(tile), C.byref(waves), tl, wv, ms, 32), "tile_info")
        return tile.value, waves.value, {(tl[k], wv[k]): ms[k] for k in range(n)}

    @property
    def compiled_operators(self) -> bool:
        """True when the element kernels run with the P2 operators compiled in (the problem's operators
        equal the shipped ones bit for bit and DGX_GENERIC_OPERATORS=1 was not set at creation)."""
        return bool(L.check(self.lib.dgx_compiled_operators(self.handle), "dgx_compiled_operators"))

    def block_info(self) -> dict:
        """Tiling of the blocked kernel: steps per launch, tiles, own elements and lanes (threads holding an
        element) in total; empty for the other step kernels."""
        k, nt, own, lanes = C.c_int(), C.c_long(), C.c_long(), C.c_long()
        L.check(self.lib.dgx_block_info(self.handle, C.byref(k), C.byref(nt), C.byref(own), C.byref(lanes)),
                "dgx_block_info")
        if not k.value:
            return {}
        return {"steps_per_launch": k.value, "tiles": nt.value, "own": own.value, "lanes": lanes.value,
                "recomputed_fraction": (lanes.value - own.value) / max(own.value, 1)}

    def kernel_times(self, n: int) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        L.check(self.lib.dgx_kernel_times(self.handle, n, C.byref(a), C.byref(b)), "dgx_kernel_times")
        return a.value, b.value
