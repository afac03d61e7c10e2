"""Python side of the request loop (include/dgx_serve.h): a Server handle, request construction on a
resident mesh, and an open-loop Poisson request stream with latency percentiles."""
from __future__ import annotations

import ctypes as C
import dataclasses
import time

import numpy as np

from . import lib as L
from .mesh import Mesh
from .problem import PI_REFERENCE, Model, Sponge


@dataclasses.dataclass
class ResidentMesh:
    """A mesh made resident on the server, with what is needed to build requests for it."""
    mesh_id: int
    mesh: Mesh
    model: Model
    dt: float
    locator: C.c_void_p
    keep: list


class Server:
    def __init__(self, precision: int = L.DGX_FP32, hbm_budget: int = 0, steps_per_block: int = 100,
                 ops: L.Operators | None = None, group_shots: int = 16):
        self.lib = L.load()
        self.ops = ops or L.read_operators()
        cfg = L.ServerConfig(precision, hbm_budget, steps_per_block, group_shots)
        self.group_shots = group_shots
        self.handle = C.c_void_p()
        L.check(self.lib.dgx_server_create(C.byref(cfg), C.byref(self.ops), C.byref(self.handle)), "server_create")
        self.meshes: list[ResidentMesh] = []
        self.steps_per_block = steps_per_block

    def close(self):
        for m in self.meshes:
            if m.locator:
                self.lib.dgx_locator_destroy(m.locator)
                m.locator = C.c_void_p()
        if self.handle:
            self.lib.dgx_server_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, mesh: Mesh, model: Model, sponge: Sponge | None, dt: float | None = None) -> ResidentMesh:
        pt = np.ascontiguousarray(mesh.point, np.float64)
        tr = np.ascontiguousarray(mesh.triangle, np.int64)
        nb = np.ascontiguousarray(mesh.neighbor, np.int64)
        la, mi, de = (np.ascontiguousarray(a, np.float64) for a in (model.lam, model.miu, model.den))
        ab = None
        if sponge is not None:
            ab = np.ones(mesh.tri * 6)
            sp = sponge.params()
            L.check(self.lib.dgx_sponge(mesh.tri, L.dp(pt), L.lp(tr), L.dp(np.ascontiguousarray(
                self.ops.array("FeketeNew"))), C.byref(sp), L.dp(ab)), "dgx_sponge")
        if dt is None:
            dt = self.lib.dgx_cfl_dt(mesh.tri, L.dp(pt), L.lp(tr), L.dp(np.ascontiguousarray(model.vp, np.float64)))
        p = L.ProblemC(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), L.lp(nb), L.dp(la), L.dp(mi), L.dp(de),
                       L.dp(ab) if ab is not None else L.dptr(), dt, 0, 1, 0, 0, L.dptr(), 0,
                       C.POINTER(L.Source)(), 0, C.POINTER(L.Receiver)(), 0, 0)
        mid = C.c_int()
        L.check(self.lib.dgx_server_add_mesh(self.handle, C.byref(p), C.byref(mid)), "server_add_mesh")
        loc = C.c_void_p()
        L.check(self.lib.dgx_locator_create(mesh.tri, mesh.point_num, L.dp(pt), L.lp(tr), C.byref(loc)), "locator")
        rm = ResidentMesh(mid.value, mesh, model, float(dt), loc, [pt, tr])
        self.meshes.append(rm)
        return rm

    def make_request(self, rm: ResidentMesh, shots_xz, f0s, nsteps: int, receivers_xz, kind=L.DGX_SRC_EXPLOSIVE,
                     pi: float = PI_REFERENCE, model: Model | None = None, dt: float | None = None):
        """Locate sources and receivers and build the FP64 wavelet rows (one Ricker per shot).  `model`: the
        request's own material model (per element, caller order) instead of the mesh's; its dt is the CFL
        step of that model unless `dt` is given."""
        n = len(shots_xz)
        if model is not None and dt is None:
            pt = np.ascontiguousarray(rm.mesh.point, np.float64)
            tr = np.ascontiguousarray(rm.mesh.triangle, np.int64)
            dt = self.lib.dgx_cfl_dt(rm.mesh.tri, L.dp(pt), L.lp(tr), L.dp(np.ascontiguousarray(model.vp, np.float64)))
        step = rm.dt if model is None else float(dt)
        srcs = (L.Source * n)()
        wav = np.zeros((n, nsteps))
        for k, ((x, z), f0) in enumerate(zip(shots_xz, f0s)):
            L.check(self.lib.dgx_locate_source(rm.locator, C.byref(self.ops), float(x), float(z), kind, k, k,
                                               C.byref(srcs[k])), "locate_source")
            t0 = self.lib.dgx_ricker_delay(step, float(f0), pi)
            L.check(self.lib.dgx_wavelet(nsteps, step, t0, float(f0), pi, L.dp(wav[k])), "wavelet")
        rxz = np.ascontiguousarray(np.asarray(receivers_xz, np.float64).reshape(-1, 2))
        recs = (L.Receiver * max(1, len(rxz)))()
        if len(rxz):
            L.check(self.lib.dgx_locate_receivers(rm.locator, len(rxz), L.dp(rxz), recs), "locate_receivers")
        req = dict(mesh=rm.mesh_id, srcs=srcs, wav=wav, recs=recs, nrec=len(rxz), nsteps=nsteps, nshots=n, dt=step)
        if model is not None:
            req["model"] = [np.ascontiguousarray(a, np.float64) for a in (model.lam, model.miu, model.den)]
        return req

    def submit(self, req: dict, arrival: float | None = None) -> int:
        arrival = self.now() if arrival is None else arrival
        mats = req.get("model")
        r = L.RequestC(req["mesh"], req["nshots"], req["nsteps"], req["srcs"], L.dp(req["wav"]), req["nrec"],
                       req["recs"], arrival, *([L.dp(a) for a in mats] if mats else [L.dptr()] * 3),
                       req["dt"] if mats else 0.0)
        rid = C.c_long()
        L.check(self.lib.dgx_server_submit(self.handle, C.byref(r), C.byref(rid)), "server_submit")
        return rid.value

    def advance(self, blocks: int = 1) -> int:
        return L.check(self.lib.dgx_server_advance(self.handle, blocks), "server_advance")

    def status(self, rid: int):
        st, ta, td = C.c_int(), C.c_double(), C.c_double()
        L.check(self.lib.dgx_server_status(self.handle, rid, C.byref(st), C.byref(ta), C.byref(td)), "status")
        return st.value, ta.value, td.value

    def result(self, rid: int, req: dict):
        n, s, r = req["nsteps"], req["nshots"], req["nrec"]
        x = np.zeros((s, n, max(r, 1)))
        z = np.zeros((s, n, max(r, 1)))
        L.check(self.lib.dgx_server_result(self.handle, rid, L.dp(x), L.dp(z)), "server_result")
        return x[..., :r], z[..., :r]

    def release(self, rid: int):
        L.check(self.lib.dgx_server_release(self.handle, rid), "server_release")

    def stats(self) -> dict:
        st = L.ServerStats()
        L.check(self.lib.dgx_server_get_stats(self.handle, C.byref(st)), "server_stats")
        return {f: getattr(st, f) for f, _ in L.ServerStats._fields_}

    def now(self) -> float:
        return self.lib.dgx_server_now()

    def drain(self):
        while True:
            s = self.stats()
            if s["queued"] == 0 and s["running"] == 0:
                return
            self.advance(1)
            t = self.stats()
            if t["running"] == 0 and t["queued"] == s["queued"] and t["done"] == s["done"]:
                raise L.DgxError("queued requests cannot be admitted (HBM budget)")


def run_open_loop(server: Server, requests: list[dict], arrivals: np.ndarray, keep_results: bool = False,
                  progress_every: float = 10.0):
    """Submit request i when the server's clock passes t_start + arrivals[i]; advance whenever work is
    running, otherwise wait for the next arrival.  Returns per-request latencies (completion minus
    arrival, seconds), the wall time and the server statistics."""
    import sys
    order = np.argsort(arrivals)
    t_start = server.now()
    t_log = t_start
    ids = {}
    nxt = 0
    done = set()
    results = {}
    while len(done) < len(requests):
        now = server.now()
        while nxt < len(order) and t_start + arrivals[order[nxt]] <= now:
            i = int(order[nxt])
            ids[i] = server.submit(requests[i], t_start + arrivals[i])
            nxt += 1
        st = server.stats()
        if st["running"] == 0 and st["queued"] == 0:
            if nxt < len(order):
                time.sleep(max(0.0, min(0.05, t_start + arrivals[order[nxt]] - server.now())))
            continue
        server.advance(1)
        if progress_every and server.now() - t_log > progress_every:
            t_log = server.now()
            s = server.stats()
            print(f"  t={t_log - t_start:.1f}s submitted={nxt} done={len(done)} running={s['running']} "
                  f"queued={s['queued']} step={s['global_step']}", file=sys.stderr, flush=True)
        for i, rid in ids.items():
            if i in done:
                continue
            state, ta, td = server.status(rid)
            if state == L.__dict__.get("DGX_REQ_DONE", 2):
                done.add(i)
                if keep_results:
                    results[i] = server.result(rid, requests[i])
                server.release(rid) if not keep_results else None
                requests[i]["latency"] = td - (t_start + arrivals[i])
                requests[i]["t_done"] = td - t_start
    wall = server.now() - t_start
    lat = np.array([requests[i]["latency"] for i in range(len(requests))])
    return lat, wall, server.stats(), results
