"""Shared test helpers: run a workload through the oracle (CPU, reference layout) and through libdgx.

The oracle side derives everything itself (operator files, sponge, dt, wavelet, source and receiver
location) so that it stays independent of the product's host setup; the tests compare the two."""
from __future__ import annotations

import numpy as np

import dgx
import oracle_py as O
from dgx import lib as L

OPDIR = L.OPERATOR_DIR


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


class OracleCase:
    def __init__(self, w: dgx.configs.Workload, pi: float = 3.1415):
        self.w = w
        self.ops = O.read_operators(OPDIR)
        sp = w.sponge
        self.sponge = O.SpongeParams(sp.bound_x, sp.bound_z, sp.model_w, sp.model_h, sp.beta, sp.pi)
        m = w.mesh
        self.lay = O.setup_reference_layout(m.point, m.triangle, m.neighbor, w.model.lam, w.model.miu, w.model.den,
                                            self.ops, self.sponge)
        self.dt = O.time_step(m.point, m.triangle, w.model.vp)
        self.freqs = []
        self.src = []
        for s in w.sources:
            if s.f0 not in self.freqs:
                self.freqs.append(s.f0)
            e, fun = O.locate_source(m.point, m.triangle, self.ops["V"], s.x, s.z)
            self.src.append((e, s.kind, fun, s.shot, self.freqs.index(s.f0)))
        self.t0 = [O.ricker_delay(self.dt, f, pi) for f in self.freqs]
        self.wav = np.stack([O.wavelet(w.nsteps, self.dt, t0, f, pi) for t0, f in zip(self.t0, self.freqs)])
        self.rec = [O.locate_receiver(m.point, m.triangle, x, z) for x, z in w.receivers]

    def run(self, shot: int = 0, nsteps: int | None = None, snap_every: int = 0):
        """Returns (rec_x [n, nrec], rec_z, {step: fields[5, Tri*6]})."""
        n = nsteps or self.w.nsteps
        srcs = [(e, k, f) for (e, k, f, sh, wi) in self.src if sh == shot]
        wi = [wi for (e, k, f, sh, wi) in self.src if sh == shot]
        assert len(set(wi)) <= 1
        run = O.Run(self.lay, self.dt, srcs, self.wav[wi[0] if wi else 0], self.rec)
        xs, zs, snaps = [], [], {}
        chunk = snap_every or n
        done = 0
        while done < n:
            k = min(chunk, n - done)
            x, z = run.run(k)
            xs.append(x)
            zs.append(z)
            done += k
            if snap_every:
                snaps[done] = run.f.copy()
        return np.concatenate(xs), np.concatenate(zs), snaps


def product_setup(w: dgx.configs.Workload, nshots=None) -> dgx.Setup:
    return dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge, nshots=nshots)


def product_run(setup: dgx.Setup, precision=L.DGX_FP64, nsteps=None, snap_every=0, shot=0, chunk=None,
                step_kernel=L.DGX_STEP_AUTO, **solver_kw):
    s = dgx.Solver(setup, precision, step_kernel=step_kernel, **solver_kw)
    n = nsteps or setup.nsteps
    snaps = {}
    step = snap_every or chunk or n
    done = 0
    while done < n:
        k = min(step, n - done)
        s.step(k)
        done += k
        if snap_every:
            snaps[done] = s.fields(shot)
    x, z = s.traces(shot)
    s.close()
    return x[:n], z[:n], snaps
