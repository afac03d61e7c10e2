"""Request loop: requests that carry their own material model, and admission under a failed allocation.

* Two material models on one resident mesh (the mesh's own two-media model and a gradient model with its own
  CFL time step): each request's traces are bit-identical alone and in a mix with requests of the other
  model, in FP32 and FP64, and the FP64 server reproduces the oracle run on that request's model to the plain
  path's bar (the reference sets materials per element, DG_leap_Frog_OpenMP.c:236-265, and derives dt from
  them, fun_DG.c:179-199).
* A device allocation that fails while a request is admitted leaves every request queued or running as it
  was: the call reports DGX_ENOMEM, the next calls admit and finish every request, and the results equal
  those of a server that never failed.
"""
import numpy as np
import pytest

import dgx
import dgx.serve as S
from dgx import lib as L
from helpers import OracleCase, rel_l2

pytestmark = pytest.mark.gpu


def mesh_and_models():
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    m = dgx.structured(40, 40, 50.0, diagonal="random", seed=3, sponge_box=sp.box(),
                       attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    own = dgx.Model.by_attribute(m.attribute)
    depth = 2000.0 - m.centroids()[:, 1]
    vp = 2200.0 + 1.2 * depth
    grad = dgx.Model.from_velocities(vp, vp / np.sqrt(3.0), 310.0 * vp ** 0.25)
    return sp, m, own, grad


# (model: 0 own / 1 gradient, shots, steps, f0)
SPECS = [(0, 3, 180, 15.0), (1, 2, 220, 18.0), (1, 4, 130, 12.0), (0, 1, 90, 25.0)]


def make(sv, rm, grad, spec, k):
    mo, ns, nsteps, f0 = spec
    rng = np.random.default_rng(700 + k)
    xz = np.stack([rng.uniform(300, 1700, ns), rng.uniform(600, 1400, ns)], 1)
    rec = np.stack([np.linspace(200, 1800, 13), np.full(13, 1450.0)], 1)
    req = sv.make_request(rm, xz, [f0] * ns, nsteps, rec, model=grad if mo == 1 else None)
    return req, xz, rec


def server(precision, group=8):
    sp, m, own, grad = mesh_and_models()
    sv = S.Server(precision, 0, 48 if group == 1 else 50, group_shots=group)
    rm = sv.add_mesh(m, own, sp)
    return sv, rm, sp, m, own, grad


@pytest.mark.parametrize("group", [8, 1])
@pytest.mark.parametrize("precision", [L.DGX_FP32, L.DGX_FP64])
def test_two_models_on_one_mesh_alone_and_mixed(precision, group):
    print(f"case serve two models, precision {precision}, group {group}", flush=True)
    sv, rm, sp, m, own, grad = server(precision, group)
    built = [make(sv, rm, grad, spec, k) for k, spec in enumerate(SPECS)]
    assert built[1][0]["dt"] != built[0][0]["dt"]  # the gradient model has its own CFL step
    ids = [sv.submit(built[0][0]), sv.submit(built[1][0])]
    sv.advance(1)
    ids += [sv.submit(built[2][0]), sv.submit(built[3][0])]
    sv.drain()
    mixed = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
    sv.close()
    for k, spec in enumerate(SPECS):
        sv, rm, sp, m, own, grad = server(precision, group)
        req, _, _ = make(sv, rm, grad, spec, k)
        rid = sv.submit(req)
        sv.drain()
        ax, az = sv.result(rid, req)
        sv.close()
        assert np.array_equal(mixed[k][0], ax) and np.array_equal(mixed[k][1], az), k
    if precision != L.DGX_FP64:
        return
    worst = 0.0
    for k in (1, 3):  # one request of each model against the oracle on its own model and dt
        mo, ns, nsteps, f0 = SPECS[k]
        _, xz, rec = built[k]
        srcs = [dgx.SourceSpec(float(a), float(b), L.DGX_SRC_EXPLOSIVE, s, f0) for s, (a, b) in enumerate(xz)]
        oc = OracleCase(dgx.configs.Workload("models", m, grad if mo == 1 else own, sp, srcs, rec, nsteps, 0))
        assert oc.dt == built[k][0]["dt"]
        for s in range(ns):
            ox, oz, _ = oc.run(shot=s)
            worst = max(worst, rel_l2(mixed[k][0][s], ox), rel_l2(mixed[k][1][s], oz))
    print(f"two-model fp64 serving vs oracle: worst trace rel L2 {worst:.3e}")
    assert worst <= 1e-12


def test_groups_hold_one_model():
    """Two 2-shot requests on different models of one mesh, 8-shot groups: two groups, one per model; a third
    request of the first model joins that model's group."""
    sv, rm, sp, m, own, grad = server(L.DGX_FP32)
    a = make(sv, rm, grad, (0, 2, 100, 15.0), 1)[0]
    b = make(sv, rm, grad, (1, 2, 100, 15.0), 2)[0]
    c = make(sv, rm, grad, (0, 2, 100, 15.0), 3)[0]
    for r in (a, b, c):
        sv.submit(r)
    sv.advance(1)
    st = sv.stats()
    assert st["running"] == 3 and st["batches"] == 2, st
    sv.drain()
    sv.close()


def test_failed_allocation_leaves_requests_intact(monkeypatch):
    """The third request-buffer allocation fails (the second request's wavelet): the call reports it and
    places nothing of that request; further calls finish all three, with the results of a server without
    the failure."""
    print("case serve failed allocation", flush=True)
    specs = [(0, 1, 120, 15.0), (0, 1, 150, 20.0), (1, 1, 110, 12.0)]

    def run(fail):
        if fail:
            monkeypatch.setenv("DGX_SERVE_FAIL_ALLOC", "3")
        sv, rm, sp, m, own, grad = server(L.DGX_FP32)
        monkeypatch.delenv("DGX_SERVE_FAIL_ALLOC", raising=False)
        built = [make(sv, rm, grad, spec, 40 + k) for k, spec in enumerate(specs)]
        ids = [sv.submit(b[0]) for b in built]
        if fail:
            with pytest.raises(L.DgxError, match=r"\(-4\)"):
                sv.advance(1)
            st = sv.stats()
            assert st["running"] == 1 and st["queued"] == 2, st
        sv.drain()
        assert all(sv.status(i)[0] == 2 for i in ids)
        out = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
        sv.close()
        return out

    failed, clean = run(True), run(False)
    for k in range(len(specs)):
        assert np.array_equal(failed[k][0], clean[k][0]) and np.array_equal(failed[k][1], clean[k][1]), k
