"""The measured workloads on their own meshes, at full size, on the GPU.

Each test builds the workload's mesh and model with the same seeded generator the measurements use and
checks the solver on it at a reduced step count:

* gradient 8 km x 4 km, 640k jittered triangles: plain FP64 against the oracle (the reference's loops,
  OpenMP_Function.c:25-306) to 1e-12 on traces and all five fields; FP32 against the oracle to 1e-4 on the
  trace set, on every trace at or above 0.1 % of the peak trace norm, and on every trace against that peak;
* slope 6 km x 3 km, ~1.2M Delaunay triangles, water over three dipping layers, 64 FP32 shots in the
  shot-innermost layout and on the blocked kernel (the FP32 default): sampled shots against FP64 single-shot
  GPU runs to the per-trace bar; sources and receivers located on the GPU equal the host locator's; one shot
  FP64 against the oracle to 1e-12;
* the eight resident serving meshes (20k - 2M triangles, random layered models): short requests on several
  meshes are bit-identical alone and in the mix; one request served in FP64 matches the oracle to 1e-12;
* carbonate 4 km x 3 km, ~2.4M triangles refined to 1 m around 300 vugs and 50 fractures: one shot FP64
  against the oracle to 1e-12; shots served in FP32 against FP64 runs to 1e-4.

Receiver sets include a cross through the source, so that traces carry the direct wave within the reduced
step counts (the workloads' surface receivers see it only after thousands of steps).
"""
import numpy as np
import pytest

import dgx
import dgx.serve as S
from dgx import lib as L
from dgx.accuracy import snapshot_errors, trace_errors
from helpers import OracleCase, product_run, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-4
# Per-trace FP32 bar: a trace (one component of one receiver) whose FP64 norm is at least FLOOR of the set's
# largest trace norm is within FP32_TOL of its own norm; every trace is within FP32_TOL of that largest norm.
# FP32 rounding leaves an absolute error of about the same size on every trace (well under 1e-6 of the peak on
# these workloads), so a trace's relative error grows as its amplitude falls: at 0.1-1 % of the peak it
# measures up to 2.5e-5, and traces still nearly silent (below 0.1 %) reach relative errors of order 1.
FLOOR = 1e-3


def cross(x, z, half=200.0, step=10.0):
    """Receivers on a horizontal and a vertical line through (x, z)."""
    t = np.arange(-half, half + step / 2, step)
    return np.concatenate([np.stack([x + t, np.full(t.size, z)], 1), np.stack([np.full(t.size, x), z + t], 1)])


def with_receivers(w, rec, nsteps):
    return dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, w.sources, rec, nsteps, 0)


def check_fp32_traces(x, z, xr, zr, what):
    e = trace_errors(x, z, xr, zr, floor=FLOOR)
    print(f"{what}: {e}", flush=True)
    assert e["all"] is not None and e["all"] <= FP32_TOL
    assert e["worst"] is not None and e["worst"] <= FP32_TOL
    assert e["worst_vs_peak"] <= FP32_TOL
    return e


@pytest.fixture(scope="module")
def gradient():
    w = dgx.configs.gradient_8x4km(1200)
    src = w.sources[0]
    rec = np.concatenate([w.receivers, cross(src.x, src.z)])
    w = with_receivers(w, rec, 1200)
    print(f"case {w.name}: {w.mesh.tri} elements, {len(rec)} receivers, 1200 steps", flush=True)
    oc = OracleCase(w)
    ox, oz, osnap = oc.run(snap_every=600)
    return w, oc, ox, oz, osnap


@pytest.mark.timeout(600)
def test_gradient_640k_fp64_against_oracle(gradient):
    w, oc, ox, oz, osnap = gradient
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    assert st.dt == oc.dt
    gx, gz, gsnap = product_run(st, L.DGX_FP64, snap_every=600)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"gradient fp64 traces rel L2: Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= FP64_TOL and ez <= FP64_TOL
    for k, f in osnap.items():
        for c in range(5):
            e = rel_l2(gsnap[k][c], f[c])
            assert e <= FP64_TOL, (k, c, e)


@pytest.mark.timeout(600)
def test_gradient_640k_reference_arithmetic_bit_identical(gradient):
    """DGX_FP64_REFERENCE on 640,000 elements: traces and both snapshots equal the oracle's bit for bit."""
    w, oc, ox, oz, osnap = gradient
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    assert st.dt == oc.dt and np.array_equal(st.absorb_fun, oc.lay.absorb_fun)
    for s, (e, kind, fun, shot, wi) in zip(st.sources, oc.src):
        assert s.elem == e and np.array_equal(np.array(list(s.fun)), fun)
    for r, (e, wgt) in zip(st.receivers, oc.rec):
        assert r.elem == e and np.array_equal(np.array(list(r.weight)), wgt)
    gx, gz, gsnap = product_run(st, L.DGX_FP64_REFERENCE, snap_every=600)
    assert np.array_equal(gx, ox) and np.array_equal(gz, oz)
    for k, f in osnap.items():
        for c in range(5):
            assert np.array_equal(gsnap[k][c], f[c]), (k, c)


@pytest.mark.timeout(600)
def test_gradient_640k_all_4000_steps():
    """The 640,000-element gradient case over all 4000 steps (32 s of oracle time): the reference-arithmetic
    precision equals the oracle bit for bit at every 1000th step and on every trace sample.  The FP64 default
    (regrouped sums, fused multiply-adds) drifts from it by rounding that grows with the step count: measured
    1.40e-12 / 1.28e-12 on the traces and at most 1.35e-12 on the step-4000 fields, above the 1e-12 bar that it
    meets on the 1200-step runs; it is held here to 3e-12."""
    w = dgx.configs.gradient_8x4km(4000)
    print(f"case {w.name}: {w.mesh.tri} elements, 4000 steps", flush=True)
    oc = OracleCase(w)
    ox, oz, osnap = oc.run(snap_every=1000)
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    gx, gz, gsnap = product_run(st, L.DGX_FP64_REFERENCE, snap_every=1000)
    assert np.array_equal(gx, ox) and np.array_equal(gz, oz)
    for k, f in osnap.items():
        for c in range(5):
            assert np.array_equal(gsnap[k][c], f[c]), (k, c)
    gx, gz, gsnap = product_run(st, L.DGX_FP64, snap_every=1000)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    worst = max(rel_l2(gsnap[k][c], f[c]) for k, f in osnap.items() for c in range(5))
    print(f"fp64 default after 4000 steps: traces {ex:.3e} {ez:.3e}, fields at most {worst:.3e}", flush=True)
    assert max(ex, ez, worst) <= 3e-12


@pytest.mark.timeout(600)
def test_gradient_640k_fp32_against_oracle(gradient):
    w, oc, ox, oz, osnap = gradient
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    gx, gz, gsnap = product_run(st, L.DGX_FP32, snap_every=600)
    check_fp32_traces(gx, gz, ox, oz, "gradient fp32 vs oracle")
    print("gradient fp32 snapshots:", snapshot_errors(gsnap[1200], osnap[1200]), flush=True)


@pytest.mark.timeout(900)
def test_slope_1200k_64_shots():
    w = dgx.configs.slope_6x3km(1500, nshots=64)
    print(f"case {w.name}: {w.mesh.tri} elements, 64 shots, 1500 steps", flush=True)
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    # the 64 sources and 2001 receivers located on the GPU equal the host locator's, bit for bit
    sd = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge, locate="device")
    assert all(bytes(a) == bytes(b) for a, b in zip(st.sources, sd.sources)) and len(sd.sources) == 64
    assert all(bytes(a) == bytes(b) for a, b in zip(st.receivers, sd.receivers)) and len(sd.receivers) == 2001
    # the shot-innermost half-step kernels, and the default FP32 path (the blocked kernel, one row of
    # workgroups per shot)
    b = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER)
    b.step(w.nsteps)
    k2 = dgx.Solver(st, L.DGX_FP32)
    assert k2.block_info()["steps_per_launch"] == 3  # the two field copies of 64 shots exceed the Infinity Cache
    k2.step(w.nsteps)
    for k in (0, 21, 42, 63):
        s = w.sources[k]
        one = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge,
                                   [dgx.SourceSpec(s.x, s.z, s.kind, 0, s.f0)], w.receivers, w.nsteps, 0)
        st1 = dgx.build_setup(one.mesh, one.model, one.sources, one.receivers, one.nsteps, sponge=one.sponge, dt=st.dt)
        x64, z64, _ = product_run(st1, L.DGX_FP64)
        x32, z32 = b.traces(k)
        check_fp32_traces(x32, z32, x64, z64, f"slope shot {k} batched fp32 vs fp64")
        x32, z32 = k2.traces(k)
        check_fp32_traces(x32, z32, x64, z64, f"slope shot {k} blocked fp32 vs fp64")
    b.close()
    k2.close()
    # one shot, FP64, against the oracle
    s = w.sources[21]
    one = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, [dgx.SourceSpec(s.x, s.z, s.kind, 0, s.f0)],
                               np.concatenate([w.receivers[::10], cross(s.x, s.z - 60.0, 50.0)]), 150, 0)
    oc = OracleCase(one)
    ox, oz, osnap = oc.run(snap_every=150)
    st1 = dgx.build_setup(one.mesh, one.model, one.sources, one.receivers, one.nsteps, sponge=one.sponge)
    assert st1.dt == oc.dt
    gx, gz, gsnap = product_run(st1, L.DGX_FP64, snap_every=150)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"slope shot 21 fp64 vs oracle: Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= FP64_TOL and ez <= FP64_TOL
    for c in range(5):
        assert rel_l2(gsnap[150][c], osnap[150][c]) <= FP64_TOL, c


@pytest.mark.timeout(900)
def test_serving_eight_meshes():
    ms = dgx.configs.serving_meshes()
    print(f"case serving meshes {[m.tri for m, _, _ in ms]}", flush=True)
    # (mesh, shots, steps, f0): one or more requests on the smallest, middle and largest meshes
    specs = [(0, 2, 260, 20.0), (7, 3, 200, 12.0), (4, 1, 230, 25.0), (7, 1, 150, 30.0), (2, 4, 180, 15.0),
             (5, 2, 210, 18.0)]

    def make(sv, rms, k, spec):
        mi, ns, nsteps, f0 = spec
        mesh = ms[mi][0]
        W, H = mesh.point[:, 0].max(), mesh.point[:, 1].max()
        rng = np.random.default_rng(900 + k)
        xz = np.stack([rng.uniform(150, W - 150, ns), H - rng.uniform(100, min(600, H - 200), ns)], 1)
        rec = np.concatenate([np.stack([np.linspace(120.0, W - 120.0, 21), np.full(21, H - 5.0)], 1),
                              cross(xz[0, 0], xz[0, 1], 60.0)])
        return sv.make_request(rms[mi], xz, [f0] * ns, nsteps, rec), xz, rec

    sv = S.Server(L.DGX_FP32, 0, 50, group_shots=4)
    rms = [sv.add_mesh(m, md, sp) for m, md, sp in ms]
    built = [make(sv, rms, k, spec) for k, spec in enumerate(specs)]
    ids = [sv.submit(b[0]) for b in built[:3]]
    sv.advance(1)
    ids += [sv.submit(b[0]) for b in built[3:]]
    sv.drain()
    mixed = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
    sv.close()
    for k in (0, 1, 3, 5):
        sv = S.Server(L.DGX_FP32, 0, 50, group_shots=4)
        mi = specs[k][0]
        rm = sv.add_mesh(*ms[mi])
        rms1 = {mi: rm}
        req, _, _ = make(sv, rms1, k, specs[k])
        rid = sv.submit(req)
        sv.drain()
        ax, az = sv.result(rid, req)
        sv.close()
        assert np.array_equal(mixed[k][0], ax) and np.array_equal(mixed[k][1], az), k
    # one request served in FP64 against the oracle (120k-element mesh)
    sv = S.Server(L.DGX_FP64, 0, 50, group_shots=4)
    rms = {2: sv.add_mesh(*ms[2])}
    req, xz, rec = make(sv, rms, 4, (2, 2, 180, 15.0))
    rid = sv.submit(req)
    sv.drain()
    x, z = sv.result(rid, req)
    sv.close()
    mesh, model, sp = ms[2]
    srcs = [dgx.SourceSpec(float(a), float(b), L.DGX_SRC_EXPLOSIVE, s, 15.0) for s, (a, b) in enumerate(xz)]
    oc = OracleCase(dgx.configs.Workload("serving_120k", mesh, model, sp, srcs, rec, 180, 0))
    worst = 0.0
    for s in range(len(xz)):
        ox, oz, _ = oc.run(shot=s)
        worst = max(worst, rel_l2(x[s], ox), rel_l2(z[s], oz))
    print(f"serving 120k fp64 vs oracle: worst trace rel L2 {worst:.3e}", flush=True)
    assert worst <= FP64_TOL


@pytest.mark.timeout(900)
def test_carbonate_2400k():
    w = dgx.configs.carbonate_4x3km(1000, nshots=256)
    print(f"case {w.name}: {w.mesh.tri} elements", flush=True)
    # one shot FP64 against the oracle, 100 steps
    s = w.sources[128]
    one = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, [dgx.SourceSpec(s.x, s.z, s.kind, 0, s.f0)],
                               np.concatenate([w.receivers[::8], cross(s.x, s.z, 20.0, 2.0)]), 100, 0)
    oc = OracleCase(one)
    ox, oz, osnap = oc.run(snap_every=100)
    st1 = dgx.build_setup(one.mesh, one.model, one.sources, one.receivers, 100, sponge=one.sponge)
    assert st1.dt == oc.dt
    gx, gz, gsnap = product_run(st1, L.DGX_FP64, snap_every=100)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"carbonate fp64 vs oracle: Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= FP64_TOL and ez <= FP64_TOL
    for c in range(5):
        assert rel_l2(gsnap[100][c], osnap[100][c]) <= FP64_TOL, c
    # the reference-arithmetic precision on the same inputs (source vector, receiver weights and sponge equal the
    # oracle's bit for bit): the oracle's traces and fields exactly
    assert np.array_equal(np.array(list(st1.sources[0].fun)), oc.src[0][2])
    assert all(np.array_equal(np.array(list(r.weight)), wgt) for r, (e, wgt) in zip(st1.receivers, oc.rec))
    gx, gz, gsnap = product_run(st1, L.DGX_FP64_REFERENCE, snap_every=100)
    assert np.array_equal(gx, ox) and np.array_equal(gz, oz)
    for c in range(5):
        assert np.array_equal(gsnap[100][c], osnap[100][c]), c
    # four shots served in FP32 (one request, 1000 steps) against FP64 single-shot runs
    picks = [0, 85, 170, 255]
    sv = S.Server(L.DGX_FP32, 0, 100, group_shots=4)
    rm = sv.add_mesh(w.mesh, w.model, w.sponge)
    xz = np.array([[w.sources[k].x, w.sources[k].z] for k in picks])
    rec = np.concatenate([w.receivers] + [cross(a, b, 30.0, 3.0) for a, b in xz])
    req = sv.make_request(rm, xz, [30.0] * 4, 1000, rec)
    rid = sv.submit(req)
    sv.drain()
    x32, z32 = sv.result(rid, req)
    sv.close()
    for i, k in enumerate(picks):
        s = w.sources[k]
        st = dgx.build_setup(w.mesh, w.model, [dgx.SourceSpec(s.x, s.z, s.kind, 0, s.f0)], rec, 1000,
                             sponge=w.sponge, dt=rm.dt)
        x64, z64, _ = product_run(st, L.DGX_FP64)
        check_fp32_traces(x32[i], z32[i], x64, z64, f"carbonate shot {k} served fp32 vs fp64")
