"""GPU parity: libdgx (HIP, gfx950) against the CPU oracle on the same inputs.

Bars: FP64 plain path relative L2 <= 1e-12 on every trace set and snapshot; FP32 path relative L2 <= 1e-4
on traces versus the FP64 oracle."""
import numpy as np
import pytest

import dgx
import oracle_py as O
from dgx import lib as L
from helpers import OracleCase, product_run, product_setup, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-4


@pytest.fixture(scope="module")
def homog_case():
    w = dgx.configs.homogeneous_2km(1500)
    print(f"case homogeneous_2km tri={w.mesh.tri} steps={w.nsteps}", flush=True)
    oc = OracleCase(w)
    ox, oz, osnap = oc.run(snap_every=250)
    return w, oc, ox, oz, osnap


@pytest.mark.parametrize("step_kernel", [L.DGX_STEP_AUTO, L.DGX_STEP_HALF])
def test_homogeneous_fp64_traces_and_snapshots(homog_case, step_kernel):
    """FP64 against the oracle: the default kernel (the blocked kernel on this 20k-element mesh) and the
    half-step pair."""
    w, oc, ox, oz, osnap = homog_case
    st = product_setup(w)
    assert st.dt == oc.dt
    gx, gz, gsnap = product_run(st, L.DGX_FP64, snap_every=250, step_kernel=step_kernel)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"fp64 (step kernel {step_kernel}) traces rel L2: Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= FP64_TOL and ez <= FP64_TOL
    for k, f in osnap.items():
        for c in range(5):
            e = rel_l2(gsnap[k][c], f[c])
            assert e <= FP64_TOL, (k, c, e)


@pytest.mark.parametrize("step_kernel", [L.DGX_STEP_AUTO, L.DGX_STEP_HALF])
def test_homogeneous_fp32_traces(homog_case, step_kernel):
    """FP32 against the FP64 oracle: the default kernel (the blocked kernel for FP32) and the half-step pair."""
    w, oc, ox, oz, osnap = homog_case
    st = product_setup(w)
    gx, gz, gsnap = product_run(st, L.DGX_FP32, snap_every=250, step_kernel=step_kernel)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"fp32 (step kernel {step_kernel}) traces rel L2: Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= FP32_TOL and ez <= FP32_TOL
    for k, f in osnap.items():
        print(f"fp32 snapshot {k}: Vx {rel_l2(gsnap[k][0], f[0]):.3e} Vz {rel_l2(gsnap[k][1], f[1]):.3e}")


def two_media_workload(nsteps=400):
    """The reference's own configuration: water (attr 1) over rock (attr 3), vertical force at
    (1200, 1200), one vertex receiver at (1600, 1600), dg.h sponge on a 2 km square."""
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(40, 40, 50.0, diagonal="random", seed=3, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    model = dgx.Model.by_attribute(mesh.attribute)
    src = [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0)]
    return dgx.configs.Workload("two_media", mesh, model, sp, src, np.array([[1600.0, 1600.0]]), nsteps, 100)


def test_reference_configuration_fp64():
    w = two_media_workload(400)
    oc = OracleCase(w)
    ox, oz, osnap = oc.run(snap_every=100)
    gx, gz, gsnap = product_run(product_setup(w), L.DGX_FP64, snap_every=100)
    assert rel_l2(gx, ox) <= FP64_TOL and rel_l2(gz, oz) <= FP64_TOL
    for k, f in osnap.items():
        for c in range(5):
            assert rel_l2(gsnap[k][c], f[c]) <= FP64_TOL, (k, c)


def test_split_entry_points_match_oracle_functions():
    """dgx_absorb_boundary / dgx_source_kernel / dgx_lu_velocity / dgx_lu_stress_elastic on a random
    state against the oracle's restatement of each reference function."""
    w = two_media_workload(10)
    oc = OracleCase(w)
    st = product_setup(w)
    rng = np.random.default_rng(7)
    T = w.mesh.tri
    f0 = rng.standard_normal((5, T * 6)) * np.array([1, 1, 1e6, 1e6, 1e6])[:, None]
    s = dgx.Solver(st, L.DGX_FP64)
    s.set_fields(f0)
    ref = f0.copy()
    s.absorb_boundary()
    O.absorb(oc.lay, ref)
    assert np.array_equal(s.fields(), ref)
    e, kind, fun, _, _ = oc.src[0]
    s.source_kernel(0.37)
    ref[1, (e - 1) * 6:(e - 1) * 6 + 6] += fun * 0.37
    assert rel_l2(s.fields()[1], ref[1]) <= 1e-15
    s.lu_velocity()
    O.lu_velocity(oc.lay, oc.dt, ref)
    g = s.fields()
    for c in (0, 1):
        assert rel_l2(g[c], ref[c]) <= 1e-13, c
    s.set_fields(ref)
    s.lu_stress_elastic()
    O.lu_stress(oc.lay, oc.dt, ref)
    g = s.fields()
    for c in (2, 3, 4):
        assert rel_l2(g[c], ref[c]) <= 1e-13, c
    s.close()


def test_graph_replay_equals_eager():
    w = two_media_workload(96)
    st = product_setup(w)
    a = product_run(st, L.DGX_FP64, chunk=96)       # 3 graph replays
    b = product_run(st, L.DGX_FP64, chunk=7)        # eager launches only (chunks < 32)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("precision,tol", [(L.DGX_FP64, 1e-13), (L.DGX_FP32, 1e-5)])
def test_full_step_kernel_equals_half_step_kernels(precision, tol):
    """One full-step kernel per step (two field copies, ring velocities recomputed per tile) against the
    two in-place half-step kernels: traces and every snapshot, at step counts that leave the state in
    either copy and mix eager steps with graph replays."""
    w = two_media_workload(141)
    st = product_setup(w)
    print(f"case two_media 40x40 full vs half, precision {precision}", flush=True)
    full = product_run(st, precision, snap_every=47, step_kernel=L.DGX_STEP_FULL)
    half = product_run(st, precision, snap_every=47, step_kernel=L.DGX_STEP_HALF)
    for a, b in ((full[0], half[0]), (full[1], half[1])):
        assert rel_l2(a, b) <= tol
    assert sorted(full[2]) == [47, 94, 141]
    for k in full[2]:
        for c in range(5):
            e = rel_l2(full[2][k][c], half[2][k][c])
            assert e <= tol, (k, c, e)


@pytest.mark.parametrize("precision,tol", [(L.DGX_FP64, 1e-13), (L.DGX_FP32, 1e-5)])
def test_compiled_operators_equal_argument_operators(precision, tol, monkeypatch):
    """The velocity kernel with the P2 operators compiled in against the one that reads them from its
    argument.  Same operator values; with constants the compiler may fuse a different product of a sum
    into an FMA (x*2 becomes x+x, for one), so the two agree to rounding, not bit for bit: traces and
    every snapshot within the full-step bar.  A change to one lift entry that survives the cast to the
    working precision selects the argument variant."""
    import dataclasses
    w = two_media_workload(120)
    st = product_setup(w)
    print(f"case two_media compiled vs argument operators, precision {precision}", flush=True)
    s = dgx.Solver(st, precision)
    assert s.compiled_operators
    s.close()
    comp = product_run(st, precision, snap_every=40)
    monkeypatch.setenv("DGX_GENERIC_OPERATORS", "1")
    s = dgx.Solver(st, precision)
    assert not s.compiled_operators
    s.close()
    arg = product_run(st, precision, snap_every=40)
    monkeypatch.delenv("DGX_GENERIC_OPERATORS")
    ex, ez = rel_l2(comp[0], arg[0]), rel_l2(comp[1], arg[1])
    print(f"compiled vs argument operators: traces rel L2 Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= tol and ez <= tol
    assert sorted(comp[2]) == [40, 80, 120]
    for k in comp[2]:
        for c in range(5):
            e = rel_l2(comp[2][k][c], arg[2][k][c])
            assert e <= tol, (k, c, e)
    ops = L.Operators.from_buffer_copy(st.ops)
    ops.S1[0] *= 1.0 + 2.0 ** -20
    s = dgx.Solver(dataclasses.replace(st, ops=ops), precision)
    assert not s.compiled_operators
    s.close()


def test_shots_are_independent():
    w = two_media_workload(200)
    w.sources = [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0),
                 dgx.SourceSpec(800.0, 1000.0, L.DGX_SRC_EXPLOSIVE, 1, 15.0)]
    st = product_setup(w)
    both = dgx.Solver(st, L.DGX_FP64)
    both.step(200)
    for shot in (0, 1):
        w1 = two_media_workload(200)
        w1.sources = [w.sources[shot]]
        w1.sources[0].shot = 0
        alone = product_run(product_setup(w1), L.DGX_FP64)
        x, z = both.traces(shot)
        assert np.array_equal(x, alone[0]) and np.array_equal(z, alone[1]), shot
    both.close()


def test_split_record_and_explosive_source_match_oracle():
    """dgx_record (Record, OpenMP_Function.c:61-72, with interpolation weights) on a random state against the
    oracle's orc_record; dgx_source_kernel with an explosive source (fun*w into Txx and Tzz) against the
    oracle's orc_source_kernel applied to both stresses, as orc_step does."""
    import ctypes as C
    w = two_media_workload(10)
    w.sources = [dgx.SourceSpec(800.0, 1000.0, L.DGX_SRC_EXPLOSIVE, 0, 15.0)]
    w.receivers = np.array([[1600.0, 1600.0], [333.3, 1211.7], [1000.0, 500.0], [1742.0, 1890.0]])
    oc = OracleCase(w)
    st = product_setup(w)
    rng = np.random.default_rng(21)
    T = w.mesh.tri
    f0 = rng.standard_normal((5, T * 6)) * np.array([1, 1, 1e6, 1e6, 1e6])[:, None]
    s = dgx.Solver(st, L.DGX_FP64)
    s.set_fields(f0)
    s.record(4)
    x, z = s.traces(0)
    run = O.Run(oc.lay, oc.dt, [], oc.wav[:1], oc.rec)
    ox, oz = run.record(f0)
    # the same six products summed in the same order; the device may fuse them into FMAs
    assert rel_l2(x[4], ox) <= 1e-15 and rel_l2(z[4], oz) <= 1e-15
    assert not np.any(x[:4]) and not np.any(x[5:])
    e, kind, fun, _, _ = oc.src[0]
    assert kind == L.DGX_SRC_EXPLOSIVE
    s.source_kernel(-0.61)
    ref = f0.copy()
    for c in (2, 3):
        O.oracle().orc_source_kernel(O.dp(ref[c]), C.c_long(e), O.dp(np.ascontiguousarray(fun)), C.c_double(-0.61))
    g = s.fields()
    assert np.array_equal(g[[0, 1, 4]], ref[[0, 1, 4]])
    for c in (2, 3):
        assert rel_l2(g[c], ref[c]) <= 1e-15, c
    s.close()
