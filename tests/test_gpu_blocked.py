"""The blocked kernel (several whole steps per launch, rings of elements recomputed per tile) against the
half-step kernel pair and the oracle.

* K = 1, 2, 3 steps per launch, FP64 and FP32, on the reference's two-media configuration (sponge on all
  sides, a force source) with a second shot carrying an explosive source: traces and all five fields at
  step counts that mix graph replays (16 launches) with eager launches of fewer than K steps, against the
  half-step kernels (FP64 to 1e-13, FP32 to 1e-5: the same arithmetic, contracted into FMAs differently);
* FP64 on the 2 km homogeneous case (1500 steps) against the oracle to the plain path's 1e-12;
* FP32 on the 640k-element gradient mesh (1000 tiles of ~640 elements) against the half-step kernels;
* the split entry points act on the state the blocked steps leave (it alternates between two copies).
"""
import numpy as np
import pytest

import dgx
from dgx import lib as L
from helpers import OracleCase, product_run, product_setup, rel_l2
from test_gpu_parity import two_media_workload

pytestmark = pytest.mark.gpu


def two_shot_workload(nsteps):
    w = two_media_workload(nsteps)
    w.sources = [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0),
                 dgx.SourceSpec(700.0, 900.0, L.DGX_SRC_EXPLOSIVE, 1, 20.0)]
    w.receivers = np.array([[1600.0, 1600.0], [333.3, 1211.7], [1000.0, 500.0], [150.0, 1750.0]])
    return w


def run(st, precision, snap_every, shot, **kw):
    return product_run(st, precision, snap_every=snap_every, shot=shot, **kw)


@pytest.mark.parametrize("K", [1, 2, 3])
@pytest.mark.parametrize("precision,tol", [(L.DGX_FP64, 1e-13), (L.DGX_FP32, 1e-5)])
def test_blocked_equals_half_step(K, precision, tol):
    print(f"case blocked K={K} precision {precision} two media, 2 shots", flush=True)
    w = two_shot_workload(161)
    st = product_setup(w)
    for shot in (0, 1):
        blk = run(st, precision, 53, shot, step_kernel=L.DGX_STEP_BLOCKED, block_steps=K)
        half = run(st, precision, 53, shot, step_kernel=L.DGX_STEP_HALF)
        for a, b in ((blk[0], half[0]), (blk[1], half[1])):
            assert rel_l2(a, b) <= tol, (shot, rel_l2(a, b))
        assert sorted(blk[2]) == [53, 106, 159, 161]
        for k in blk[2]:
            for c in range(5):
                e = rel_l2(blk[2][k][c], half[2][k][c])
                assert e <= tol, (shot, k, c, e)


def test_blocked_fp64_matches_oracle():
    w = dgx.configs.homogeneous_2km(1500)
    print(f"case blocked K=2 fp64 homogeneous_2km {w.mesh.tri} elements, 1500 steps", flush=True)
    oc = OracleCase(w)
    ox, oz, osnap = oc.run(snap_every=500)
    gx, gz, gsnap = product_run(product_setup(w), L.DGX_FP64, snap_every=500, step_kernel=L.DGX_STEP_BLOCKED,
                                block_steps=2)
    ex, ez = rel_l2(gx, ox), rel_l2(gz, oz)
    print(f"blocked fp64 vs oracle: traces Vx {ex:.3e} Vz {ez:.3e}", flush=True)
    assert ex <= 1e-12 and ez <= 1e-12
    for k, f in osnap.items():
        for c in range(5):
            assert rel_l2(gsnap[k][c], f[c]) <= 1e-12, (k, c)


@pytest.mark.timeout(600)
def test_blocked_fp32_gradient_640k():
    w = dgx.configs.gradient_8x4km(300)
    src = w.sources[0]
    t = np.arange(-100.0, 101.0, 10.0)
    w.receivers = np.concatenate([w.receivers[::20], np.stack([src.x + t, np.full(t.size, src.z)], 1)])
    print(f"case blocked K=2 fp32 {w.name}, 300 steps", flush=True)
    st = product_setup(w)
    s = dgx.Solver(st, L.DGX_FP32, step_kernel=L.DGX_STEP_BLOCKED, block_steps=2)
    info = s.block_info()
    s.close()
    print(f"tiling: {info}", flush=True)
    assert info["steps_per_launch"] == 2 and info["own"] == w.mesh.tri
    blk = product_run(st, L.DGX_FP32, snap_every=300, step_kernel=L.DGX_STEP_BLOCKED, block_steps=2)
    half = product_run(st, L.DGX_FP32, snap_every=300, step_kernel=L.DGX_STEP_HALF)
    e = max(rel_l2(blk[0], half[0]), rel_l2(blk[1], half[1]))
    print(f"blocked vs half-step fp32: traces {e:.3e}", flush=True)
    assert e <= 1e-5
    for c in range(5):
        assert rel_l2(blk[2][300][c], half[2][300][c]) <= 1e-5, c


def test_split_entry_points_after_blocked_steps():
    """After an odd number of blocked launches the state lives in the second copy: dgx_get_fields,
    dgx_set_fields and the split kernels act on it, and stepping on from a state set by the caller gives the
    half-step kernels' result."""
    w = two_media_workload(40)
    st = product_setup(w)
    b = dgx.Solver(st, L.DGX_FP64, step_kernel=L.DGX_STEP_BLOCKED, block_steps=2)
    h = dgx.Solver(st, L.DGX_FP64, step_kernel=L.DGX_STEP_HALF)
    b.step(6)  # three launches
    h.step(6)
    assert rel_l2(b.fields(), h.fields()) <= 1e-13
    rng = np.random.default_rng(3)
    f = rng.standard_normal((5, w.mesh.tri * 6)) * np.array([1, 1, 1e6, 1e6, 1e6])[:, None]
    b.set_fields(f)
    h.set_fields(f)
    b.absorb_boundary()
    h.absorb_boundary()
    assert np.array_equal(b.fields(), h.fields())
    b.step(9)
    h.step(9)
    g1, g2 = b.fields(), h.fields()
    for c in range(5):
        assert rel_l2(g1[c], g2[c]) <= 1e-13, c
    b.close()
    h.close()


def test_compensated_fp32_removes_accumulation_error():
    """The 2 km homogeneous case at a quarter of its CFL step (6000 steps over the same 4.6 s).  Storing each
    value once per step in FP32 adds a rounding error that random-walks with the step count (about sqrt(N)); the
    compensated FP32 update (a low word per value, TwoSum) keeps what the store drops, leaving only the rounding
    inside one step's increment, which does not grow with N.  Errors against the FP64 run on the same dt, taken
    in absolute terms and divided by the run's peak field norm (the sponge drains the field toward the end, so
    an error relative to the last snapshot alone grows as the field dies, not as rounding accumulates):
    snapshots at steps 3000 and 6000 and the traces.  Compensated: within 1e-5 and at most 0.6 of plain FP32's
    error (measured 0.33 on snapshots, 0.42 on traces)."""
    w = dgx.configs.homogeneous_2km(6000)
    w.receivers = np.concatenate([w.receivers[::10], np.array([[1000.0 + d, 1500.0] for d in (-150, -90, 60, 120)])])
    st0 = product_setup(w)
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, 6000, sponge=w.sponge, dt=st0.dt / 4)
    print(f"case compensated fp32, homogeneous_2km at dt/4, 6000 steps", flush=True)
    ref = product_run(st, L.DGX_FP64, snap_every=3000)
    plain = product_run(st, L.DGX_FP32, snap_every=3000, step_kernel=L.DGX_STEP_BLOCKED)
    comp = product_run(st, L.DGX_FP32_COMPENSATED, snap_every=3000, step_kernel=L.DGX_STEP_BLOCKED)

    def snap_err(run):
        worst = 0.0
        for c in range(5):
            peak = max(np.linalg.norm(ref[2][k][c]) for k in (3000, 6000))
            worst = max(worst, max(np.linalg.norm(run[2][k][c] - ref[2][k][c]) for k in (3000, 6000)) / peak)
        return worst

    ep = max(rel_l2(plain[0], ref[0]), rel_l2(plain[1], ref[1]))
    ec = max(rel_l2(comp[0], ref[0]), rel_l2(comp[1], ref[1]))
    sp, sc = snap_err(plain), snap_err(comp)
    print(f"traces: fp32 {ep:.3e} compensated {ec:.3e}; snapshots/peak: fp32 {sp:.3e} compensated {sc:.3e}",
          flush=True)
    assert ec <= 1e-5 and sc <= 1e-5
    assert ec <= 0.6 * ep and sc <= 0.6 * sp


def test_compensated_fields_round_trip_and_split_refused():
    w = two_media_workload(20)
    st = product_setup(w)
    s = dgx.Solver(st, L.DGX_FP32_COMPENSATED)
    rng = np.random.default_rng(9)
    f = rng.standard_normal((5, w.mesh.tri * 6)) * np.array([1, 1, 1e6, 1e6, 1e6])[:, None]
    s.set_fields(f)
    g = s.fields()
    for c in range(5):  # high + low word hold the value to about 2^-48
        assert rel_l2(g[c], f[c]) <= 1e-13, c
    with pytest.raises(L.DgxError):
        s.lu_velocity()
    with pytest.raises(L.DgxError):
        s.record(0)
    s.step(5)
    h = dgx.Solver(st, L.DGX_FP32, step_kernel=L.DGX_STEP_BLOCKED)
    h.set_fields(f)
    h.step(5)
    a, b = s.fields(), h.fields()
    for c in range(5):
        assert rel_l2(a[c], b[c]) <= 1e-5, c
    s.close()
    h.close()
