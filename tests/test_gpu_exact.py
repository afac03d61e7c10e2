"""DGX_FP64_REFERENCE against the oracle, bit for bit.

The oracle equals the reference build bit for bit on the reference's own fixtures (test_oracle.py); this mode
evaluates the reference's per-element arrays in the reference's order without fused multiply-adds, so its fields
and traces must equal the oracle's exactly (np.array_equal: equal values, the sign of exact zeros aside) at every
snapshot, on every step's trace, and through the split entry points."""
import numpy as np
import pytest

import dgx
import oracle_py as O
from dgx import lib as L
from helpers import OracleCase, product_run, product_setup
from test_gpu_parity import two_media_workload

pytestmark = pytest.mark.gpu
EXACT = L.DGX_FP64_REFERENCE


def same_inputs(w, oc):
    """The product's host setup gives the oracle's dt, wavelets, sponge, source vectors and receiver weights."""
    st = product_setup(w)
    assert st.dt == oc.dt and np.array_equal(st.wavelets, oc.wav)
    if oc.lay.absorb_fun is not None:
        assert np.array_equal(st.absorb_fun, oc.lay.absorb_fun)
    for s, (e, kind, fun, shot, wi) in zip(st.sources, oc.src):
        assert s.elem == e and s.kind == kind and np.array_equal(np.array(list(s.fun)), fun)
    for r, (e, wgt) in zip(st.receivers, oc.rec):
        assert r.elem == e and np.array_equal(np.array(list(r.weight)), wgt)
    return st


def assert_identical(g, o, every):
    gx, gz, gsnap = g
    ox, oz, osnap = o
    bad_x = np.flatnonzero(np.any(gx != ox, axis=1))
    bad_z = np.flatnonzero(np.any(gz != oz, axis=1))
    assert bad_x.size == 0 and bad_z.size == 0, ("first differing trace steps", bad_x[:3], bad_z[:3])
    assert sorted(gsnap) == sorted(osnap)
    for k, f in osnap.items():
        for c in range(5):
            assert np.array_equal(gsnap[k][c], f[c]), (k, c, np.abs(gsnap[k][c] - f[c]).max())


def test_exact_homogeneous_1500_steps():
    """The 2 km homogeneous case (20,000 elements, explosive vertex source, 101 receivers): all 1500 steps."""
    w = dgx.configs.homogeneous_2km(1500)
    print(f"case {w.name} exact, {w.nsteps} steps", flush=True)
    oc = OracleCase(w)
    st = same_inputs(w, oc)
    o = oc.run(snap_every=250)
    g = product_run(st, EXACT, snap_every=250)
    assert_identical(g, o, 250)


def test_exact_reference_configuration_sponge():
    """The reference's own configuration (water over rock, sponge, one vertex receiver), 400 steps."""
    w = two_media_workload(400)
    print("case two_media exact, 400 steps", flush=True)
    oc = OracleCase(w)
    st = same_inputs(w, oc)
    assert_identical(product_run(st, EXACT, snap_every=100), oc.run(snap_every=100), 100)


def test_exact_several_sources_and_interpolating_receivers():
    """Force and explosive sources of one shot (two of them on one element, which must add in list order),
    receivers off the vertices (P2 weights), a velocity gradient with sponge; eager steps and graph replays
    mixed by a snapshot cadence of 45."""
    w = dgx.configs.gradient_8x4km(270, nx=120, nz=60, h=50.0, source_x=3000.0, source_depth=800.0)
    s0 = w.sources[0]
    w.sources = [s0, dgx.SourceSpec(s0.x, s0.z, L.DGX_SRC_EXPLOSIVE, 0, s0.f0),
                 dgx.SourceSpec(s0.x, s0.z, L.DGX_SRC_FORCE_Z, 0, s0.f0),
                 dgx.SourceSpec(1800.0, 1500.0, L.DGX_SRC_EXPLOSIVE, 0, s0.f0),
                 dgx.SourceSpec(4700.0, 2500.0, L.DGX_SRC_FORCE_Z, 0, s0.f0)]
    w.receivers = np.array([[2900.0, 2150.3], [2333.3, 2521.7], [1850.7, 1460.2], [4650.1, 2480.9]])
    print("case gradient 6x3 km exact, 5 sources, 270 steps", flush=True)
    oc = OracleCase(w)
    st = same_inputs(w, oc)
    assert_identical(product_run(st, EXACT, snap_every=45), oc.run(snap_every=45), 45)


def test_exact_split_entry_points():
    """dgx_absorb_boundary, dgx_source_kernel, dgx_lu_velocity, dgx_lu_stress_elastic and dgx_record on a random
    state, each against the oracle's restatement of the reference function, bit for bit."""
    w = two_media_workload(10)
    w.sources = [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0),
                 dgx.SourceSpec(800.0, 1000.0, L.DGX_SRC_EXPLOSIVE, 0, 15.0)]
    w.receivers = np.array([[1600.0, 1600.0], [333.3, 1211.7]])
    oc = OracleCase(w)
    st = same_inputs(w, oc)
    T = w.mesh.tri
    rng = np.random.default_rng(7)
    f0 = rng.standard_normal((5, T * 6)) * np.array([1, 1, 1e6, 1e6, 1e6])[:, None]
    s = dgx.Solver(st, EXACT)
    s.set_fields(f0)
    ref = f0.copy()
    s.absorb_boundary()
    O.absorb(oc.lay, ref)
    assert np.array_equal(s.fields(), ref)
    s.source_kernel(0.37)
    for e, kind, fun, _, _ in oc.src:
        for c in ((1,) if kind == 0 else (2, 3)):
            ref[c, (e - 1) * 6:e * 6] = ref[c, (e - 1) * 6:e * 6] + fun * 0.37
    assert np.array_equal(s.fields(), ref)
    s.lu_velocity()
    O.lu_velocity(oc.lay, oc.dt, ref)
    assert np.array_equal(s.fields(), ref)
    s.lu_stress_elastic()
    O.lu_stress(oc.lay, oc.dt, ref)
    assert np.array_equal(s.fields(), ref)
    s.record(3)
    x, z = s.traces()
    for k, (e, wgt) in enumerate(oc.rec):
        ax = az = 0.0
        for i in range(6):
            ax = ax + wgt[i] * ref[0, (e - 1) * 6 + i]
            az = az + wgt[i] * ref[1, (e - 1) * 6 + i]
        assert x[3, k] == ax and z[3, k] == az, k
    s.close()


def test_exact_two_shots_are_independent():
    """Shot 1 of a two-shot exact solver equals a one-shot run of its source alone, bit for bit."""
    w = two_media_workload(120)
    w.sources = [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0),
                 dgx.SourceSpec(800.0, 1000.0, L.DGX_SRC_EXPLOSIVE, 1, 15.0)]
    both = dgx.Solver(product_setup(w), EXACT)
    both.step(120)
    w1 = two_media_workload(120)
    w1.sources = [dgx.SourceSpec(800.0, 1000.0, L.DGX_SRC_EXPLOSIVE, 0, 15.0)]
    alone = product_run(product_setup(w1), EXACT)
    x, z = both.traces(1)
    assert np.array_equal(x, alone[0]) and np.array_equal(z, alone[1])
    both.close()
