"""Shot-innermost (batched) layout: each shot of a batched run equals its own single-shot run on the
element layout (FP64, rounding-level), FP32 batched traces stay within 1e-4 of the FP64 oracle, and the
autotuner leaves the solver in its initial state."""
import numpy as np
import pytest

import dgx
from dgx import lib as L
from helpers import OracleCase, product_run, product_setup, rel_l2

pytestmark = pytest.mark.gpu


def batched_workload(nshots, nsteps=200):
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(40, 40, 50.0, diagonal="random", seed=3, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    model = dgx.Model.by_attribute(mesh.attribute)
    xs = np.linspace(300.0, 1700.0, nshots)
    src = [dgx.SourceSpec(float(x), 1200.0 + 37.0 * (k % 5), L.DGX_SRC_EXPLOSIVE if k % 2 else L.DGX_SRC_FORCE_Z,
                          k, 15.0) for k, x in enumerate(xs)]
    rec = np.stack([np.linspace(150.0, 1850.0, 23), np.full(23, 1460.0)], axis=1)
    return dgx.configs.Workload("batched", mesh, model, sp, src, rec, nsteps, 0)


def single_shot(w, k):
    w1 = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, [dgx.SourceSpec(w.sources[k].x, w.sources[k].z,
                              w.sources[k].kind, 0, w.sources[k].f0)], w.receivers, w.nsteps, 0)
    return product_run(product_setup(w1), L.DGX_FP64)


@pytest.mark.parametrize("tile", [8, 16, 64])
def test_batched_fp64_equals_single_shot_runs(tile):
    w = batched_workload(16 if tile < 64 else 64, 150)
    st = product_setup(w)
    s = dgx.Solver(st, L.DGX_FP64, L.DGX_LAYOUT_SHOT_INNER, tile)
    assert s.layout_info()[:2] == (L.DGX_LAYOUT_SHOT_INNER, tile)
    s.step(w.nsteps)
    for k in (0, 1, len(w.sources) // 2, len(w.sources) - 1):
        x1, z1, _ = single_shot(w, k)
        x, z = s.traces(k)
        assert rel_l2(x, x1) <= 1e-13 and rel_l2(z, z1) <= 1e-13, (tile, k)
    f = s.fields(3)
    s.close()
    w3 = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, [dgx.SourceSpec(w.sources[3].x, w.sources[3].z,
                              w.sources[3].kind, 0, 15.0)], w.receivers, w.nsteps, 0)
    one = dgx.Solver(product_setup(w3), L.DGX_FP64, L.DGX_LAYOUT_ELEMENT)
    one.step(w.nsteps)
    g = one.fields(0)
    one.close()
    for c in range(5):
        assert rel_l2(f[c], g[c]) <= 1e-13, c


@pytest.mark.parametrize("nshots,tile", [(8, 8), (16, 16), (32, 32)])
def test_interleaved_rows_equal_one_element_rows(nshots, tile):
    """A shot tile that holds every shot of a row stores the 64/tile elements of a wavefront interleaved in
    each node row; half that tile keeps one element per row.  FP64 traces and fields of the two agree to
    rounding, and fields written with set_fields read back unchanged."""
    print(f"case interleaved rows, {nshots} shots, tile {tile}", flush=True)
    w = batched_workload(nshots, 120)
    st = product_setup(w)
    runs = []
    for t in (tile, tile // 2):
        s = dgx.Solver(st, L.DGX_FP64, L.DGX_LAYOUT_SHOT_INNER, t)
        s.step(w.nsteps)
        runs.append(([s.traces(k) for k in range(nshots)], [s.fields(k) for k in (0, nshots - 1)]))
        if t == tile:
            f = s.fields(1) * 0.5 + 1.0
            s.set_fields(f, 1)
            assert np.array_equal(s.fields(1), f)
            assert np.array_equal(s.fields(0), runs[-1][1][0])
        s.close()
    (ta, fa), (tb, fb) = runs
    for k in range(nshots):
        assert rel_l2(ta[k][0], tb[k][0]) <= 1e-13 and rel_l2(ta[k][1], tb[k][1]) <= 1e-13, k
    for x, y in zip(fa, fb):
        assert np.abs(x - y).max() <= 1e-13 * max(np.abs(y).max(), 1e-300)


def test_batched_fp32_against_fp64_oracle():
    w = batched_workload(16, 300)
    oc = OracleCase(w)
    s = dgx.Solver(product_setup(w), L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, 16)
    s.step(w.nsteps)
    worst = 0.0
    for k in (0, 5, 15):
        ox, oz, _ = oc.run(shot=k)
        x, z = s.traces(k)
        worst = max(worst, rel_l2(x, ox), rel_l2(z, oz))
    s.close()
    print(f"batched fp32 worst trace rel L2 vs fp64 oracle: {worst:.3e}")
    assert worst <= 1e-4


def test_workgroup_size_does_not_change_results():
    """The wavefronts per workgroup only regroup elements into workgroups: bit-identical traces."""
    w = batched_workload(16, 100)
    st = product_setup(w)
    ref = None
    for waves in (2, 4, 8):
        s = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, 16, wg_waves=waves)
        assert s.tile_info()[:2] == (16, waves)
        s.step(w.nsteps)
        tr = np.stack([np.stack(s.traces(k)) for k in (0, 15)])
        s.close()
        if ref is None:
            ref = tr
        assert np.array_equal(tr, ref), waves


def test_autotuner_restores_initial_state():
    w = batched_workload(32, 120)
    st = product_setup(w)
    auto = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, 0)
    lay, tile, _ = auto.layout_info()
    tile2, waves, table = auto.tile_info()
    print("autotune table (ms/step):", table, "chosen", tile, waves)
    # shot tiles 4 and 8 with 4 wavefronts, 16 and 32 with 2, 4 and 8 (32 padded shots: no tile 64)
    assert lay == L.DGX_LAYOUT_SHOT_INNER and tile == tile2 and (tile, waves) in table and len(table) == 8
    assert auto.current_step == 0
    assert not np.any(auto.fields(7))
    auto.step(w.nsteps)
    fixed = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, tile, wg_waves=waves)
    fixed.step(w.nsteps)
    for k in (0, 31):
        assert np.array_equal(auto.traces(k)[0], fixed.traces(k)[0])
    auto.close()
    fixed.close()
