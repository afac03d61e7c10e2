"""dgx_get_fields_noalloc: the fields of one shot through the library's static staging array, with no
allocation, against dgx_get_fields (the snapshot path, which stages through buffers it allocates) on the same
solver state: bit for bit, over calls in a row on meshes that shrink from one call to the next (the array keeps
the previous call's larger contents), over FP64, FP32 shot-innermost and compensated FP32, with subsets of the
fields, and with the size limit refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import dgx
from dgx import lib as L

pytestmark = pytest.mark.gpu


def workload(n, nshots, nsteps):
    """n x n cells of 20 m (2 n^2 triangles), water over rock, a sponge, explosive sources (all five fields
    carry signal)."""
    W = 20.0 * n
    sp = dgx.Sponge.for_domain(W, W, 100.0)
    mesh = dgx.structured(n, n, 20.0, diagonal="random", seed=5, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 0.7 * W, 1, 3))
    model = dgx.Model.by_attribute(mesh.attribute)
    srcs = [dgx.SourceSpec(W * (0.3 + 0.2 * k), W * 0.45, L.DGX_SRC_EXPLOSIVE, k, 25.0) for k in range(nshots)]
    return dgx.build_setup(mesh, model, srcs, np.array([[W / 2, W / 2]]), nsteps, sponge=sp, nshots=nshots)


def noalloc(s, shot, mask):
    T = s.setup.mesh.tri
    out = np.full((5, T * 6), np.nan)
    ptrs = [L.dp(out[c]) if (mask >> c) & 1 else L.dptr() for c in range(5)]
    rc = s.lib.dgx_get_fields_noalloc(s.handle, shot, *ptrs)
    return rc, out


def check(s, shot, mask, what):
    print(f"case noalloc {what}: {s.setup.mesh.tri} elements, shot {shot}, mask {mask}", flush=True)
    rc, got = noalloc(s, shot, mask)
    assert rc == L.DGX_OK, s.lib.dgx_last_error()
    ref = s.fields(shot)
    for c in range(5):
        if (mask >> c) & 1:
            assert np.array_equal(got[c], ref[c]), (what, c)
            assert np.any(ref[c] != 0), (what, c)  # the comparison saw signal
        else:
            assert np.all(np.isnan(got[c])), (what, c)  # fields outside the request are not written


def test_noalloc_matches_get_fields_on_shrinking_meshes():
    # 1: FP64 element layout, 28,800 elements, all five fields of shot 1
    s = dgx.Solver(workload(120, 2, 60), L.DGX_FP64)
    s.step(60)
    check(s, 1, 31, "fp64 element layout")
    s.close()
    # 2: smaller, FP32 shot-innermost (3 shots), two of the fields of shot 2
    s = dgx.Solver(workload(50, 3, 40), L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER)
    s.step(40)
    check(s, 2, (1 << 1) | (1 << 4), "fp32 shot-innermost")
    s.close()
    # 3: smaller still, compensated FP32 on the blocked kernel (high + low words reordered together)
    s = dgx.Solver(workload(30, 1, 30), L.DGX_FP32_COMPENSATED, step_kernel=L.DGX_STEP_BLOCKED)
    s.step(30)
    check(s, 0, 31, "compensated fp32")
    s.close()


def test_noalloc_size_limit():
    """270 x 270 cells: 145,800 elements, 7.0 MB per field: all five (35 MB) exceed the 32 MiB array and are
    refused with nothing done; four fit, and a call after the refusal works."""
    s = dgx.Solver(workload(270, 1, 20), L.DGX_FP32)
    s.step(20)
    T = s.setup.mesh.tri
    assert 5 * T * 6 * 8 > L.DGX_NOALLOC_SCRATCH_BYTES >= 4 * T * 6 * 8
    rc, got = noalloc(s, 0, 31)
    assert rc == L.DGX_EARG and b"static array" in s.lib.dgx_last_error()
    assert np.all(np.isnan(got))
    check(s, 0, 15, "four fields of the largest mesh")
    rc, _ = noalloc(s, 1, 1)  # no such shot
    assert rc == L.DGX_EARG
    s.close()
