"""Snapshots (include/dgx.h: dgx_get_fields, dgx_snapshot_begin / dgx_snapshot_wait): the fields come back
in the caller's element order, reordered on the device and downloaded through pinned memory.

* a snapshot begun at step n and collected after more steps were enqueued equals the fields of a twin
  solver read at step n, bit for bit, and the stepping it overlapped is unchanged;
* get_fields while a snapshot is in flight (the host-side reorder path) returns the same values;
* the element and shot-innermost layouts, FP64 and FP32;
* a second begin before the wait, and a wait without a begin, are refused (DGX_ESTATE).
The values themselves are checked against the oracle by the parity tests, which read snapshots through
dgx_get_fields."""
import numpy as np
import pytest

import dgx
from dgx import lib as L
from helpers import product_setup

pytestmark = pytest.mark.gpu


def workload(nshots):
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(30, 30, 2000.0 / 30, diagonal="random", seed=11, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    model = dgx.Model.by_attribute(mesh.attribute)
    src = [dgx.SourceSpec(400.0 + 90.0 * k, 1150.0, L.DGX_SRC_EXPLOSIVE, k, 15.0) for k in range(nshots)]
    rec = np.stack([np.linspace(150.0, 1850.0, 9), np.full(9, 1460.0)], axis=1)
    return dgx.configs.Workload("snap", mesh, model, sp, src, rec, 200, 0)


@pytest.mark.parametrize("precision,layout,nshots", [
    (L.DGX_FP64, L.DGX_LAYOUT_ELEMENT, 1), (L.DGX_FP32, L.DGX_LAYOUT_ELEMENT, 1),
    (L.DGX_FP64, L.DGX_LAYOUT_SHOT_INNER, 16), (L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, 16)])
def test_snapshot_overlapping_steps_equals_fields_at_its_step(precision, layout, nshots):
    print(f"case snapshot precision {precision} layout {layout} shots {nshots}", flush=True)
    st = product_setup(workload(nshots))
    shot = nshots - 1
    a = dgx.Solver(st, precision, layout, 16 if nshots > 1 else 0)
    b = dgx.Solver(st, precision, layout, 16 if nshots > 1 else 0)
    a.step(90)
    b.step(90)
    a.snapshot_begin(shot, 31)
    a.step(37)                       # enqueued while the snapshot is in flight
    inflight = a.fields(shot)        # host-side path while the snapshot owns the buffers
    snap = a.snapshot_wait()
    ref = b.fields(shot)             # the twin at step 90, through the snapshot path
    assert np.array_equal(snap, ref)
    b.step(37)
    assert np.array_equal(inflight, b.fields(shot))
    assert np.array_equal(a.fields(shot), b.fields(shot))
    assert np.any(snap != 0.0)
    a.close()
    b.close()


def test_partial_mask_and_state_errors():
    print("case snapshot mask and errors", flush=True)
    st = product_setup(workload(1))
    s = dgx.Solver(st, L.DGX_FP64)
    s.step(60)
    full = s.fields(0)
    s.snapshot_begin(0, 0b10010)     # Vz and Txz only
    with pytest.raises(L.DgxError):
        s.snapshot_begin(0, 1)       # one snapshot in flight per solver
    part = s.snapshot_wait()
    assert np.array_equal(part[[1, 4]], full[[1, 4]])
    assert not np.any(part[[0, 2, 3]])
    with pytest.raises(L.DgxError):
        s.snapshot_wait()            # nothing in flight
    with pytest.raises(L.DgxError):
        s.snapshot_begin(0, 0)       # empty mask
    s.close()
