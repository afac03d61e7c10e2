"""Trace and snapshot error measures (dgx.accuracy): the numbers every FP32 bar is stated in."""
import numpy as np

from dgx.accuracy import BANDS, rel, trace_errors


def test_rel_silent_reference_is_none_and_nonfinite_is_nan():
    assert rel(np.ones(4), np.zeros(4)) is None
    assert np.isnan(rel(np.array([1.0, np.nan]), np.ones(2)))
    assert rel(np.array([3.0, 4.0]), np.array([0.0, 4.0])) == 0.75


def test_trace_errors_per_component_floor_and_bands():
    """Receiver 0 is strong (norm 10 on Vz, 5 on Vx), receiver 1 weak (norm 0.05 on Vz, silent Vx); an absolute
    error of 1e-3 on every component gives 1e-4 / 2e-4 relative on the strong traces, 2e-2 on the weak one, and
    1e-4 of the peak everywhere."""
    n = 100
    t = np.arange(n)
    xr = np.zeros((n, 2))
    zr = np.zeros((n, 2))
    xr[:, 0] = np.sin(t)
    zr[:, 0] = np.cos(t)
    xr[:, 0] *= 5.0 / np.linalg.norm(xr[:, 0])
    zr[:, 0] *= 10.0 / np.linalg.norm(zr[:, 0])
    zr[:, 1] = 0.05 / np.sqrt(n)
    err = np.full(n, 1e-3 / np.sqrt(n))
    x, z = xr + err[:, None], zr + err[:, None]
    e = trace_errors(x, z, xr, zr)
    assert e["ntraces"] == 4 and np.isclose(e["peak"], 10.0)
    # floor 0.1: Vx0 (0.5 of the peak) and Vz0 are kept; the weak Vz1 (0.005) and silent Vx1 are not
    assert e["kept"] == 2
    assert np.isclose(e["worst"], 1e-3 / 5.0) and e["worst_trace"] == ["Vx", 0] and np.isclose(e["worst_norm"], 0.5)
    assert np.isclose(e["worst_vs_peak"], 1e-3 / 10.0)
    bands = {tuple(b["band"]): b for b in e["bands"]}
    assert [tuple(b["band"]) for b in e["bands"]] == list(zip(BANDS[1:], BANDS[:-1]))
    assert bands[(0.1, 1.0)]["traces"] == 2
    assert bands[(1e-3, 1e-2)]["traces"] == 1 and np.isclose(bands[(1e-3, 1e-2)]["worst"], 1e-3 / 0.05)
    assert bands[(1e-2, 0.1)]["worst"] is None
    # a lower floor takes the weak trace in and makes it the worst
    e2 = trace_errors(x, z, xr, zr, floor=1e-3)
    assert e2["kept"] == 3 and e2["worst_trace"] == ["Vz", 1] and np.isclose(e2["worst"], 2e-2)
    # the set error is the max over the two components
    assert np.isclose(e["all"], max(rel(x, xr), rel(z, zr)))


def test_trace_errors_silent_and_nonfinite():
    z0 = np.zeros((10, 3))
    e = trace_errors(z0 + 1.0, z0, z0, z0)
    assert e["all"] is None and e["worst"] is None and e["peak"] == 0.0
    xr = np.ones((10, 3))
    bad = xr.copy()
    bad[3, 1] = np.inf
    e = trace_errors(bad, xr, xr, xr)
    assert np.isnan(e["worst"]) and np.isnan(e["worst_vs_peak"]) and np.isnan(e["all"])
