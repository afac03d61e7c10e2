"""Relative L2 errors of traces and snapshots against a reference run, with the amplitudes they refer to.

A relative error is only meaningful where the reference carries signal: `rel` returns None when the
reference is identically zero.  `trace_errors` treats every component of every receiver (Vx or Vz of one
receiver over all steps) as one trace and reports, against the largest reference trace norm of the set (the
peak): the worst trace's own relative error among traces at or above `floor` of the peak, the worst trace's
error divided by the peak (every trace), and a table of the worst relative error per amplitude band, so that
every number can be read against the amplitude it was measured on.  FP32 rounding leaves an absolute error
of about the same size on strong and weak traces alike, so a trace's relative error grows as its amplitude
falls; the table shows that directly.
"""
from __future__ import annotations

import numpy as np


def rel(u, v) -> float | None:
    """||u - v|| / ||v||; None when v is zero everywhere (nothing to compare against), nan when either holds
    a non-finite value."""
    u, v = np.asarray(u, np.float64), np.asarray(v, np.float64)
    if not (np.all(np.isfinite(u)) and np.all(np.isfinite(v))):
        return float("nan")
    nv = np.linalg.norm(v)
    return float(np.linalg.norm(u - v) / nv) if nv > 0 else None


def trace_norms(x, z) -> np.ndarray:
    """Per-receiver norm of the (Vx, Vz) trace pair: x, z are [nsteps, nrec]."""
    x, z = np.asarray(x, np.float64), np.asarray(z, np.float64)
    return np.sqrt((x ** 2).sum(0) + (z ** 2).sum(0))


BANDS = (1.0, 0.1, 1e-2, 1e-3, 0.0)  # amplitude band edges (fraction of the peak trace norm)


def trace_errors(x, z, xr, zr, floor: float = 0.1) -> dict:
    """Errors of traces (x, z) [nsteps, nrec] against reference traces (xr, zr); a trace is one component
    (Vx or Vz) of one receiver.

    all:           max over Vx, Vz of the relative L2 error of the whole trace set (None if silent);
    worst:         the largest relative error ||t_j - r_j|| / ||r_j|| among traces with ||r_j|| >= floor * peak;
    worst_norm:    that trace's norm divided by the peak; worst_trace: [component, receiver];
    worst_vs_peak: the largest ||t_j - r_j|| / peak over every trace;
    bands:         per amplitude band (lo, hi] of ||r_j|| / peak: traces in it and their worst relative error;
    kept, ntraces: traces at or above the floor, and all traces;
    peak:          the largest reference trace norm (absolute)."""
    t = np.concatenate([np.asarray(x, np.float64), np.asarray(z, np.float64)], 1)
    r = np.concatenate([np.asarray(xr, np.float64), np.asarray(zr, np.float64)], 1)
    nrec = t.shape[1] // 2
    out = {"all": None, "worst": None, "worst_norm": None, "worst_trace": None, "worst_vs_peak": None,
           "bands": [], "kept": 0, "ntraces": int(r.shape[1]), "floor": floor, "peak": 0.0}
    a, b = rel(x, xr), rel(z, zr)
    vals = [e for e in (a, b) if e is not None]
    out["all"] = max(vals) if vals else None
    n = np.sqrt((r ** 2).sum(0))
    if not n.size or n.max() <= 0:
        return out
    peak = float(n.max())
    out["peak"] = peak
    if not np.all(np.isfinite(t)):
        out["worst"] = out["worst_vs_peak"] = float("nan")
        return out
    d = np.sqrt(((t - r) ** 2).sum(0))
    out["worst_vs_peak"] = float(d.max() / peak)
    frac = n / peak
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(n > 0, d / np.where(n > 0, n, 1.0), 0.0)
    keep = np.nonzero(frac >= floor)[0]
    out["kept"] = int(keep.size)
    j = int(keep[np.argmax(e[keep])])
    out["worst"] = float(e[j])
    out["worst_norm"] = float(frac[j])
    out["worst_trace"] = ["Vx" if j < nrec else "Vz", j % nrec]
    for hi, lo in zip(BANDS[:-1], BANDS[1:]):
        sel = (frac > lo) & (frac <= hi)
        out["bands"].append({"band": [lo, hi], "traces": int(sel.sum()),
                             "worst": float(e[sel].max()) if sel.any() else None})
    return out


FIELDS = ("Vx", "Vz", "Txx", "Tzz", "Txz")


def snapshot_errors(f, fr) -> dict:
    """Per-field relative L2 error of fields f [5, Tri*6] against fr, each with the reference's largest
    absolute value."""
    return {FIELDS[c]: {"rel_l2": rel(f[c], fr[c]), "max_abs": float(np.max(np.abs(fr[c])))} for c in range(5)}
