"""Relative L2 errors of traces and snapshots against a reference run, with the amplitudes they refer to.

A relative error is only meaningful where the reference carries signal: `rel` returns None when the
reference is identically zero, and `trace_errors` reports the worst single trace only among traces whose
reference norm is at least `floor` times the largest trace norm of the set, together with that trace's norm
ratio, so that every number can be read against the amplitude it was measured on.
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


def trace_errors(x, z, xr, zr, floor: float = 1e-2) -> dict:
    """Errors of traces (x, z) [nsteps, nrec] against reference traces (xr, zr).

    all:          max over Vx, Vz of the relative L2 error of the whole trace set (None if silent);
    worst:        the largest per-receiver error max(rel(x_j, xr_j), rel(z_j, zr_j)) among receivers whose
                  reference norm is >= floor * the largest one;
    worst_norm:   that receiver's norm divided by the largest;
    kept, nrec:   receivers above the floor, and all receivers;
    peak:         the largest reference trace norm (absolute)."""
    n = trace_norms(xr, zr)
    out = {"all": None, "worst": None, "worst_norm": None, "kept": 0, "nrec": int(n.size), "floor": floor,
           "peak": float(n.max()) if n.size else 0.0}
    a, b = rel(x, xr), rel(z, zr)
    vals = [e for e in (a, b) if e is not None]
    out["all"] = max(vals) if vals else None
    if not n.size or n.max() <= 0:
        return out
    keep = np.nonzero(n >= floor * n.max())[0]
    out["kept"] = int(keep.size)
    worst, where = -1.0, -1
    for j in keep:
        e = max(v for v in (rel(x[:, j], xr[:, j]), rel(z[:, j], zr[:, j]), 0.0) if v is not None)
        if e > worst:
            worst, where = e, j
    if where >= 0:
        out["worst"] = worst
        out["worst_norm"] = float(n[where] / n.max())
        out["worst_receiver"] = int(where)
    return out


FIELDS = ("Vx", "Vz", "Txx", "Tzz", "Txz")


def snapshot_errors(f, fr) -> dict:
    """Per-field relative L2 error of fields f [5, Tri*6] against fr, each with the reference's largest
    absolute value."""
    return {FIELDS[c]: {"rel_l2": rel(f[c], fr[c]), "max_abs": float(np.max(np.abs(fr[c])))} for c in range(5)}
