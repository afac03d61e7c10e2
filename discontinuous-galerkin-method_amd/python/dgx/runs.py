"""Full runs of the measured workloads, one JSON line each:

    python -m dgx.runs gradient  [--steps 4000]             single shot 640k elements: FP64 vs FP32
    python -m dgx.runs slope     [--steps 5000 --shots 64]  batched FP32 (default path; --layout inner: the
                                                            shot-innermost kernels), FP64 reference on 4 shots
    python -m dgx.runs serving   [--requests 2000 --rate 20] open-loop request stream on 8 meshes
    python -m dgx.runs carbonate [--shots 256 --steps 8000] 256 shots through the request loop, FP64 on 4

The serving runs take --group (1: single-shot slots on the blocked kernel, the default; 4, 8, 16, 32 or 64 shots per
slot group).

Progress lines go to stderr while a run is going."""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs
from .accuracy import snapshot_errors, trace_errors
from . import lib as L
from .problem import SourceSpec, Solver, build_setup


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_run(setup, precision, layout=L.DGX_LAYOUT_AUTO, chunk=500):
    s = Solver(setup, precision, layout)
    s.synchronize()
    t = time.perf_counter()
    done = 0
    while done < setup.nsteps:
        k = min(chunk, setup.nsteps - done)
        s.step(k)
        s.synchronize()
        done += k
        log(f"  {done}/{setup.nsteps} steps, {time.perf_counter() - t:.1f} s")
    sec = time.perf_counter() - t
    return s, sec


def single_shot_reference(w, k, dt, nsteps):
    """FP64 plain single-shot run of shot k of a workload (the path parity-tested against the CPU oracle)."""
    src = w.sources[k]
    one = configs.Workload(w.name, w.mesh, w.model, w.sponge, [SourceSpec(src.x, src.z, src.kind, 0, src.f0)],
                           w.receivers, nsteps, 0)
    st = build_setup(one.mesh, one.model, one.sources, one.receivers, nsteps, sponge=one.sponge, dt=dt)
    s, sec = timed_run(st, L.DGX_FP64, L.DGX_LAYOUT_ELEMENT)
    out = s.traces(0), s.fields(0)
    s.close()
    return out, sec


def run_gradient(a):
    w = configs.gradient_8x4km(a.steps)
    st = build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    s64, sec64 = timed_run(st, L.DGX_FP64)
    x64, z64 = s64.traces(0)
    f64 = s64.fields(0)
    s64.close()
    s32, sec32 = timed_run(st, L.DGX_FP32)
    x32, z32 = s32.traces(0)
    f32 = s32.fields(0)
    s32.close()
    dof = w.mesh.tri * 30 * a.steps
    return {"workload": w.name, "steps": a.steps, "dt": st.dt, "fp64_dof_updates_per_s": dof / sec64,
            "fp32_dof_updates_per_s": dof / sec32, "fp32_traces": trace_errors(x32, z32, x64, z64),
            "fp32_snapshot": snapshot_errors(f32, f64)}


def run_slope(a):
    w = configs.slope_6x3km(a.steps, nshots=a.shots)
    log(f"slope mesh {w.mesh.tri} elements")
    st = build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    layout = L.DGX_LAYOUT_SHOT_INNER if a.layout == "inner" else L.DGX_LAYOUT_AUTO
    s32, sec32 = timed_run(st, L.DGX_FP32, layout)
    path = (f"shot-innermost half-step kernels, tile {s32.tile_info()[0]}" if a.layout == "inner" else
            f"blocked kernel, {s32.block_info()['steps_per_launch']} steps per launch")
    picks = [0, a.shots // 3, (2 * a.shots) // 3, a.shots - 1]
    checks, sec64 = [], 0.0
    for k in picks:
        ((x64, z64), f64), t = single_shot_reference(w, k, st.dt, a.steps)
        sec64 += t
        x32, z32 = s32.traces(k)
        f32 = s32.fields(k)
        checks.append({"shot": k, "traces": trace_errors(x32, z32, x64, z64),
                       "snapshot": snapshot_errors(f32, f64),
                       "fp64_finite": bool(np.all(np.isfinite(f64))), "fp32_finite": bool(np.all(np.isfinite(f32)))})
        log(f"  shot {k}: {checks[-1]}")
    s32.close()
    dof = w.mesh.tri * 30 * a.shots * a.steps
    return {"workload": w.name, "elements": w.mesh.tri, "shots": a.shots, "steps": a.steps, "dt": st.dt,
            "batched_fp32_dof_updates_per_s": dof / sec32, "path": path,
            "plain_fp64_dof_updates_per_s": w.mesh.tri * 30 * a.steps * len(picks) / sec64,
            "fp64_reference_checks": checks}


def serve_stream(meshes, specs, arrivals, precision=L.DGX_FP32, steps_per_block=96, group=16):
    from . import serve as SV
    sv = SV.Server(precision, 0, steps_per_block, group_shots=group)
    rms = [sv.add_mesh(m, md, sp) for m, md, sp in meshes]
    reqs = [sv.make_request(rms[r["mesh"]], r["xz"], [r["f0"]] * len(r["xz"]), r["nsteps"], r["rec"]) for r in specs]
    log(f"serving {len(reqs)} requests on meshes {[m.tri for m, _, _ in meshes]}")
    lat, wall, st, res = SV.run_open_loop(sv, reqs, arrivals, keep_results=True)
    sv.close()
    return lat, wall, st, res, reqs


def run_serving(a):
    ms = configs.serving_meshes()
    arrivals, specs = configs.serving_requests(ms, n=a.requests, rate=a.rate)
    lat, wall, st, _, _ = serve_stream(ms, specs, arrivals, group=a.group)
    return {"requests": a.requests, "rate_per_s": a.rate, "meshes": [m.tri for m, _, _ in ms], "group_shots": a.group,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "wall_s": wall, "dof_updates": st["dof_updates"], "dof_updates_per_s": st["dof_updates"] / wall,
            "offered_dof_updates_per_s": float(sum(len(r["xz"]) * r["nsteps"] * ms[r["mesh"]][0].tri * 30
                                                   for r in specs) / arrivals[-1])}


def run_carbonate(a):
    w = configs.carbonate_4x3km(a.steps, nshots=a.shots)
    T = w.mesh.tri
    log(f"carbonate mesh {T} elements")
    st = build_setup(w.mesh, w.model, w.sources[:1], w.receivers, 1, sponge=w.sponge)
    per_req = 16  # shots per request
    specs, arrivals = [], []
    for g in range(0, a.shots, per_req):
        xz = np.array([[s.x, s.z] for s in w.sources[g:g + per_req]])
        specs.append(dict(mesh=0, xz=xz, f0=30.0, nsteps=a.steps, rec=w.receivers))
        arrivals.append(0.0)
    lat, wall, stats, res, reqs = serve_stream([(w.mesh, w.model, w.sponge)], specs, np.array(arrivals), group=a.group)
    picks = [0, a.shots // 3, (2 * a.shots) // 3, a.shots - 1]
    errs = []
    for k in picks:
        ((x64, z64), _), _ = single_shot_reference(w, k, st.dt, a.steps)
        x32, z32 = res[k // per_req][0][k % per_req], res[k // per_req][1][k % per_req]
        errs.append({"shot": k, "traces": trace_errors(x32, z32, x64, z64)})
    return {"workload": w.name, "elements": T, "shots": a.shots, "steps": a.steps, "dt": st.dt, "group_shots": a.group,
            "served_fp32_dof_updates_per_s": stats["dof_updates"] / wall, "wall_s": wall,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "fp32_vs_fp64_trace_rel_l2_per_checked_shot": errs}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["gradient", "slope", "serving", "carbonate"])
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--shots", type=int, default=0)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--group", type=int, default=0,
                    help="shots per slot group of the server, 1 (single-shot slots on the blocked kernel, the default), "
                         "4, 8, 16, 32 or 64")
    ap.add_argument("--layout", choices=["auto", "inner"], default="auto",
                    help="slope: the default path (blocked kernel) or the shot-innermost half-step kernels")
    a = ap.parse_args(argv)
    defaults = {"gradient": (4000, 1), "slope": (5000, 64), "serving": (0, 0), "carbonate": (8000, 256)}
    a.steps = a.steps or defaults[a.which][0]
    a.shots = a.shots or defaults[a.which][1]
    a.group = a.group or 1
    fn = {"gradient": run_gradient, "slope": run_slope, "serving": run_serving, "carbonate": run_carbonate}[a.which]
    print(json.dumps(fn(a)), flush=True)


if __name__ == "__main__":
    main()
