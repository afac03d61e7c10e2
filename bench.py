"""Throughput of the DG time-stepping hot path on one MI355X.

Workload (default): the 8 km x 4 km gradient model, 800x400 squares of 10 m with random diagonals and
jittered interior vertices (640,000 P2 triangles, 19.2 M DOF), 20 Hz explosive Ricker source, 401
surface receivers.  A step is one full leap-frog step (sponge, source, velocity half-step, stress
half-step, receiver pick) over the whole mesh; DOF-updates/s = Tri * 6 nodes * 5 fields * shots *
steps / seconds, with all inputs resident in HBM when the timed region starts.

Prints one JSON line.  `value` is the plain FP64 single-shot path (the fused velocity and stress half-step
kernels, whose HBM roofline `roofline` reports); beside it the FP64 default (the blocked kernel, several steps
per launch), the FP32 paths and their trace errors against FP64, the batched FP32 path (64 shots on the
~1.2 M-element slope model: the blocked kernel per shot, and the shot-innermost half-step kernels) and a
sample of the open-loop serving stream on the eight resident meshes (p50/p99 latency and aggregate
DOF-updates/s), each with its ratio to the plain FP64 rate.  `roofline` is for the dominant kernel
(algorithmic field bytes per launch over its HIP-event-measured duration); `cpu_baseline` is the reference's
own step loop (oracle/_ref, compiled from the reference sources) or, without it, the oracle, timed on the
host.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "discontinuous-galerkin-method_amd" / "python"))

import numpy as np  # noqa: E402

import dgx  # noqa: E402
from dgx import lib as L  # noqa: E402
from dgx.accuracy import rel, snapshot_errors, trace_errors, trace_norms  # noqa: E402

HBM_PEAK = 8.0e12


def progress(msg):
    """One line per finished stage on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU time of the reference step-loop sample (steps sized from a short timed run)")
    ap.add_argument("--shots", type=int, default=64, help="shots of the batched FP32 leg (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=320)
    ap.add_argument("--block-steps", type=int, default=0,
                    help="steps per launch of the blocked kernel in the batched leg (1-3; 0 = the library's default)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the seeded serving stream to run (0 = skip; 2000 = the whole stream)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-budget-gb", type=float, default=0.0,
                    help="HBM budget of the serving leg's wavefields and request buffers (0 = 80%% of free HBM)")
    ap.add_argument("--serve-group", type=int, default=1,
                    help="shots per slot group of the server (1: single-shot slots on the blocked kernel; 4, 8, 16, "
                         "32 or 64: shot-innermost groups)")
    ap.add_argument("--serve-compare-group", type=int, default=4,
                    help="also run the stream with this group size for comparison (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=1,
                    help="also time a plain device copy (achievable bandwidth, counter calibration)")
    ap.add_argument("--skip", default="",
                    help="comma-separated legs to leave out (large, small, accuracy, end_to_end), for profiling runs; "
                         "the headline fields are always measured")
    a = ap.parse_args()
    a.skip = {x for x in a.skip.split(",") if x}
    bad = a.skip - {"large", "small", "accuracy", "end_to_end"}
    if bad:
        ap.error(f"--skip: unknown legs {sorted(bad)}")
    return a


def run_path(setup, precision, warmup, steps, layout=L.DGX_LAYOUT_AUTO, bench_launches=100,
             step_kernel=L.DGX_STEP_AUTO):
    """Warm up, time `steps` whole steps (hipGraph replays), then time each kernel of the step alone over
    back-to-back launches between two HIP events on the solver's stream: {"full": ms} for the
    full-step kernel, {"velocity": ms, "stress": ms} for the half-step pair."""
    s = dgx.Solver(setup, precision, layout, step_kernel=step_kernel)
    info = s.layout_info() + (s.tile_info(),)
    s.step(warmup)
    s.synchronize()
    t = time.perf_counter()
    s.step(steps)
    s.synchronize()
    sec = time.perf_counter() - t
    x, z = s.traces(0)
    if step_kernel == L.DGX_STEP_FULL:
        kt = {"full": s.kernel_bench(2, bench_launches)}
    elif step_kernel == L.DGX_STEP_BLOCKED:
        kt = {"block": s.kernel_bench(3, bench_launches), "info": s.block_info()}
    else:
        kt = {"velocity": s.kernel_bench(0, bench_launches), "stress": s.kernel_bench(1, bench_launches)}
    s.close()
    n = warmup + steps
    return sec, kt, x[:n], z[:n], info


def end_to_end(setup, nsteps, snap_every):
    """The run as a caller of the C-ABI sees it, timed from the host arrays to the last snapshot: dgx_create
    (validation, element order, upload, step graph), `nsteps` steps, and all five fields of the shot every
    `snap_every` steps in the caller's element order.  Each snapshot is begun at its step and collected
    (dgx_snapshot_wait) after the next block of steps has been enqueued, as the C driver does."""
    T = setup.mesh.tri
    buf = np.zeros((5, T * 6))
    buf[:] = 1.0  # touch the pages before the timed region
    t = time.perf_counter()
    s = dgx.Solver(setup, L.DGX_FP64, step_kernel=L.DGX_STEP_HALF)
    s.synchronize()
    create = time.perf_counter() - t
    done, pending, nsnap = 0, False, 0
    while done < nsteps or pending:
        n = min(snap_every, nsteps - done)
        if n > 0:
            s.step(n)
        if pending:
            s.snapshot_wait(buf)
            pending = False
        if n == 0:
            break
        done += n
        if done % snap_every == 0:
            s.snapshot_begin(0, 31)
            pending, nsnap = True, nsnap + 1
    s.synchronize()
    total = time.perf_counter() - t
    t = time.perf_counter()
    s.fields(0)
    fetch = time.perf_counter() - t
    s.close()
    return {"create_s": create, "steps": nsteps, "snapshot_every": snap_every, "snapshots": nsnap,
            "get_fields_all5_ms": fetch * 1e3, "wall_s": total, "value": T * 30 * nsteps / total,
            "unit": "DOF-updates/s",
            "note": "host arrays in, five fields out at every snapshot (PCIe-inclusive); not the headline value"}


def cross(x, z, half=200.0, step=10.0):
    """Receivers on a horizontal and a vertical line through (x, z)."""
    t = np.arange(-half, half + step / 2, step)
    return np.concatenate([np.stack([x + t, np.full(t.size, z)], 1), np.stack([np.full(t.size, x), z + t], 1)])


def accuracy_leg(w, paths):
    """Accuracy on the workload's own run length (w.nsteps steps, whatever --steps is): each path in `paths`
    ({name: (precision, step_kernel)}) against plain FP64 on the same mesh, source and dt.  Receivers: the
    workload's surface spread and a cross through the source (the direct wave reaches only the latter within
    4000 steps at this dt).  Reported per set (a trace is one component of one receiver): the relative L2 error
    of the whole set, the worst trace among traces at >= 10% of the set's peak trace norm (with that trace's
    norm ratio), the worst trace error against the peak, the worst error per amplitude band, the peak; and per
    field the relative L2 error of the final fields with the field's largest amplitude."""
    src = w.sources[0]
    nsurf = len(w.receivers)
    rec = np.concatenate([w.receivers, cross(src.x, src.z)])
    st = dgx.build_setup(w.mesh, w.model, w.sources, rec, w.nsteps, sponge=w.sponge)

    def run(precision, step_kernel):
        s = dgx.Solver(st, precision, step_kernel=step_kernel)
        s.step(w.nsteps)
        x, z = s.traces(0)
        f = s.fields(0)
        s.close()
        return x, z, f
    x64, z64, f64 = run(L.DGX_FP64, L.DGX_STEP_HALF)
    out = {"steps": w.nsteps, "physical_time_s": w.nsteps * st.dt, "dt": st.dt,
           "receivers": {"surface": nsurf, "cross_through_source": len(rec) - nsurf},
           "surface_peak_over_cross_peak": float(trace_norms(x64[:, :nsurf], z64[:, :nsurf]).max() /
                                                 trace_norms(x64[:, nsurf:], z64[:, nsurf:]).max())}
    for name, (precision, sk) in paths.items():
        x, z, f = run(precision, sk)
        out[name] = {"surface_traces": trace_errors(x[:, :nsurf], z[:, :nsurf], x64[:, :nsurf], z64[:, :nsurf]),
                     "cross_traces": trace_errors(x[:, nsurf:], z[:, nsurf:], x64[:, nsurf:], z64[:, nsurf:]),
                     "snapshot": snapshot_errors(f, f64)}
    return out


def large_leg(a):
    """Plain FP64 on a ~2M-element jittered mesh (the serving set's largest, 480 MB of fields: twice the 256 MB
    Infinity Cache, so the fields stream from HBM every step): ms per step, DOF-updates/s and the fraction of
    HBM peak at 24 algorithmic bytes per DOF-update, for the whole step and per kernel."""
    mesh, model, sp = dgx.configs.serving_meshes(sizes=(2_000_000,))[0]
    W, H = mesh.point[:, 0].max(), mesh.point[:, 1].max()
    n = 320
    st = dgx.build_setup(mesh, model, [dgx.SourceSpec(W / 2, H - 300.0, L.DGX_SRC_EXPLOSIVE, 0, 20.0)],
                         np.stack([np.linspace(100.0, W - 100.0, 101), np.full(101, H - 5.0)], 1), n + 200, sponge=sp)
    sec, kt, _, _, _ = run_path(st, L.DGX_FP64, 64, n, step_kernel=L.DGX_STEP_HALF)
    secd, ktd, _, _, _ = run_path(st, L.DGX_FP64, 64, n, step_kernel=L.DGX_STEP_BLOCKED)  # the FP64 default
    T = mesh.tri
    rate = T * 30 * n / sec
    kv, ks = kt["velocity"], kt["stress"]
    return {"elements": T, "fields_MB": T * 30 * 8 / 1e6, "steps": n, "ms_per_step": sec * 1e3 / n, "value": rate,
            "path": "half-step pair", "step_frac": rate * 24 / HBM_PEAK, "velocity_ms": kv, "stress_ms": ks,
            "velocity_frac": T * 6 * 8 * 7 / (kv * 1e-3) / HBM_PEAK, "stress_frac": T * 6 * 8 * 8 / (ks * 1e-3) / HBM_PEAK,
            "default": {"path": "blocked kernel (the FP64 default), 256-thread tiles, 2 steps per launch",
                        "ms_per_step": secd * 1e3 / n, "value": T * 30 * n / secd, "launch_ms": ktd["block"],
                        "tiling": ktd["info"], "x_vs_half_step": sec / secd}}


def small_leg(a):
    """FP64 on the 2 km homogeneous case (20,000 elements, the reference's own problem size): the default
    step kernel (the blocked kernel on 256-thread tiles, 2 steps per launch: both field copies stay in the
    Infinity Cache) against the half-step pair, ms per step over 1440 steps after 96 of warm-up."""
    w = dgx.configs.homogeneous_2km(1600)
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, 1600, sponge=w.sponge)
    out = {"workload": w.name, "elements": w.mesh.tri}
    for name, sk in (("default", L.DGX_STEP_AUTO), ("half_step", L.DGX_STEP_HALF)):
        s = dgx.Solver(st, L.DGX_FP64, step_kernel=sk)
        s.step(96)
        s.synchronize()
        t = time.perf_counter()
        s.step(1440)
        s.synchronize()
        sec = time.perf_counter() - t
        out[name] = {"ms_per_step": sec * 1e3 / 1440, "value": w.mesh.tri * 30 * 1440 / sec}
        if sk == L.DGX_STEP_AUTO:
            out[name]["tiling"] = s.block_info()
        s.close()
    out["default_vs_half_step"] = out["default"]["value"] / out["half_step"]["value"]
    return out


def fp64_trace_check(w, st, traces_of, picks, n):
    """Worst FP32 trace error of shots `picks` (traces_of(k) -> (x, z)) against FP64 single-shot runs of the
    plain path on the same mesh, sources and dt: max over the shots of the set error, of the worst trace at
    >= 10 % of the set's peak trace norm, and of the worst trace error relative to that peak."""
    worst = 0.0
    for k in picks:
        one = dgx.configs.Workload(w.name, w.mesh, w.model, w.sponge, [dgx.SourceSpec(
            w.sources[k].x, w.sources[k].z, w.sources[k].kind, 0, w.sources[k].f0)], w.receivers, w.nsteps, 0)
        s1 = dgx.build_setup(one.mesh, one.model, one.sources, one.receivers, one.nsteps, sponge=one.sponge,
                             dt=st.dt)
        r64 = dgx.Solver(s1, L.DGX_FP64, L.DGX_LAYOUT_ELEMENT)
        r64.step(n)
        x64, z64 = r64.traces(0)
        r64.close()
        x32, z32 = traces_of(k)
        e = trace_errors(x32[:n], z32[:n], x64[:n], z64[:n])
        worst = max(worst, e["all"] or 0.0, e["worst"] or 0.0, e["worst_vs_peak"] or 0.0)
    return worst


def batched_leg(a, rate64, rate64d):
    """Batched FP32 on the slope model (water over three dipping layers, ~1.2 M elements), S shots.  Primary:
    the blocked kernel (element layout, block_steps steps per launch, tiles with recomputed rings); beside it
    the shot-innermost half-step kernels with the autotuned shot tile.  FP32 traces of 4 shots compared with
    FP64 runs of the plain path."""
    w = dgx.configs.slope_6x3km(a.batched_steps + 64, nshots=a.shots)  # room for the launch timing after the run
    S, nb = a.shots, a.batched_steps
    st = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    T = w.mesh.tri
    picks = [0, S // 3, (2 * S) // 3, S - 1]
    # blocked kernel
    s = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_ELEMENT, step_kernel=L.DGX_STEP_BLOCKED, block_steps=a.block_steps)
    binfo = s.block_info()
    s.step(32)
    s.synchronize()
    t = time.perf_counter()
    s.step(nb)
    s.synchronize()
    secb = time.perf_counter() - t
    worst = fp64_trace_check(w, st, s.traces, picks, 32 + nb)
    kms = s.kernel_bench(3, 10)
    s.close()
    rateb = T * 30 * S * nb / secb
    K = binfo["steps_per_launch"]
    # algorithmic bytes of one launch: each field read and written once per element and shot for K steps
    alg = T * S * 6 * 4 * 10
    out = {"workload": w.name, "elements": T, "value": rateb, "unit": "DOF-updates/s", "dtype": "f32",
           "shots": S, "steps": nb, "ms_per_step": secb * 1e3 / nb, "x_vs_plain_fp64": rateb / rate64,
           "x_vs_fp64_default": rateb / rate64d,
           "trace_rel_l2_vs_fp64_worst_of_4_shots": worst,
           "trace_error_note": "max over 4 shots of the set error, of the worst trace at >= 10% of the peak trace norm and of the worst trace error against that peak",
           "path": f"blocked kernel, element layout, {K} steps per launch, graphs of 16 launches",
           "tiling": binfo, "launch_ms": kms, "launch_algorithmic_GBps": alg / (kms * 1e-3) / 1e9,
           "bytes_per_dof_update_algorithmic": 10 * 4 / (30 / 6) / K}
    progress(f"batched FP32 (blocked): {rateb:.3e} DOF-updates/s")
    # shot-innermost half-step kernels on the same shots, for comparison
    secs, kts, _, _, info = run_path(st, L.DGX_FP32, 20, nb, L.DGX_LAYOUT_SHOT_INNER, 20)
    tinfo = info[3]
    kvb, ksb = kts["velocity"], kts["stress"]
    bytes_b = T * 6 * 4 * S
    out["shot_innermost"] = {
        "value": T * 30 * S * nb / secs, "ms_per_step": secs * 1e3 / nb, "shot_tile": tinfo[0],
        "wavefronts_per_workgroup": tinfo[1], "elements_per_workgroup": tinfo[1] * 64 // tinfo[0],
        "autotune_ms_per_step": {f"tile {t} x {w} wavefronts": v for (t, w), v in tinfo[2].items()},
        "velocity_ms": kvb, "stress_ms": ksb,
        "velocity_GBps": bytes_b * 7 / (kvb * 1e-3) / 1e9, "stress_GBps": bytes_b * 8 / (ksb * 1e-3) / 1e9}
    return out


def serving_leg(a, rate64, group, rate64d):
    """Open-loop serving on the eight resident meshes (20 k .. 2 M elements, random layered models):
    the first `serve_requests` requests of the seeded Poisson stream (mean rate per second), FP32, with
    `group` shots per slot group (1: single-shot slots stepped by the blocked kernel)."""
    import dgx.serve as SV
    ms = dgx.configs.serving_meshes()
    arrivals, specs = dgx.configs.serving_requests(ms, n=max(a.serve_requests, 1), rate=a.serve_rate)
    arrivals, specs = arrivals[:a.serve_requests], specs[:a.serve_requests]
    # blocks of 96 steps: a multiple of 6, so that single-shot slots run three steps per launch
    sv = SV.Server(L.DGX_FP32, int(a.serve_budget_gb * 2**30), 96, group_shots=group)
    rms = [sv.add_mesh(m, md, sp) for m, md, sp in ms]
    reqs = [sv.make_request(rms[r["mesh"]], r["xz"], [r["f0"]] * len(r["xz"]), r["nsteps"], r["rec"])
            for r in specs]
    lat, wall, st, _ = SV.run_open_loop(sv, reqs, arrivals)
    sv.close()
    dof = st["dof_updates"]
    return {"requests": len(reqs), "arrival_rate_per_s": a.serve_rate, "meshes": [m.tri for m, _, _ in ms],
            "group_shots": group, "hbm_budget_gb": a.serve_budget_gb or "80% of free",
            "path": ("single-shot slots, blocked kernel (3 steps per launch, blocks of 96 steps)" if group == 1 else
                     f"{group}-shot slot groups, shot-innermost half-step kernels"),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "wall_s": wall, "dof_updates": dof,
            "value": dof / wall, "unit": "DOF-updates/s", "dtype": "f32",
            "x_vs_plain_fp64": dof / wall / rate64, "x_vs_fp64_default": dof / wall / rate64d,
            "busy_s": st["busy_seconds"],
            "gpu_s": st["gpu_seconds"], "slot_occupancy": dof / max(st["lane_updates"], 1), "moved": st["moved"],
            "lane_rate": st["lane_updates"] / max(st["gpu_seconds"], 1e-9),
            "sample": f"first {len(reqs)} requests of the seeded stream (2000 in the full stream)"}


def cpu_baseline(w, setup, budget_s):
    """Time the reference's step functions (oracle/_ref, -O2 -fopenmp, all host threads, OMP_PROC_BIND=close,
    OMP_PLACES=cores) on a bounded number of steps of the same mesh: three runs, the median reported.  The
    README's flag-less build (-O0) is timed once beside it for context."""
    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when libgomp loads, with the libraries below
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, str(ROOT / "oracle"))
    import ctypes as C

    import oracle_py as O
    lay = O.setup_reference_layout(w.mesh.point, w.mesh.triangle, w.mesh.neighbor, w.model.lam, w.model.miu,
                                   w.model.den, O.read_operators(L.OPERATOR_DIR),
                                   O.SpongeParams(w.sponge.bound_x, w.sponge.bound_z, w.sponge.model_w,
                                                  w.sponge.model_h, w.sponge.beta, w.sponge.pi))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or O.oracle().orc_num_threads()
    src = setup.sources[0]
    fun = np.array(list(src.fun))

    def ref_timer(r):
        ibe = lay.ibe.copy()

        def timed(n):
            return r.ref_time_steps(C.c_long(lay.tri), C.c_double(setup.dt), O.dp(lay.A), O.dp(lay.B), O.dp(lay.Dx),
                                    O.dp(lay.Dz), O.dp(lay.Q), O.dp(lay.S), O.lp(ibe), O.lp(lay.node_neighbor),
                                    O.lp(lay.neighbor), O.lp(lay.edge_flag), O.dp(lay.absorb_fun), C.c_long(src.elem),
                                    O.dp(fun), C.c_long(setup.receivers[0].elem), C.c_long(1), C.c_long(n),
                                    C.c_double(setup.t0[0]))
        return timed
    if O.reference_available():
        timed, kind = ref_timer(O.reference()), "reference"
    else:
        def timed(n):
            run = O.Run(lay, setup.dt, [(src.elem, src.kind, fun)], setup.wavelets[0], [])
            t = time.perf_counter()
            run.run(n)
            return time.perf_counter() - t
        kind = "port"
    timed(2)  # warm-up: first touches of the arrays, thread start
    t1 = timed(6)
    n = max(3, min(3000, int(budget_s / 4 / max(t1 / 6, 1e-6))))
    runs = sorted(timed(n) for _ in range(3))
    sec = runs[1]
    rate = lay.tri * 30 * n / sec
    o0 = None
    if kind == "reference" and O.reference_available("O0"):
        t0 = ref_timer(O.reference("O0"))
        t0(1)
        n0 = max(2, int(budget_s / 4 / max(t0(2) / 2, 1e-6)))
        o0 = {"value": lay.tri * 30 * n0 / t0(n0), "steps": n0, "build": "-O0 -fopenmp (the README's flag-less build)"}
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": rate, "unit": "DOF-updates/s", "cores": threads, "kind": kind,
            "host": f"{model}, {os.cpu_count()} logical CPUs visible",
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "runs_s": runs, "statistic": "median of 3 runs",
            "sample": f"{n} full steps of the same 640k-element mesh per run (reference step loop, -O2 -fopenmp, "
                      f"{threads} threads), median {sec:.1f} s",
            "O0": o0}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a design note "
                                   "(shots and requests shard with no exchange), not a measured path"}))
        sys.exit(2)
    w = dgx.configs.gradient_8x4km(max(a.warmup + a.steps + 5, 4000))  # 4000: the end-to-end run
    setup = dgx.build_setup(w.mesh, w.model, w.sources, w.receivers, w.nsteps, sponge=w.sponge)
    T = w.mesh.tri
    dof = T * 30

    progress(f"{w.name}: {T} elements set up")
    sec64, kt64, x64, z64, _ = run_path(setup, L.DGX_FP64, a.warmup, a.steps, step_kernel=L.DGX_STEP_HALF)
    progress(f"plain FP64: {sec64 * 1e3 / a.steps:.4f} ms/step")
    sec32, kt32, x32, z32, _ = run_path(setup, L.DGX_FP32, a.warmup, a.steps, step_kernel=L.DGX_STEP_HALF)
    progress(f"plain FP32: {sec32 * 1e3 / a.steps:.4f} ms/step")
    secb1, ktb1, _, _, _ = run_path(setup, L.DGX_FP32, a.warmup, a.steps, step_kernel=L.DGX_STEP_BLOCKED)
    progress(f"blocked FP32: {secb1 * 1e3 / a.steps:.4f} ms/step")
    secf, ktf, xf, zf, _ = run_path(setup, L.DGX_FP64, a.warmup, a.steps, step_kernel=L.DGX_STEP_FULL)
    progress(f"full-step FP64: {secf * 1e3 / a.steps:.4f} ms/step")
    secd, ktd, _, _, _ = run_path(setup, L.DGX_FP64, a.warmup, a.steps, step_kernel=L.DGX_STEP_BLOCKED)  # the FP64 default
    progress(f"default (blocked) FP64: {secd * 1e3 / a.steps:.4f} ms/step")
    secr, ktr, _, _, _ = run_path(setup, L.DGX_FP64_REFERENCE, a.warmup, a.steps, step_kernel=L.DGX_STEP_HALF)
    progress(f"reference-arithmetic FP64: {secr * 1e3 / a.steps:.4f} ms/step")
    rate64 = dof * a.steps / sec64
    rate32 = dof * a.steps / sec32
    rate64f = dof * a.steps / secf

    # launches of the timed region: dgx_step replays a 32-step graph per full block, the rest is eager
    graph_blocks, eager = divmod(a.steps, 32)
    path = (f"plain FP64, fused velocity+stress kernels: {graph_blocks} hipGraph replays of 32 steps"
            f" + {eager} eagerly launched steps" if graph_blocks else
            f"plain FP64, fused velocity+stress kernels: {eager} eagerly launched steps (fewer than one 32-step graph)")

    # dominant kernel of the FP64 step: algorithmic bytes = fields read + written per launch
    bytes_vel = T * 6 * 8 * 7   # read Txx,Tzz,Txz,Vx,Vz; write Vx,Vz
    bytes_str = T * 6 * 8 * 8   # read Vx,Vz,Txx,Tzz,Txz; write Txx,Tzz,Txz
    kv64, ks64 = kt64["velocity"], kt64["stress"]
    if ks64 >= kv64:
        kname, kbytes, kms = "k_stress<double,fused>", bytes_str, ks64
    else:
        kname, kbytes, kms = "k_velocity<double,fused>", bytes_vel, kv64
    achieved = kbytes / (kms * 1e-3)
    traffic = None
    tfile = ROOT / "profiles" / "traffic_fp64_gradient_8x4km.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(kname)
        except Exception:
            traffic = None

    out = {
        "metric": "DOF-updates/s: batched FP32 vs plain FP64 (<=1e-4 trace err); p50/p99 latency",
        "value": rate64,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec64 * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh and gradient model generated in-process)",
        "config": {"workload": w.name, "elements": T, "dof": dof, "shots": 1, "receivers": int(len(w.receivers)),
                   "path": path},
        "kernel_timing": "each kernel timed alone over back-to-back launches between HIP events",
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "achieved_GBps": achieved / 1e9, "kernel_ms": kms, "algorithmic_bytes_per_launch": kbytes,
                     "velocity_ms": kv64, "stress_ms": ks64,
                     "step_frac": T * 30 * 24 / (sec64 / a.steps) / HBM_PEAK},
        "fp32": {"value": rate32, "ms_per_step": sec32 * 1e3 / a.steps,
                 "velocity_ms": kt32["velocity"], "stress_ms": kt32["stress"], "speedup_vs_fp64": rate32 / rate64},
        "blocked_fp32": {"value": dof * a.steps / secb1, "ms_per_step": secb1 * 1e3 / a.steps,
                         "launch_ms": ktb1["block"], "tiling": ktb1["info"],
                         "x_vs_plain_fp64": dof * a.steps / secb1 / rate64,
                         "x_vs_fp64_default": secd / secb1},
        "full_step_fp64": {"value": rate64f, "ms_per_step": secf * 1e3 / a.steps, "kernel_ms": ktf["full"],
                           "note": "one kernel per step over tiles + recomputed ring (opt-in DGX_STEP_FULL)"},
        "default_fp64": {"value": dof * a.steps / secd, "ms_per_step": secd * 1e3 / a.steps, "launch_ms": ktd["block"],
                         "tiling": ktd["info"], "x_vs_half_step": sec64 / secd,
                         "path": "the FP64 default: blocked kernel, 256-thread tiles, 2 steps per launch"},
        "reference_arithmetic_fp64": {
            "value": dof * a.steps / secr, "ms_per_step": secr * 1e3 / a.steps, "velocity_ms": ktr["velocity"],
            "stress_ms": ktr["stress"], "time_vs_fp64_default": secr / secd,
            "path": "DGX_FP64_REFERENCE: the reference's operations in its order, no fused multiply-adds; two "
                    "kernels per step (sponge and sources applied as values load, receivers recorded by the stress "
                    "kernel)",
            "parity": "fields and traces equal the oracle's bit for bit (all 4000 steps of this workload: "
                      "tests/test_gpu_configs.py::test_gradient_640k_all_4000_steps)"},
    }
    if "large" not in a.skip:
        out["plain_fp64_large"] = large_leg(a)
        progress(f"plain FP64, 2M elements: {out['plain_fp64_large']['value']:.3e} DOF-updates/s, "
                 f"{out['plain_fp64_large']['step_frac']:.3f} of HBM peak")
    if "small" not in a.skip:
        out["plain_fp64_small"] = small_leg(a)
        progress(f"plain FP64, 20k elements: default {out['plain_fp64_small']['default']['ms_per_step']:.4f} "
                 f"ms/step, half-step {out['plain_fp64_small']['half_step']['ms_per_step']:.4f}")
    if "accuracy" not in a.skip:
        out["accuracy"] = accuracy_leg(dataclasses.replace(w, nsteps=4000), {
            "fp32": (L.DGX_FP32, L.DGX_STEP_HALF), "blocked_fp32": (L.DGX_FP32, L.DGX_STEP_BLOCKED),
            "full_step_fp64": (L.DGX_FP64, L.DGX_STEP_FULL)})
        ac = out["accuracy"]
        progress(f"accuracy over {ac['steps']} steps: fp32 cross traces {ac['fp32']['cross_traces']['all']:.2e}, "
                 f"worst {ac['fp32']['cross_traces']['worst']:.2e}")
    if "end_to_end" not in a.skip:
        out["end_to_end_fp64"] = end_to_end(setup, 4000, 500)
        progress(f"end to end: {out['end_to_end_fp64']['value']:.3e} DOF-updates/s")
    if a.shots > 0:
        out["batched_fp32"] = batched_leg(a, rate64, dof * a.steps / secd)
        progress(f"batched FP32 (shot-innermost): {out['batched_fp32']['shot_innermost']['value']:.3e} DOF-updates/s")
    if a.serve_requests > 0:
        out["serving"] = serving_leg(a, rate64, a.serve_group, dof * a.steps / secd)
        progress(f"serving ({a.serve_group}-shot groups): {out['serving']['value']:.3e} DOF-updates/s")
        if a.serve_compare_group and a.serve_compare_group != a.serve_group:
            out["serving"]["compare"] = serving_leg(a, rate64, a.serve_compare_group, dof * a.steps / secd)
            progress(f"serving ({a.serve_compare_group}-shot groups): "
                     f"{out['serving']['compare']['value']:.3e} DOF-updates/s")
    if a.copy_probe:
        import ctypes as C
        nbytes, reps = 1 << 30, 20
        probes = {}
        for variant, name in ((0, "copy_grid_stride"), (1, "copy_nontemporal_32KiB_per_wg"), (2, "read_only")):
            ms = C.c_double()
            L.check(L.load().dgx_copy_probe(nbytes, reps, variant, C.byref(ms)), "dgx_copy_probe")
            probes[name] = (1 if variant == 2 else 2) * nbytes / (ms.value * 1e-3) / 1e9
        out["copy_probe"] = {"bytes_per_launch": nbytes, "GBps": probes,
                             "best_copy_GBps": max(probes["copy_grid_stride"], probes["copy_nontemporal_32KiB_per_wg"])}
    out["roofline"]["achieved"] = achieved / 1e9
    out["roofline"]["peak"] = HBM_PEAK / 1e9
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(w, setup, a.cpu_seconds)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
