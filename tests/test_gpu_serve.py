"""Request loop (include/dgx_serve.h): continuous batching must not change any result.

* every request's traces are bit-identical alone and in a staggered mix with other requests on the same
  and on other meshes;
* they equal the standalone shot-innermost FP32 solver's traces for the same shots to FP32 rounding;
* they are within 1e-4 relative L2 of the FP64 oracle (FP32 serving), and within 1e-12 of it when the
  server runs in FP64;
* a tight memory budget queues requests instead of failing, and all of them complete;
* a request's shots may spread over several slot groups of its mesh (also when it has more shots than one
  group holds), and may move between them when groups are compacted, without changing its results.

The first two hold for every slot-group size (4, 8, 16, 32 and 64 shots per group, which is also the serving
kernel's shot tile) and for single-shot slots (group 1), which the blocked kernel steps three steps per launch
in FP32 (blocks of 48 steps), two in FP64 (compared with the standalone solver's blocked kernel).  The first is also checked in FP64, for single-shot
slots (256-thread tiles) and 16-shot groups.
"""
import numpy as np
import pytest

import dgx
import dgx.serve as S
from dgx import lib as L
from helpers import OracleCase, rel_l2

pytestmark = pytest.mark.gpu


def meshes():
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    m0 = dgx.structured(40, 40, 50.0, diagonal="random", seed=3, sponge_box=sp.box(),
                        attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    m1 = dgx.structured(25, 25, 80.0, diagonal="same", jitter=10.0, seed=9, sponge_box=sp.box(), attribute=3)
    return sp, [(m0, dgx.Model.by_attribute(m0.attribute)), (m1, dgx.Model.by_attribute(m1.attribute))]


REQS = [  # (mesh, shots, steps, f0)
    (0, 3, 230, 15.0), (1, 5, 170, 20.0), (0, 16, 120, 12.0), (0, 2, 260, 25.0), (1, 1, 90, 30.0),
]


def build(server, rms, spec, k):
    mi, ns, nsteps, f0 = spec
    rng = np.random.default_rng(100 + k)
    xz = np.stack([rng.uniform(300, 1700, ns), rng.uniform(600, 1400, ns)], 1)
    rec = np.stack([np.linspace(200, 1800, 17), np.full(17, 1450.0 + 10 * k)], 1)
    return server.make_request(rms[mi], xz, [f0] * ns, nsteps, rec), xz, rec


def new_server(precision=L.DGX_FP32, budget=0, K=50, group=16):
    if group == 1:
        K = K // 12 * 12  # single-shot slots: an even number of 2- or 3-step launches per block
    sp, ms = meshes()
    sv = S.Server(precision, budget, K, group_shots=group)
    rms = [sv.add_mesh(m, md, sp) for m, md in ms]
    return sv, rms, sp, ms


def run_alone(k, group=16, precision=L.DGX_FP32):
    sv, rms, _, _ = new_server(precision=precision, group=group)
    req, xz, rec = build(sv, rms, REQS[k], k)
    rid = sv.submit(req)
    sv.drain()
    out = sv.result(rid, req)
    sv.close()
    return out, xz, rec


@pytest.mark.parametrize("group,precision", [(g, L.DGX_FP32) for g in (1, 4, 8, 16, 32, 64)]
                         + [(1, L.DGX_FP64), (16, L.DGX_FP64)])
def test_results_identical_alone_and_in_the_mix(group, precision):
    print(f"case serve mix, group {group}, precision {precision}", flush=True)
    sv, rms, _, _ = new_server(precision=precision, group=group)
    built = [build(sv, rms, spec, k) for k, spec in enumerate(REQS)]
    ids = [sv.submit(built[0][0]), sv.submit(built[1][0])]
    sv.advance(1)
    ids.append(sv.submit(built[2][0]))
    sv.advance(2)
    ids += [sv.submit(built[3][0]), sv.submit(built[4][0])]
    sv.drain()
    st = sv.stats()
    assert st["done"] == len(REQS) and st["running"] == 0
    mixed = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
    sv.close()
    for k in range(len(REQS)):
        (ax, az), _, _ = run_alone(k, group, precision)
        assert np.array_equal(mixed[k][0], ax) and np.array_equal(mixed[k][1], az), k


def test_single_shot_slots_two_and_three_steps_per_launch_agree():
    """FP32 single-shot slots run three steps per launch when the block of steps is a multiple of 6 (48) and two
    otherwise (52); every element's update is the same code on the same inputs either way (the rings recompute
    what their owning tile computes), so the traces agree bit for bit."""
    print("case serve single-shot slots, 2 and 3 steps per launch", flush=True)
    out = []
    for K in (48, 52):
        sp, ms = meshes()
        sv = S.Server(L.DGX_FP32, 0, K, group_shots=1)
        rms = [sv.add_mesh(m, md, sp) for m, md in ms]
        reqs = [build(sv, rms, REQS[k], k) for k in (0, 1)]
        ids = [sv.submit(r[0]) for r in reqs]
        sv.drain()
        out.append([sv.result(rid, r[0]) for rid, r in zip(ids, reqs)])
        sv.close()
    for (ax, az), (bx, bz) in zip(*out):
        assert np.array_equal(ax, bx) and np.array_equal(az, bz)


def test_requests_spread_over_groups():
    """Three requests of 10, 10 and 12 shots on one mesh with 16-shot groups fill exactly two groups (the
    second and third request take the free slots of the first group before a new one is opened), and each
    request's traces equal its traces when it runs alone in a group of its own."""
    print("case serve spread over groups, group 16", flush=True)
    specs = [(0, 10, 140, 15.0), (0, 10, 200, 20.0), (0, 12, 170, 12.0)]
    sv, rms, _, _ = new_server(group=16)
    built = [build(sv, rms, spec, 10 + k) for k, spec in enumerate(specs)]
    ids = [sv.submit(b[0]) for b in built]
    sv.advance(1)
    st = sv.stats()
    assert st["running"] == 3 and st["batches"] == 2, st
    sv.drain()
    spread = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
    sv.close()
    for k, spec in enumerate(specs):
        sv, rms, _, _ = new_server(group=16)
        req, _, _ = build(sv, rms, spec, 10 + k)
        rid = sv.submit(req)
        sv.drain()
        ax, az = sv.result(rid, req)
        sv.close()
        assert np.array_equal(spread[k][0], ax) and np.array_equal(spread[k][1], az), k


def test_request_larger_than_a_group():
    """With 8-shot groups, a 20-shot request opens three groups at once, and each of its shots equals the same
    shot served as a request of its own, bit for bit.  Requests above DGX_SERVE_MAX_SHOTS are refused."""
    print("case serve request over three groups, group 8", flush=True)
    spec = (0, 20, 150, 15.0)
    sv, rms, _, _ = new_server(group=8)
    req, xz, rec = build(sv, rms, spec, 30)
    rid = sv.submit(req)
    sv.advance(1)
    st = sv.stats()
    assert st["running"] == 1 and st["batches"] == 3, st
    sv.drain()
    x, z = sv.result(rid, req)
    with pytest.raises(L.DgxError):
        sv.submit(sv.make_request(rms[0], [(900.0, 900.0)] * 65, [15.0] * 65, 10, rec))
    sv.close()
    sv, rms, _, _ = new_server(group=8)
    singles = [sv.make_request(rms[0], xz[s:s + 1], [spec[3]], spec[2], rec) for s in range(spec[1])]
    ids = [sv.submit(r) for r in singles]
    sv.drain()
    for s in range(spec[1]):
        sx, sz = sv.result(ids[s], singles[s])
        assert np.array_equal(x[s], sx[0]) and np.array_equal(z[s], sz[0]), s
    sv.close()


def test_compaction_moves_shots_without_changing_results():
    """16-shot groups, blocks of 50 steps: A (10 shots, 200 steps) fills group 1, B (10 shots, 60 steps)
    takes the other 6 slots of group 1 and 4 of group 2, C (6 shots, 200 steps) joins group 2.  When B
    retires, A and C (16 shots) fit one group, so C's 6 shots move into group 1.  Every request's traces
    equal its traces alone, bit for bit."""
    print("case serve compaction, group 16", flush=True)
    specs = [(0, 10, 200, 15.0), (0, 10, 60, 20.0), (0, 6, 200, 12.0)]
    sv, rms, _, _ = new_server(group=16, K=50)
    built = [build(sv, rms, spec, 20 + k) for k, spec in enumerate(specs)]
    ids = [sv.submit(b[0]) for b in built]
    sv.advance(2)
    assert sv.stats()["moved"] == 0
    sv.advance(1)
    st = sv.stats()
    assert st["done"] == 1 and st["running"] == 2 and st["moved"] == 6, st
    sv.drain()
    out = [sv.result(rid, built[k][0]) for k, rid in enumerate(ids)]
    assert sv.stats()["moved"] == 6
    sv.close()
    for k, spec in enumerate(specs):
        sv, rms, _, _ = new_server(group=16, K=50)
        req, _, _ = build(sv, rms, spec, 20 + k)
        rid = sv.submit(req)
        sv.drain()
        ax, az = sv.result(rid, req)
        sv.close()
        assert np.array_equal(out[k][0], ax) and np.array_equal(out[k][1], az), k


@pytest.mark.parametrize("group", [1, 4, 8, 16, 32, 64])
def test_request_equals_standalone_solver_and_oracle(group):
    print(f"case serve vs standalone and oracle, group {group}", flush=True)
    k = 0
    (x, z), xz, rec = run_alone(k, group)
    sp, ms = meshes()
    mesh, model = ms[REQS[k][0]]
    nsteps, f0 = REQS[k][2], REQS[k][3]
    srcs = [dgx.SourceSpec(float(a), float(b), L.DGX_SRC_EXPLOSIVE, s, f0) for s, (a, b) in enumerate(xz)]
    st = dgx.build_setup(mesh, model, srcs, rec, nsteps, sponge=sp)
    if group == 1:  # the same tile code as the solver's blocked kernel (3 steps per launch, 512 threads)
        solver = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_ELEMENT, step_kernel=L.DGX_STEP_BLOCKED, block_steps=3)
    else:
        solver = dgx.Solver(st, L.DGX_FP32, L.DGX_LAYOUT_SHOT_INNER, group)
    solver.step(nsteps)
    for s in range(len(xz)):
        gx, gz = solver.traces(s)
        # same arithmetic in two kernels (the compiler may contract FMAs differently): FP32 rounding level
        e = max(rel_l2(x[s], gx), rel_l2(z[s], gz))
        print(f"shot {s}: served vs standalone solver {e:.3e}")
        assert e <= 1e-5, s
    solver.close()
    w = dgx.configs.Workload("serve", mesh, model, sp, srcs, rec, nsteps, 0)
    oc = OracleCase(w)
    worst = 0.0
    for s in range(len(xz)):
        ox, oz, _ = oc.run(shot=s)
        worst = max(worst, rel_l2(x[s], ox), rel_l2(z[s], oz))
    print(f"served fp32 vs fp64 oracle: worst trace rel L2 {worst:.3e}")
    assert worst <= 1e-4


@pytest.mark.parametrize("group", [1, 4, 8, 16, 32, 64])
def test_fp64_serving_matches_oracle(group):
    """FP64 serving reproduces the oracle to the plain path's parity bar, a request mixed with others."""
    print(f"case fp64 serve vs oracle, group {group}", flush=True)
    sv, rms, sp, ms = new_server(precision=L.DGX_FP64, group=group)
    reqs = [build(sv, rms, REQS[k], k) for k in (0, 1, 3)]
    ids = [sv.submit(r[0]) for r in reqs[:2]]
    sv.advance(1)
    ids.append(sv.submit(reqs[2][0]))  # joins while the others run
    sv.drain()
    worst = 0.0
    for (req, xz, rec), rid, k in zip(reqs, ids, (0, 1, 3)):
        x, z = sv.result(rid, req)
        mi, _, nsteps, f0 = REQS[k]
        mesh, model = ms[mi]
        srcs = [dgx.SourceSpec(float(a), float(b), L.DGX_SRC_EXPLOSIVE, s, f0) for s, (a, b) in enumerate(xz)]
        oc = OracleCase(dgx.configs.Workload("serve64", mesh, model, sp, srcs, rec, nsteps, 0))
        for s in range(len(xz)):
            ox, oz, _ = oc.run(shot=s)
            worst = max(worst, rel_l2(x[s], ox), rel_l2(z[s], oz))
    sv.close()
    print(f"served fp64 vs oracle: worst trace rel L2 {worst:.3e}")
    assert worst <= 1e-12


def test_many_slot_groups():
    """80 requests of 16 shots on one mesh fill 80 slot groups (80 batches in one launch): every sampled
    request equals the same request served alone, bit for bit."""
    print("case many slot groups", flush=True)
    sp, ms = meshes()

    def req(sv, rms, k):
        rng = np.random.default_rng(500 + k)
        xz = np.stack([rng.uniform(300, 1700, 16), rng.uniform(600, 1400, 16)], 1)
        rec = np.stack([np.linspace(200, 1800, 5), np.full(5, 1450.0)], 1)
        return sv.make_request(rms[1], xz, [20.0] * 16, 40 + (k % 7) * 10, rec)

    sv, rms, _, _ = new_server(group=16, K=25)
    reqs = [req(sv, rms, k) for k in range(80)]
    ids = [sv.submit(r) for r in reqs]
    sv.advance(1)
    assert sv.stats()["running"] == 80
    sv.drain()
    mixed = {k: sv.result(ids[k], reqs[k]) for k in (0, 41, 63, 64, 79)}
    sv.close()
    for k, (x, z) in mixed.items():
        sv1, rms1, _, _ = new_server(group=16, K=25)
        r = req(sv1, rms1, k)
        rid = sv1.submit(r)
        sv1.drain()
        x1, z1 = sv1.result(rid, r)
        sv1.close()
        assert np.array_equal(x, x1) and np.array_equal(z, z1), k


def test_tight_budget_queues_and_completes():
    sp, ms = meshes()
    T0 = ms[0][0].tri
    budget = 5 * T0 * 6 * 16 * 4 + 4 * T0 + 2_000_000  # one slot group of mesh 0 plus small buffers
    sv = S.Server(L.DGX_FP32, budget, 50)
    rm0 = sv.add_mesh(ms[0][0], ms[0][1], sp)
    reqs = [sv.make_request(rm0, [(500.0 + 100 * k, 900.0)], [15.0], 120, [[1000.0, 1500.0]]) for k in range(6)]
    reqs.append(sv.make_request(rm0, [(700.0, 800.0)] * 12, [15.0] * 12, 100, [[1000.0, 1500.0]]))
    ids = [sv.submit(r) for r in reqs]
    sv.advance(1)
    st = sv.stats()
    assert st["batches"] == 1 and st["queued"] >= 1  # the 12-shot request must wait for slots
    sv.drain()
    assert all(sv.status(i)[0] == 2 for i in ids)
    sv.close()


def test_idle_groups_are_reclaimed_for_another_mesh():
    """With room for one slot group only, a request on a second mesh still runs once the first mesh's
    group is idle: idle groups are freed when a queued request does not fit."""
    sp, ms = meshes()
    T0, T1 = ms[0][0].tri, ms[1][0].tri
    budget = 5 * max(T0, T1) * 6 * 16 * 4 + 4 * max(T0, T1) + 2_000_000
    sv = S.Server(L.DGX_FP32, budget, 50)
    rm0, rm1 = sv.add_mesh(ms[0][0], ms[0][1], sp), sv.add_mesh(ms[1][0], ms[1][1], sp)
    a = sv.submit(sv.make_request(rm0, [(900.0, 900.0)], [15.0], 100, [[1000.0, 1500.0]]))
    sv.drain()
    b = sv.submit(sv.make_request(rm1, [(900.0, 900.0)], [15.0], 100, [[1000.0, 1500.0]]))
    sv.drain()
    assert sv.status(a)[0] == 2 and sv.status(b)[0] == 2
    assert sv.stats()["batches"] == 1
    sv.close()


def test_request_that_can_never_fit_is_rejected():
    sp, ms = meshes()
    sv = S.Server(L.DGX_FP32, 1_000_000, 50)  # smaller than one slot group
    rm0 = sv.add_mesh(ms[0][0], ms[0][1], sp)
    with pytest.raises(L.DgxError):
        sv.submit(sv.make_request(rm0, [(900.0, 900.0)], [15.0], 100, [[1000.0, 1500.0]]))
    assert sv.stats()["queued"] == 0
    sv.close()
