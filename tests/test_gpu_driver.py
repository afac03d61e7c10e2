"""The C driver (bin/dgx_run) end to end: reference-format mesh files in, the reference's snapshot and
seismogram files out, compared with the oracle (binary twins at 1e-12, text at its 14-decimal
quantisation) and, in the reference-arithmetic precision, with the reference build's own outputs (bit for
bit)."""
import pathlib
import subprocess

import numpy as np
import pytest

import dgx
from dgx import lib as L
from helpers import OracleCase, rel_l2

pytestmark = pytest.mark.gpu

BIN = L.PKG_DIR / "bin" / "dgx_run"


def test_driver_matches_oracle(tmp_path):
    assert BIN.exists(), "bin/dgx_run not built"
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(40, 40, 50.0, diagonal="random", seed=3, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    dgx.formats.write_mesh(tmp_path / "model.1", mesh)
    out = tmp_path / "output_files"
    out.mkdir()
    (tmp_path / "run.cfg").write_text(
        f"mesh {tmp_path / 'model.1'}\noperators {L.OPERATOR_DIR}\noutput {out}\nsteps 300\nsnapshot 100\n"
        "source 1200 1200 force\nreceiver 1600 1600\nreceiver_line 300 1450 1700 1450 8\n")
    r = subprocess.run([str(BIN), str(tmp_path / "run.cfg")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the oracle on the mesh as the driver read it back from the 8-digit text files
    lib = L.load()
    T, P = mesh.tri, mesh.point_num
    pt, tr, nb, at = np.zeros(P * 2), np.zeros(T * 3, np.int64), np.zeros(T * 3, np.int64), np.zeros(T, np.int64)
    paths = [str(tmp_path / f"model.1.{x}").encode() for x in ("ele", "node", "neigh")]
    L.check(lib.dgx_read_mesh(*paths, T, P, L.dp(pt), L.lp(tr), L.lp(nb), L.lp(at)))
    m2 = dgx.Mesh(pt.reshape(-1, 2), tr.reshape(-1, 3), nb.reshape(-1, 3), at)
    rec = np.array([[1600.0, 1600.0]] + [[x, 1450.0] for x in np.linspace(300, 1700, 8)])
    w = dgx.configs.Workload("driver", m2, dgx.Model.by_attribute(at), sp,
                             [dgx.SourceSpec(1200.0, 1200.0, L.DGX_SRC_FORCE_Z, 0, 15.0)], rec, 300, 100)
    ox, oz, snaps = OracleCase(w).run(snap_every=100)
    for k in (100, 200, 300):
        vx = np.fromfile(out / f"Vx{k}.bin")
        vz = np.fromfile(out / f"Vz{k}.bin")
        assert rel_l2(vx, snaps[k][0]) <= 1e-12 and rel_l2(vz, snaps[k][1]) <= 1e-12, k
        txt = np.loadtxt(out / f"Vx{k}")
        assert np.all(np.abs(txt - vx) <= 5.1e-15 + 1e-15 * np.abs(vx))  # "%0.14f" quantisation
    for j in range(len(rec)):
        gx = np.fromfile(out / f"record_x{j}.bin")
        gz = np.fromfile(out / f"record_z{j}.bin")
        assert rel_l2(gx, ox[:, j]) <= 1e-12 and rel_l2(gz, oz[:, j]) <= 1e-12, j
    allnode = np.loadtxt(out / "All_node.txt")
    assert allnode.shape == (T * 6, 2)


def test_driver_binary_mesh_equals_text_mesh(tmp_path):
    """mesh_binary holding the arrays the text reader produces gives bit-identical outputs."""
    assert BIN.exists(), "bin/dgx_run not built"
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    mesh = dgx.structured(30, 30, 2000.0 / 30, diagonal="random", seed=4, sponge_box=sp.box(),
                          attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    dgx.formats.write_mesh(tmp_path / "model.1", mesh)
    lib = L.load()
    T, P = mesh.tri, mesh.point_num
    pt, tr, nb, at = np.zeros(P * 2), np.zeros(T * 3, np.int64), np.zeros(T * 3, np.int64), np.zeros(T, np.int64)
    paths = [str(tmp_path / f"model.1.{x}").encode() for x in ("ele", "node", "neigh")]
    L.check(lib.dgx_read_mesh(*paths, T, P, L.dp(pt), L.lp(tr), L.lp(nb), L.lp(at)))
    L.check(lib.dgx_write_mesh_binary(str(tmp_path / "model.dgxm").encode(), T, P, L.dp(pt), L.lp(tr), L.lp(nb),
                                      L.lp(at)))
    outs = []
    for name, line in (("text", f"mesh {tmp_path / 'model.1'}"), ("bin", f"mesh_binary {tmp_path / 'model.dgxm'}")):
        out = tmp_path / name
        out.mkdir()
        (tmp_path / f"{name}.cfg").write_text(
            f"{line}\noperators {L.OPERATOR_DIR}\noutput {out}\nsteps 120\nsnapshot 60\n"
            "source 1200 1200 force\nreceiver_line 300 1450 1700 1450 4\n")
        print(f"case driver {name} mesh", flush=True)
        r = subprocess.run([str(BIN), str(tmp_path / f"{name}.cfg")], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(out)
    for f in ["Vx60.bin", "Vz120.bin", "record_x0.bin", "record_z3.bin"]:
        assert open(outs[0] / f, "rb").read() == open(outs[1] / f, "rb").read(), f


GOLDEN = sorted((pathlib.Path(__file__).parent / "golden").glob("ref_*.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_driver_reference_precision_reproduces_the_reference_build(tmp_path, path):
    """`precision reference`: the driver's snapshot and seismogram twins equal the outputs the reference's own
    build wrote for the same mesh (the committed fixtures), bit for bit; the mesh goes in losslessly as
    mesh_binary, and the sponge, materials, frequency and Pi are the driver's defaults, which are the
    reference's."""
    assert BIN.exists(), "bin/dgx_run not built"
    d = np.load(path, allow_pickle=False)
    T, P = len(d["triangle"]), len(d["point"])
    arrays = [np.ascontiguousarray(d[k].ravel(), t) for k, t in
              (("point", np.float64), ("triangle", np.int64), ("neighbor", np.int64), ("attribute", np.int64))]
    L.check(L.load().dgx_write_mesh_binary(str(tmp_path / "model.dgxm").encode(), T, P, L.dp(arrays[0]),
                                           L.lp(arrays[1]), L.lp(arrays[2]), L.lp(arrays[3])))
    n, every = int(d["nsteps"]), int(d["snap_every"])
    out = tmp_path / "out"
    out.mkdir()
    (tmp_path / "run.cfg").write_text(
        f"mesh_binary {tmp_path / 'model.dgxm'}\noperators {L.OPERATOR_DIR}\noutput {out}\nsteps {n}\n"
        f"snapshot {every}\nsource {float(d['src'][0])!r} {float(d['src'][1])!r} force\n"
        f"receiver {float(d['rec'][0])!r} {float(d['rec'][1])!r}\n"
        "precision reference\n")
    print(f"case driver reference precision {path.stem}", flush=True)
    r = subprocess.run([str(BIN), str(tmp_path / "run.cfg")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for k in range(n // every):
        step = (k + 1) * every
        assert np.array_equal(np.fromfile(out / f"Vx{step}.bin"), d["snaps"][k][0]), step
        assert np.array_equal(np.fromfile(out / f"Vz{step}.bin"), d["snaps"][k][1]), step
    assert np.array_equal(np.fromfile(out / "record_x.bin"), d["rec_x"])
    assert np.array_equal(np.fromfile(out / "record_z.bin"), d["rec_z"])
