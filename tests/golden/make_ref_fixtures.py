"""Writes the golden fixtures of the reference's own build (oracle/_ref/libref.so, compiled from the
reference sources by oracle/Makefile) for small meshes.  Run in a checkout where the reference exists:

    make -C oracle && python tests/golden/make_ref_fixtures.py

Each .npz holds only data: the input mesh and materials (so the case is self-contained), the scalars and
small vectors of the reference's setup, SHA-256 digests of its large setup arrays (Dx, Dz, Q, A, B,
absorb_fun, Neighbor, Edge_flag, Node_Neighbor), the receiver traces, and full field snapshots.
The reference build uses its own dg.h constants: 2 km square with 100 m sponge, Fp=15, Pi=3.1415,
one source (vertical force, Vz) and one receiver at mesh vertices.
"""
from __future__ import annotations

import hashlib
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "discontinuous-galerkin-method_amd" / "python"), str(ROOT / "oracle")]

import dgx  # noqa: E402
import oracle_py as O  # noqa: E402

OPDIR = ROOT / "discontinuous-galerkin-method_amd" / "input_files" / "precomputed_matrix"
HASHED = ("Dx", "Dz", "Q", "A", "B", "absorb_fun", "jacobi", "jacobi_edge", "n1", "n2", "n3", "neighbor",
          "edge_flag", "node_neighbor")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cases():
    sp = dgx.Sponge.for_domain(2000.0, 2000.0, 100.0)
    # reference-like: water (attr 1) over rock (attr 3), random diagonals
    m1 = dgx.structured(20, 20, 100.0, diagonal="random", seed=5, sponge_box=sp.box(),
                        attribute=lambda c: np.where(c[:, 1] > 1500, 1, 3))
    yield "two_media_800", m1, (1200.0, 1200.0), (1600.0, 1600.0), 400, 200
    # all rock, same diagonal, jittered interior (source and receiver stay on vertices)
    m2 = dgx.structured(25, 25, 80.0, diagonal="same", jitter=10.0, seed=9, sponge_box=sp.box(), attribute=3)
    p = m2.point
    src = p[np.argmin(np.hypot(p[:, 0] - 1000, p[:, 1] - 1000))]
    rec = p[np.argmin(np.hypot(p[:, 0] - 1500, p[:, 1] - 1700))]
    yield "rock_jitter_1250", m2, tuple(src), tuple(rec), 300, 300


def main():
    out_dir = pathlib.Path(__file__).resolve().parent
    for name, mesh, src, rec, nsteps, snap in cases():
        epart = (mesh.centroids()[:, 0] > 1000.0).astype(np.int64)  # Process_Num = 2 in dg.h
        res, rx, rz, snaps = O.ref_case(OPDIR, mesh, epart, src, rec, nsteps, snap)
        data = dict(point=mesh.point, triangle=mesh.triangle, neighbor=mesh.neighbor, attribute=mesh.attribute,
                    epart=epart, src=np.array(src), rec=np.array(rec), nsteps=nsteps, snap_every=snap,
                    dt=res["dt"], t0=res["t0"], source=np.array(res["source"]), receiver=np.array(res["receiver"]),
                    fun=res["fun"], rec_x=rx, rec_z=rz, snaps=snaps, absorb_num=res["absorb_num"],
                    hashed=np.array(HASHED), digests=np.array([digest(res[k]) for k in HASHED]))
        np.savez_compressed(out_dir / f"ref_{name}.npz", **data)
        print(name, mesh.tri, "elements", nsteps, "steps; dt", res["dt"], "source", res["source"])


if __name__ == "__main__":
    main()
