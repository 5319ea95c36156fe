"""Writes tests/golden/stripes_0.6_golden.json: the CPU oracle's diffusion coefficients on a real relief.

Input: stripes_0.6.obj.xz, one of the optimised 0.6 m reliefs the reference ships as data
(submission/acoustic_reliefs/0.6/stripes_0.6.obj: 19,345 vertices, 38,686 triangles), stored compressed.
The oracle (oracle/, FP64) assembles the H-matrix (leaf 32, radius factor 1.5, ACA tol 1e-6), solves three
plane waves of the 37-direction fan at 1000 Hz to BiCGSTAB tol 1e-9 and evaluates the 181 semicircle
receivers at 5 m. Takes a few minutes on 8 threads.

    python tests/golden/make_relief_obj_golden.py
"""
import json
import lzma
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from acoustic_reliefs_amd import capi, meshes  # noqa: E402
from oracle import oracle as O  # noqa: E402

F, C0, ACA_TOL, TOL, PICK = 1000.0, 343.0, 1e-6, 1e-9, [0, 18, 36]


def load(path=os.path.join(HERE, "stripes_0.6.obj.xz")):
    tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), "stripes_0.6.obj")
    with lzma.open(path, "rb") as f, open(tmp, "wb") as g:
        g.write(f.read())
    P, E = meshes.read_obj(tmp)
    os.remove(tmp)
    return P, E


def main():
    t0 = time.perf_counter()
    P, E = load()
    perm, dense, lowrank, _ = O.partition(capi.centroids(P, E), 32, 1.5)
    oc = O.Ctx(P, E[perm], dense, lowrank)
    k = 2 * np.pi * F / C0
    oc.assemble(k, ACA_TOL, 100)
    dirs = meshes.incidence_directions(37)[PICK]
    L = meshes.semicircle_receivers(5.0, 181)
    Q = oc.listener_matrix(k, L)
    d, iters = [], []
    for v in dirs:
        x, it, err, _ = oc.bicgstab(-oc.rhs_plane(k, v), TOL)
        d.append(float(O.diffusion(np.abs(Q @ x))[0]))
        iters.append(int(it))
        print("direction", v, "iters", it, "d", d[-1], flush=True)
    out = {"mesh": "stripes_0.6.obj (submission/acoustic_reliefs/0.6)", "N": int(len(E)), "V": int(len(P)),
           "frequency": F, "c0": C0, "aca_tol": ACA_TOL, "bicgstab_tol": TOL, "directions": dirs.tolist(),
           "receivers": "181 on the 5 m semicircle", "d": d, "iters": iters,
           "bbox": [P.min(axis=0).tolist(), P.max(axis=0).tolist()]}
    with open(os.path.join(HERE, "stripes_0.6_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("done in %.1f s" % (time.perf_counter() - t0))


if __name__ == "__main__":
    main()
