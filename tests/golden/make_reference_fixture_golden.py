"""Writes oracle golden values for two of the reference's own data files (stored compressed here):

* sphere_m.obj.xz -- test-data/spheres/sphere_m.obj (4,034 vertices, 8,064 triangles, radius 1), the mesh of the
  reference's sphere validation (misc/analytic_sphere.py:32-60: sphere("test-data/spheres/sphere_m.obj", 2000,
  LL=360, lrad=10)). The oracle restates sphere() (sphere.h:28-85): leaf 128, radius factor 1.5, k = 2 pi f / 344
  (constants.h:13), plane wave b = e^{-ik x}, ACA and BiCGSTAB tolerance 1e-10, listener ACA at 1e-10 on the
  circle (r cos t, 0, -r sin t). Output sphere_m_golden.json: the 360 scattered pressures, the iteration count
  and the oracle's mean absolute error against the analytic series (misc/analytic_sphere.py:22-26, c = 344).
* mountains_0.9.obj.xz -- submission/acoustic_reliefs/0.9/mountains_0.9.obj (41,995 vertices, 83,986 triangles).
  The oracle assembles the H-matrix at 250 Hz (c0 = 343, leaf 32, radius factor 1.5, ACA 1e-4), solves plane
  waves 0, 18 and 36 of the 37-direction fan to BiCGSTAB tol 1e-8 and evaluates the 181 semicircle receivers at
  5 m (exact listener matrix). Output mountains_0.9_golden.json.

Takes a few minutes on 8 threads.

    python tests/golden/make_reference_fixture_golden.py [sphere|relief]
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


def load(name):
    tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), name)
    with lzma.open(os.path.join(HERE, name + ".xz"), "rb") as f, open(tmp, "wb") as g:
        g.write(f.read())
    P, E = meshes.read_obj(tmp)
    os.remove(tmp)
    return P, E


def sphere_golden(freq=2000.0, LL=360, lrad=10.0, c0=344.0, tol=1e-10):
    t0 = time.perf_counter()
    P, E = load("sphere_m.obj")
    Er, dirb, appb, _ = O.setup(P, E, 128, 1.5)
    oc = O.Ctx(P, Er, dirb, appb)
    k = 2 * np.pi * freq / c0
    oc.assemble(k, tol, 100)
    x, it, err, _ = oc.bicgstab(-oc.rhs_plane(k, [1.0, 0.0, 0.0]), tol)
    t = np.arange(LL) / LL * 2 * np.pi
    L = np.column_stack([lrad * np.cos(t), np.zeros(LL), -lrad * np.sin(t)])
    U, V = oc.aca_listener(k, L, tol)
    y = U @ (V @ x)
    # the analytic scattered field at the same points (theta from the propagation direction +x)
    ya = O.sphere_scattered(k, 1.0, lrad, np.cos(t), nterms=100)
    mae = float(np.mean(np.abs(y - ya)))
    rmae = float(np.mean(np.abs(y - ya) / np.abs(ya)))
    out = {"mesh": "test-data/spheres/sphere_m.obj", "N": int(len(E)), "V": int(len(P)), "frequency": freq,
           "c0": c0, "LL": LL, "lrad": lrad, "tol": tol, "iters": int(it), "listener_rank": int(U.shape[1]),
           "y_re": y.real.tolist(), "y_im": y.imag.tolist(), "oracle_mae_vs_analytic": mae,
           "oracle_rel_mae_vs_analytic": rmae}
    with open(os.path.join(HERE, "sphere_m_golden.json"), "w") as f:
        json.dump(out, f)
    print("sphere_m: N=%d iters %d listener rank %d MAE %.4e rel MAE %.4e (%.0f s)" % (
        len(E), it, U.shape[1], mae, rmae, time.perf_counter() - t0), flush=True)


def relief_golden(F=250.0, c0=343.0, aca_tol=1e-4, tol=1e-8, pick=(0, 18, 36)):
    t0 = time.perf_counter()
    P, E = load("mountains_0.9.obj")
    perm, dense, lowrank, _ = O.partition(capi.centroids(P, E), 32, 1.5)
    oc = O.Ctx(P, E[perm], dense, lowrank)
    k = 2 * np.pi * F / c0
    oc.assemble(k, aca_tol, 100)
    print("assembled (%.0f s)" % (time.perf_counter() - t0), flush=True)
    dirs = meshes.incidence_directions(37)[list(pick)]
    L = meshes.semicircle_receivers(5.0, 181)
    Q = oc.listener_matrix(k, L)
    d, iters = [], []
    for v in dirs:
        x, it, err, _ = oc.bicgstab(-oc.rhs_plane(k, v), tol)
        d.append(float(O.diffusion(np.abs(Q @ x))[0]))
        iters.append(int(it))
        print("direction", v, "iters", it, "d", d[-1], "(%.0f s)" % (time.perf_counter() - t0), flush=True)
    out = {"mesh": "submission/acoustic_reliefs/0.9/mountains_0.9.obj", "N": int(len(E)), "V": int(len(P)),
           "frequency": F, "c0": c0, "aca_tol": aca_tol, "bicgstab_tol": tol, "directions": dirs.tolist(),
           "receivers": "181 on the 5 m semicircle", "d": d, "iters": iters,
           "bbox": [P.min(axis=0).tolist(), P.max(axis=0).tolist()]}
    with open(os.path.join(HERE, "mountains_0.9_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("done in %.1f s" % (time.perf_counter() - t0))


if __name__ == "__main__":
    which = sys.argv[1:] or ["sphere", "relief"]
    if "sphere" in which:
        sphere_golden()
    if "relief" in which:
        relief_golden()
