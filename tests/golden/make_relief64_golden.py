"""Writes tests/golden/relief64_golden.npz: the CPU oracle's solution of the benchmark relief workload.

Workload: the 0.6 m relief, closed box with a 0.05 m base and a 64 x 64 vertex top from seeded smoothed
noise scaled to [0, 0.1] m (meshes.relief_config(0.6, 64, 0.1, seed=0), N = 9,828); c0 = 343 m/s; leaf
32, radius factor 1.5, ACA tol 1e-4; the 37 plane waves of the 0..180 degree fan; the 181 receivers on
the 5 m semicircle (exact listener matrix). For the bands 250, 1000 and 4000 Hz the oracle (oracle/, FP64)
assembles the H-matrix and solves every direction with BiCGSTAB to a recursive residual of 1e-8, then
stores:

* d: the 3 x 37 diffusion coefficients;
* iters: the 3 x 37 BiCGSTAB iteration counts;
* x: the surface pressures (reordered element order) of directions 0, 18, 36 per band, 3 x 3 x N,
  stored as complex64 (the tests compare them at ~1e-5);
* y: the 181 receiver pressures of the same directions, 3 x 3 x 181;
* the mesh checksums the test checks before comparing.

Takes about 10-15 minutes on 8 threads.

    python tests/golden/make_relief64_golden.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from acoustic_reliefs_amd import capi, meshes  # noqa: E402
from oracle import oracle as O  # noqa: E402

BANDS, C0, ACA_TOL, TOL, PICK = [250.0, 1000.0, 4000.0], 343.0, 1e-4, 1e-8, [0, 18, 36]


def main():
    t0 = time.perf_counter()
    P, E, top = meshes.relief_config(0.6, 64, 0.1, seed=0)
    perm, dense, lowrank, _ = O.partition(capi.centroids(P, E), 32, 1.5)
    Er = E[perm]
    oc = O.Ctx(P, Er, dense, lowrank)
    dirs = meshes.incidence_directions(37)
    L = meshes.semicircle_receivers(5.0, 181)
    nb, R, N = len(BANDS), len(dirs), len(Er)
    d = np.zeros((nb, R))
    iters = np.zeros((nb, R), np.int32)
    xs = np.zeros((nb, len(PICK), N), np.complex128)
    ys = np.zeros((nb, len(PICK), len(L)), np.complex128)
    for bi, F in enumerate(BANDS):
        k = 2 * np.pi * F / C0
        oc.assemble(k, ACA_TOL, 100)
        Q = oc.listener_matrix(k, L)
        for j, v in enumerate(dirs):
            x, it, err, _ = oc.bicgstab(-oc.rhs_plane(k, v), TOL)
            y = Q @ x
            d[bi, j] = O.diffusion(np.abs(y))[0]
            iters[bi, j] = it
            if j in PICK:
                xs[bi, PICK.index(j)] = x
                ys[bi, PICK.index(j)] = y
            print("band %g Hz direction %d: iters %d d %.8f (%.0f s)" % (F, j, it, d[bi, j], time.perf_counter() - t0),
                  flush=True)
    np.savez_compressed(os.path.join(HERE, "relief64_golden.npz"), bands=np.array(BANDS), c0=C0, aca_tol=ACA_TOL,
                        bicgstab_tol=TOL, pick=np.array(PICK), directions=dirs, d=d, iters=iters, x=xs.astype(np.complex64), y=ys,
                        N=N, V=len(P), P_sum=P.sum(axis=0), E_sum=np.int64(E.astype(np.int64).sum()),
                        perm=perm.astype(np.int32), top=np.asarray(top, np.int32))
    print("done in %.1f s" % (time.perf_counter() - t0))


if __name__ == "__main__":
    main()
