"""The reference's own data files, stored compressed under tests/golden/, on the GPU path:

* sphere_m.obj (test-data/spheres, 8,064 triangles, radius 1): the reference's sphere validation protocol
  (misc/analytic_sphere.py:32-60) runs sphere("test-data/spheres/sphere_m.obj", 2000, LL=360, lrad=10). The
  drop-in's sphere() must give the oracle's scattered field (tests/golden/sphere_m_golden.json, written by
  tests/golden/make_reference_fixture_golden.py) to 1e-6 relative, and the same error against the analytic
  series (c = 344 on both sides) as the oracle. At 2000 Hz the mesh has about 2.5 elements per wavelength, so
  that error is the discretisation's (relative L2 0.30, mean absolute error 0.053); at 500 Hz (ka = 9.1) the
  result is within 1.5 % (relative L2) of the series.
* mountains_0.9.obj (submission/acoustic_reliefs/0.9, 83,986 triangles): diffusion coefficients of three plane
  waves at 250 Hz (ACA 1e-4, BiCGSTAB 1e-8, 181 receivers at 5 m) against the oracle's
  (tests/golden/mountains_0.9_golden.json), FP64 and mixed precision, within 1e-5.
"""
import json
import lzma
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _unpack(tmp_path_factory, name):
    path = str(tmp_path_factory.mktemp("obj") / name)
    with lzma.open(os.path.join(GOLDEN, name + ".xz"), "rb") as f, open(path, "wb") as g:
        g.write(f.read())
    return path


@pytest.fixture(scope="module")
def sphere_m(tmp_path_factory):
    with open(os.path.join(GOLDEN, "sphere_m_golden.json")) as f:
        gold = json.load(f)
    return _unpack(tmp_path_factory, "sphere_m.obj"), gold


def test_sphere_m_reference_protocol(sphere_m, oracle):
    from acoustic_reliefs_amd import acoustics3d_impl as m
    path, gold = sphere_m
    y = m.sphere(path, gold["frequency"], LL=gold["LL"], lrad=gold["lrad"])  # speed_of_sound 344 by default
    yo = np.array(gold["y_re"]) + 1j * np.array(gold["y_im"])
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    t = np.arange(gold["LL"]) / gold["LL"] * 2 * np.pi
    k = 2 * np.pi * gold["frequency"] / gold["c0"]
    ya = oracle.sphere_scattered(k, 1.0, gold["lrad"], np.cos(t), nterms=100)
    mae = float(np.mean(np.abs(y - ya)))
    print("sphere_m 2000 Hz: rel diff to the oracle %.2e; MAE vs analytic %.5f (oracle %.5f), rel MAE %.4f" % (
        rel, mae, gold["oracle_mae_vs_analytic"], float(np.mean(np.abs(y - ya) / np.abs(ya)))))
    assert rel <= 1e-6
    assert abs(mae - gold["oracle_mae_vs_analytic"]) <= 0.01 * gold["oracle_mae_vs_analytic"]


def test_sphere_m_vs_analytic_500hz(sphere_m, oracle):
    from acoustic_reliefs_amd import acoustics3d_impl as m
    path, gold = sphere_m
    y = m.sphere(path, 500.0, LL=360, lrad=10.0)
    t = np.arange(360) / 360 * 2 * np.pi
    ya = oracle.sphere_scattered(2 * np.pi * 500.0 / 344.0, 1.0, 10.0, np.cos(t), nterms=100)
    rel = np.linalg.norm(y - ya) / np.linalg.norm(ya)
    print("sphere_m 500 Hz: relative L2 error vs the analytic series %.4f" % rel)
    assert rel <= 0.015


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
def test_mountains_0p9_vs_oracle(tmp_path_factory, ctx, capi, storage):
    from acoustic_reliefs_amd import meshes
    with open(os.path.join(GOLDEN, "mountains_0.9_golden.json")) as f:
        gold = json.load(f)
    P, E = meshes.read_obj(_unpack(tmp_path_factory, "mountains_0.9.obj"))
    assert len(P) == gold["V"] and len(E) == gold["N"] == 83986
    assert np.allclose(P.min(axis=0), gold["bbox"][0]) and np.allclose(P.max(axis=0), gold["bbox"][1])
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    k = 2 * np.pi * gold["frequency"] / gold["c0"]
    st_id = capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED
    print("mountains_0.9 N=%d %s" % (g.N, storage), flush=True)
    d, _, _, st, info = g.band_solve(k, capi.ARH_RHS_PLANE, np.array(gold["directions"]),
                                     meshes.semicircle_receivers(5.0, 181), aca_tol=gold["aca_tol"], storage=st_id,
                                     tol=gold["bicgstab_tol"])
    print("d", d, "golden", gold["d"], "true residual %.2e" % st["max_true_res"], "H-matrix %.2f GB" % (
        info["bytes_total"] / 1e9))
    assert st["max_true_res"] <= 1e-8
    assert np.abs(d - np.array(gold["d"])).max() <= 1e-5
