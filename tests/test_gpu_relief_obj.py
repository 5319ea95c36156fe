"""A real relief from the reference's data (submission/acoustic_reliefs/0.6/stripes_0.6.obj, 38,686 triangles,
stored as tests/golden/stripes_0.6.obj.xz): read with the OBJ reader (read_obj.cpp:8-45), solved on the GPU in
FP64 and in mixed precision, diffusion coefficients against the CPU oracle's (tests/golden/stripes_0.6_golden.json,
written by tests/golden/make_relief_obj_golden.py). Tolerance 5e-6: both sides compress with ACA at 1e-6, with
pivots that differ at near-ties."""
import json
import lzma
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def relief(tmp_path_factory):
    from acoustic_reliefs_amd import meshes
    path = str(tmp_path_factory.mktemp("obj") / "stripes_0.6.obj")
    with lzma.open(os.path.join(GOLDEN, "stripes_0.6.obj.xz"), "rb") as f, open(path, "wb") as g:
        g.write(f.read())
    P, E = meshes.read_obj(path)
    with open(os.path.join(GOLDEN, "stripes_0.6_golden.json")) as f:
        gold = json.load(f)
    return P, E, gold


def test_reader_counts(relief):
    P, E, gold = relief
    assert len(P) == gold["V"] == 19345 and len(E) == gold["N"] == 38686
    assert np.allclose(P.min(axis=0), gold["bbox"][0]) and np.allclose(P.max(axis=0), gold["bbox"][1])


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
def test_real_relief_vs_oracle(relief, ctx, capi, storage):
    from acoustic_reliefs_amd import meshes
    P, E, gold = relief
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    k = 2 * np.pi * gold["frequency"] / gold["c0"]
    st_id = capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED
    print("real relief N=%d %s" % (g.N, storage), flush=True)
    d, _, _, st, _ = g.band_solve(k, capi.ARH_RHS_PLANE, np.array(gold["directions"]), meshes.semicircle_receivers(5.0, 181),
                                  aca_tol=gold["aca_tol"], storage=st_id, tol=gold["bicgstab_tol"])
    print("d", d, "golden", gold["d"], "true residual", st["max_true_res"])
    assert st["max_true_res"] <= 1e-8
    assert np.abs(d - np.array(gold["d"])).max() < 5e-6
