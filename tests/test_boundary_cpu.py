"""The drop-in boundary without a GPU: the C-ABI library loads and exports every declared entry
point, the host-side partition is bit-exact with the oracle's, argument errors come back as error
codes, and the pybind11 module mirrors the reference's API (acoustics3d.cpp:353-383)."""
import inspect
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "arh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(arh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(capi):
    syms = declared_symbols()
    assert len(syms) >= 25
    lib = capi.lib()
    for s in syms:
        assert hasattr(lib, s), s
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert bound == set(syms), set(syms) ^ bound


@pytest.mark.parametrize("leaf", [8, 32, 64, 128])
def test_partition_bit_exact_with_oracle(capi, oracle, sphere_mesh, small_relief, leaf):
    from acoustic_reliefs_amd import meshes
    bodies = [sphere_mesh, small_relief[:2], meshes.relief_config(0.3, 24, 0.05, seed=4)[:2]]
    for P, E in bodies:
        C = capi.centroids(P, E)
        perm, d, a = capi.partition(C, leaf, 1.5)
        operm, od, oa, _ = oracle.partition(C, leaf, 1.5)
        assert np.array_equal(perm, operm)
        assert np.array_equal(d, od)
        assert np.array_equal(a, oa)


def test_partition_errors(capi):
    C = np.zeros((10, 3))  # coincident centroids: the reference recursion never ends
    with pytest.raises(RuntimeError, match="coincident"):
        capi.partition(C, 4, 1.5)
    C = np.random.default_rng(0).random((50, 3))
    C[3, 1] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        capi.partition(C, 4, 1.5)
    perm = np.zeros(50, np.int32)
    nd, na = capi.c_int(), capi.c_int()
    small = np.zeros(4, np.int32)
    C = np.random.default_rng(0).random((50, 3))
    rc = capi.lib().arh_partition(capi._dp(C), 50, 4, 1.5, capi._ip(perm), capi._ip(small), 1, capi.ctypes.byref(nd),
                                  capi.ctypes.byref(na))
    assert rc == -1 and b"capacity" in capi.lib().arh_last_error(None)


def test_module_api_surface():
    from acoustic_reliefs_amd import acoustics3d_impl as m
    import acoustics3d
    assert acoustics3d.DiffBEM is m.DiffBEM and acoustics3d.sphere is m.sphere
    for name in ["set_mesh", "set_diff_points", "precompute", "value", "band_value", "values", "gradient", "pvals",
                 "get_listeners", "surface_vals", "get_mesh", "get_Hs", "set_band"]:
        assert hasattr(m.DiffBEM, name), name
    def default(doc, name):
        mm = re.search(name + r": [^=]+= (\[[^\]]*\]|[^,)]+)", doc)
        assert mm, name
        return mm.group(1).strip()

    doc = m.DiffBEM.__init__.__doc__
    want = {"cluster_size": "64", "radius_factor": "1.5", "freq_bands": "[1000.0]", "n_freqs": "1",
            "approx_ACA_tol": "0.001", "Q_ACA_tol": "0.001", "solver_tol": "0.001", "src_pt": "[0.0, 100.0, 0.0]",
            "listener_radius": "50", "listener_ds": "5", "recompute_matrices": "False"}
    for k, v in want.items():
        assert default(doc, k) == v, (k, default(doc, k))
    pos = [doc.index(o + ":") for o in want]
    assert pos == sorted(pos)  # positional order of the reference binding
    sdoc = m.sphere.__doc__
    assert default(sdoc, "LL") == "100" and default(sdoc, "lrad") == "10" and default(sdoc, "actual") == "False"
    assert m.sphere.__doc__.index("sphere_mesh:") < m.sphere.__doc__.index("freq:")


def test_module_state_errors_without_gpu(capfd):
    """Not-set guards print to stderr and return the reference's sentinels (acoustics3d.cpp:110-113)."""
    from acoustic_reliefs_amd import acoustics3d_impl as m
    d = m.DiffBEM(128, 1.5, [500.0, 1000.0], 1, 1e-5, 1e-5, 1e-5, np.array([0.0, 100.0, 0.0]), 50, 5, False)
    assert d.value() == -1
    assert d.band_value(1000.0) == -1
    assert d.values() == []
    v, g = d.gradient(np.zeros(3))
    assert v == 0.0 and len(g) == 0
    err = capfd.readouterr().err
    assert "No mesh set" in err
    d.silent = True
    d.use_actual = True
    assert d.silent and d.use_actual
    d.set_diff_points([5, 7, 5, 9])  # a repeated vertex keeps its last slot (acoustics3d.cpp:74-77)
    assert d.get_Hs() == {5: 2, 7: 1, 9: 3}
    d.set_band(2000.0)
    kw = m.DiffBEM(cluster_size=32, freq_bands=[250.0], speed_of_sound=343.0)
    assert kw.get_Hs() == {}


def test_meshes():
    from acoustic_reliefs_amd import meshes
    P, E = meshes.icosphere(3, 0.5)
    assert P.shape == (642, 3) and E.shape == (1280, 3)
    a, b, c = P[E[:, 0]], P[E[:, 1]], P[E[:, 2]]
    n = np.cross(b - a, c - a)
    assert (np.einsum("ij,ij->i", n, a + b + c) > 0).all()  # outward
    P, E, top = meshes.relief_config(0.6, 64, 0.1, seed=0)
    assert len(E) == 9828 and len(top) == 4096
    a, b, c = P[E[:, 0]], P[E[:, 1]], P[E[:, 2]]
    assert np.abs(np.cross(b - a, c - a).sum(0)).max() < 1e-12  # closed surface: normals sum to 0
    h = meshes.smooth_heights(64, 0, 2.0, 0.1)
    assert h.min() == 0.0 and abs(h.max() - 0.1) < 1e-15
    assert abs(h.sum() - 228.95392631910167) < 1e-9  # the workload's height checksum
    assert np.allclose(P[top, 1], 0.05 + h.reshape(-1))
    L = meshes.semicircle_receivers()
    assert L.shape == (181, 3) and np.allclose(np.linalg.norm(L, axis=1), 5.0)
    D = meshes.incidence_directions(37)
    assert D.shape == (37, 3) and np.allclose(np.linalg.norm(D, axis=1), 1.0) and (D[1:-1, 1] < 0).all()


def test_obj_roundtrip(tmp_path):
    from acoustic_reliefs_amd import meshes
    P, E = meshes.icosphere(1, 1.0)
    f = tmp_path / "s.obj"
    meshes.write_obj(str(f), P, E)
    P2, E2 = meshes.read_obj(str(f))
    assert np.array_equal(E, E2) and np.array_equal(P, P2)
