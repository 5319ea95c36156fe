"""The drop-in module on the GPU: DiffBEM's forward methods and sphere() against the oracle and the
analytic rigid-sphere series (calls as acoustics_pipeline.py / misc/analytic_sphere.py make them)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bem(small_relief):
    from acoustic_reliefs_amd import acoustics3d_impl as m
    P, E, top = small_relief
    d = m.DiffBEM(32, 1.5, [500.0, 1000.0, 2000.0], 1, 1e-7, 1e-7, 1e-9, np.array([0.0, 100.0, 0.0]), 50, 5, False,
                  speed_of_sound=343.0)
    d.silent = True
    Ps, Es = d.set_mesh(P, E)
    d.set_diff_points(top)
    return d, Ps, Es, top


@pytest.fixture(scope="module")
def oc(oracle, bem, small_relief):
    from acoustic_reliefs_amd import capi
    d, Ps, Es, top = bem
    C = capi.centroids(Ps, Es)
    perm, dirb, appb, _ = oracle.partition(capi.centroids(small_relief[0], small_relief[1]), 32, 1.5)
    return oracle.Ctx(Ps, Es, dirb, appb)


def test_set_mesh_contract(bem, oracle, small_relief):
    d, Ps, Es, top = bem
    P, E, _ = small_relief
    assert Ps.dtype == np.float64 and Es.dtype == np.int32 and np.array_equal(Ps, P)
    perm, _, _, _ = oracle.partition(oracle.intermediates(P, E)[0], 32, 1.5)
    assert np.array_equal(Es, E[perm])  # the caller must use the returned (reordered) Es
    L = d.get_listeners()
    assert L.shape == (1297, 3) and np.allclose(L, oracle.listener_points(50, 5))
    Pm, Em = d.get_mesh()
    assert np.array_equal(Pm, P) and np.array_equal(Em, Es)
    x = np.linspace(0, 0.01, len(top))
    Px, _ = d.get_mesh(x)
    assert np.allclose(Px[top, 1] - P[top, 1], x)


def test_values_vs_oracle(bem, oc, oracle):
    d, Ps, Es, top = bem
    vals = d.values()
    L = d.get_listeners()
    ref = [oracle.band_value_grad(oc, [2 * np.pi * F / 343.0], L, aca_tol=1e-7, solver_tol=1e-9, q_aca_tol=1e-7)
           for F in (500.0, 1000.0, 2000.0)]
    print("values", vals, "oracle", ref)
    assert np.abs(np.array(vals) - np.array(ref)).max() < 1e-6
    assert abs(d.value() - np.mean(ref)) < 1e-6
    assert abs(d.band_value(1000.0) - ref[1]) < 1e-6


def test_value_x_vs_oracle(bem, oc, oracle):
    d, Ps, Es, top = bem
    x = 0.02 * np.sin(np.arange(len(top)) * 0.37)
    v = d.value(x)
    P = Ps.copy()
    P[top, 1] += x
    oc.set_vertices(P)
    ref = np.mean([oracle.band_value_grad(oc, [2 * np.pi * F / 343.0], d.get_listeners(), aca_tol=1e-7,
                                          solver_tol=1e-9, q_aca_tol=1e-7) for F in (500.0, 1000.0, 2000.0)])
    oc.set_vertices(Ps)
    assert abs(v - ref) < 1e-6


def test_pvals_surface_vals(bem, oc, oracle):
    d, Ps, Es, top = bem
    f = 1250.0
    k = 2 * np.pi * f / 343.0
    oc.assemble(k, 1e-7, 100)
    xo, _, _, _ = oc.bicgstab(-oc.rhs_point(k, [0.0, 100.0, 0.0]), 1e-9)
    xs = d.surface_vals(f)
    assert xs.dtype == np.complex128 and rel(xs, xo) < 1e-6
    Lc = np.array([[5.0 * np.cos(t), 5.0 * np.sin(t), 0.0] for t in np.linspace(0, np.pi, 19)])
    # pvals: y = U (V x) from the listener ACA at Q_ACA_tol = 1e-7 (acoustics3d.cpp:275)
    p = d.pvals(f, Lc)
    U, V = oc.aca_listener(k, Lc, 1e-7)
    assert rel(p, U @ (V @ xo)) < 1e-6 and rel(p, oc.listener_matrix(k, Lc) @ xo) < 1e-5
    p0 = d.pvals(f)
    U, V = oc.aca_listener(k, d.get_listeners(), 1e-7)
    assert p0.shape == (1297,) and rel(p0, U @ (V @ xo)) < 1e-6
    # pvals and surface_vals take the H-matrix path whatever use_actual says (acoustics3d.cpp:240-316)
    d.use_actual = True
    try:
        assert rel(d.pvals(f, Lc), p) < 1e-12 and rel(d.surface_vals(f), xs) < 1e-12
    finally:
        d.use_actual = False


def test_use_actual_dense(bem, oc, oracle):
    d, Ps, Es, top = bem
    d.use_actual = True
    try:
        v = d.band_value(1000.0)
    finally:
        d.use_actual = False
    ref = oracle.band_value_grad(oc, [2 * np.pi * 1000.0 / 343.0], d.get_listeners(), use_actual=True)
    assert abs(v - ref) < 1e-8


def test_sphere_vs_analytic(tmp_path, oracle):
    """ac3d.sphere on an OBJ sphere (r = 1, 5120 triangles) vs the analytic scattered field, with the
    same c on both sides (the reference's script mixes 343 and 344)."""
    from acoustic_reliefs_amd import acoustics3d_impl as m
    from acoustic_reliefs_amd import meshes
    P, E = meshes.icosphere(4, 1.0)
    path = str(tmp_path / "sphere.obj")
    meshes.write_obj(path, P, E)
    f, r, LL = 300.0, 10.0, 72
    y = m.sphere(path, f, LL=LL, lrad=r, actual=False)
    k = 2 * np.pi * f / 344.0
    th = np.arange(LL) / LL * 2 * np.pi
    ya = oracle.sphere_scattered(k, 1.0, r, np.cos(th))
    print("sphere: rel L2 err vs analytic %.3e" % rel(y, ya))
    assert rel(y, ya) < 0.03
    ya2 = m.sphere(path, f, LL=LL, lrad=r, actual=True)
    assert rel(ya2, y) < 1e-6
