"""The H-matrix products against an independent product of the GPU's own stored operator, at rounding level.

y = -x/2 + sum_b D_b x[cs:ce] -> y[rs:re] + sum_a U_a (V_a x[cs:ce]) (hmatrix.h:84-140), and its
(non-conjugate) transpose, built in numpy from the dense blocks and factors the library stores
(arh_hmat_dense_block, arh_hmat_factors). The comparison does not depend on ACA pivots, so it pins the
product kernels (layouts, tiles, passes, split leaves, summation) to FP64 rounding: <= 1e-12 relative per
right-hand side for FP64 storage; <= 2e-6 for the complex-FP32 copy (entries and intermediates rounded to
FP32, FP32 accumulation).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "mixed": 2e-6}


def stored_product(h, X, transpose):
    """numpy product of the stored operator; X is (nrhs, N)."""
    g = h.geom
    Y = -0.5 * X
    for b, (rs, re, cs, ce) in enumerate(g.dense):
        D = h.dense_block(b)
        if transpose:
            Y[:, cs:ce] += X[:, rs:re] @ D
        else:
            Y[:, rs:re] += X[:, cs:ce] @ D.T
    for a, (rs, re, cs, ce) in enumerate(g.lowrank):
        U, V = h.factors(a)
        if U.shape[1] == 0:
            continue
        if transpose:
            Y[:, cs:ce] += (X[:, rs:re] @ U) @ V
        else:
            Y[:, rs:re] += (X[:, cs:ce] @ V.T) @ U.T
    return Y


def _check(capi, h, X, storage, label):
    prec = capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED
    worst = 0.0
    for tr in (False, True):
        ref = stored_product(h, X.copy(), tr)
        Y = h.matvec_np(X, transpose=tr, precision=prec)
        err = np.linalg.norm(Y - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
        print("%s %s transpose=%s: max rel err %.2e over %d rhs" % (label, storage, tr, err.max(), len(X)),
              flush=True)
        assert err.max() <= TOL[storage], (label, storage, tr, err.max())
        worst = max(worst, err.max())
    return worst


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
@pytest.mark.parametrize("leaf", [32, 128, 256])
def test_product_vs_stored_operator(ctx, capi, storage, leaf):
    """Leaf 32 (2-row-tile MFMA class), 128 (8-row-tile class), 256 (generic row-owner kernels), with 1
    and 37 right-hand sides (several passes of the wide classes)."""
    from acoustic_reliefs_amd import meshes
    P, E = meshes.icosphere(4, 0.5) if leaf > 32 else meshes.relief_config(0.3, 24, 0.05, seed=4)[:2]
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=leaf, factor=1.5)
    k = 2 * np.pi * 1500.0 / 343.0
    h = g.assemble(k, 1e-5, 100, capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED)
    rng = np.random.default_rng(leaf)
    for nrhs in (1, 37):
        X = rng.standard_normal((nrhs, g.N)) + 1j * rng.standard_normal((nrhs, g.N))
        _check(capi, h, X, storage, "leaf %d N=%d nrhs=%d" % (leaf, g.N, nrhs))


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
def test_product_vs_stored_operator_relief_4000hz(ctx, capi, storage):
    """The benchmark's heaviest operator: 0.6 m relief with a 64 x 64 top (N = 9,828), 4000 Hz,
    ACA 1e-4, leaf 32, the 37 plane waves' solutions as right-hand sides."""
    from acoustic_reliefs_amd import meshes
    P, E, _ = meshes.relief_config(0.6, 64, 0.1, seed=0)
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    assert g.N == 9828
    k = 2 * np.pi * 4000.0 / 343.0
    h = g.assemble(k, 1e-4, 100, capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED)
    X = g.rhs(k, capi.ARH_RHS_PLANE, meshes.incidence_directions(37)).numpy()
    _check(capi, h, X, storage, "relief 0.6 m 4000 Hz")
