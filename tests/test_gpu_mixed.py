"""Mixed precision on the GPU: the complex-FP32 MFMA multi-rhs product, FP64-refined solves, and the
independence of every right-hand side from the batch it is solved in.

Tolerances: FP32 storage and accumulation give ~1e-6 relative products; solves are refined to an
FP64 true residual <= 1e-6; diffusion coefficients agree with the FP64 path to <= 1e-4.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def relief(ctx, capi):
    from acoustic_reliefs_amd import meshes
    P, E, top = meshes.relief_config(0.6, 32, 0.1, seed=0)
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    k = 2 * np.pi * 1000.0 / 343.0
    h = g.assemble(k, 1e-4, 100, capi.ARH_MIXED)
    return g, h, k


def test_mfma_product_vs_fp64(relief, capi):
    g, h, k = relief
    rng = np.random.default_rng(0)
    for R in (37, 8, 5, 45):
        X = rng.standard_normal((R, g.N)) + 1j * rng.standard_normal((R, g.N))
        Y64 = h.matvec_np(X)
        Y32 = h.matvec_np(X, precision=capi.ARH_MIXED)
        e = max(rel(Y32[j], Y64[j]) for j in range(R))
        print("R=%d max rel diff fp32/fp64 %.2e" % (R, e))
        assert e < 5e-6


def test_mfma_product_batch_independent(relief, capi):
    g, h, k = relief
    rng = np.random.default_rng(1)
    X = rng.standard_normal((37, g.N)) + 1j * rng.standard_normal((37, g.N))
    Y = h.matvec_np(X, precision=capi.ARH_MIXED)
    Y8 = h.matvec_np(X[3:11], precision=capi.ARH_MIXED)
    assert np.array_equal(Y[3:11], Y8)
    Y4 = h.matvec_np(X[[36, 0, 17, 5]], precision=capi.ARH_MIXED)
    assert np.array_equal(Y[[36, 0, 17, 5]], Y4)


def test_fp32_small_batch_and_transpose(relief, capi):
    g, h, k = relief
    rng = np.random.default_rng(2)
    x = rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)
    assert rel(h.matvec_np(x, precision=capi.ARH_MIXED), h.matvec_np(x)) < 5e-6
    X = rng.standard_normal((6, g.N)) + 1j * rng.standard_normal((6, g.N))
    assert rel(h.matvec_np(X, transpose=True, precision=capi.ARH_MIXED), h.matvec_np(X, transpose=True)) < 5e-6


def test_mixed_solve_refined_and_batch_independent(relief, capi):
    from acoustic_reliefs_amd import meshes
    g, h, k = relief
    dirs = meshes.incidence_directions(37)
    B = -g.rhs(k, capi.ARH_RHS_PLANE, dirs).numpy()
    X, st = h.solve_np(B, tol=1e-6)
    print("mixed: iters %d refinements %d worst true residual %.2e" % (st["iters"], st["refinements"],
                                                                       st["max_true_res"]))
    assert (st["true_res"] <= 1e-6).all()
    # the FP64 operator's own solution
    h64 = g.assemble(k, 1e-4, 100, capi.ARH_FP64)
    X64, st64 = h64.solve_np(B, tol=1e-9, true_residual=True)
    assert max(rel(X[j], X64[j]) for j in range(37)) < 2e-5
    # alone and in the mix: bitwise identical
    Xs, _ = h.solve_np(B[[4, 20, 33, 7]], tol=1e-6)
    assert np.array_equal(Xs, X[[4, 20, 33, 7]])
    X1, _ = h.solve_np(B[12], tol=1e-6)
    assert np.array_equal(X1, X[12])  # a single rhs takes the same MFMA product


def test_band_solve_mixed_vs_fp64(relief, capi):
    from acoustic_reliefs_amd import meshes
    g, _, k = relief
    dirs = meshes.incidence_directions(37)
    L = meshes.semicircle_receivers(5.0, 181)
    d32, _, _, st32, info32 = g.band_solve(k, capi.ARH_RHS_PLANE, dirs, L, storage=capi.ARH_MIXED)
    d64, _, _, st64, info64 = g.band_solve(k, capi.ARH_RHS_PLANE, dirs, L, storage=capi.ARH_FP64)
    print("max |d32 - d64| = %.2e" % np.abs(d32 - d64).max())
    assert np.abs(d32 - d64).max() <= 1e-4
    assert st32["max_true_res"] <= 1e-6 and st64["max_true_res"] <= 1e-6
    assert info32["bytes_fp32"] > 0 and info64["bytes_fp32"] == 0


def test_compact_keeps_results(relief, capi):
    """arh_hmat_compact frees the derived product copies; the rebuilt copies give the same bits."""
    g, h, k = relief
    rng = np.random.default_rng(5)
    X = rng.standard_normal((6, g.N)) + 1j * rng.standard_normal((6, g.N))
    before = [h.matvec_np(X, transpose=t, precision=p) for t in (False, True) for p in (capi.ARH_MIXED, capi.ARH_FP64)]
    b32 = h.info()["bytes_fp32"]
    h.compact()
    assert h.info()["bytes_fp32"] < b32
    after = [h.matvec_np(X, transpose=t, precision=p) for t in (False, True) for p in (capi.ARH_MIXED, capi.ARH_FP64)]
    for a, b in zip(before, after):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("grid", [16, 17])
def test_solver_workspace_every_batch_size(ctx, capi, grid):
    """Regression: the solver's workspace pieces are 256-byte aligned, so its size is the sum of the rounded
    pieces; with a fixed slack some (N, nrhs) put the per-batch active counts past the allocation
    (N = 818, one rhs, failed with an invalid copy). Every batch size 1..40 on two small reliefs, both
    precisions, must solve to the true-residual tolerance."""
    from acoustic_reliefs_amd import meshes
    P, E, _ = meshes.relief_config(0.3, grid, 0.05, seed=1000 + grid)
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    k = 2 * np.pi * 1250.0 / 343.0
    L = meshes.semicircle_receivers(5.0, 181)
    fan = meshes.incidence_directions(40)
    print("workspace sizes: N=%d" % g.N, flush=True)
    for storage in (capi.ARH_FP64, capi.ARH_MIXED):
        for R in range(1, 41):
            d, _, _, st, _ = g.band_solve(k, capi.ARH_RHS_PLANE, fan[:R], L, storage=storage, tol=1e-6)
            assert st["max_true_res"] <= 1e-6 and np.all(np.isfinite(d)), (storage, R)
