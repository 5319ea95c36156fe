"""arh_receivers_noalloc (receiver pressures without device allocation) against arh_receivers, through the
C-ABI: forward (y = Q x) and transposed (Q^T vf), two calls in a row with the second smaller, and the size
limit. Tolerance 1e-13 relative: both evaluate the same entries (the 6-point rule of integrate.cpp:5-14 at
each receiver, integrate.cpp:46-57) and sum them in the same order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_receivers_noalloc_matches_receivers(ctx, capi):
    from acoustic_reliefs_amd import meshes
    P, E, _ = meshes.relief_config(0.3, 24, 0.05, seed=5)
    g = capi.Geometry.from_mesh(ctx, P, E, leaf=32, factor=1.5)
    k = 2 * np.pi * 1250.0 / 343.0
    rng = np.random.default_rng(3)
    L_big = meshes.semicircle_receivers(5.0, 181)
    L_small = np.vstack([meshes.semicircle_receivers(3.0, 37)[::2], [[0.1, 2.0, 0.3]]])
    for L, nrhs in ((L_big, 9), (L_small, 3)):  # the second call is smaller in LL and nrhs
        LL = len(L)
        X = rng.standard_normal((nrhs, g.N)) + 1j * rng.standard_normal((nrhs, g.N))
        Xd = ctx.to_device(X)
        want = g.receivers(k, L, Xd, nrhs=nrhs).numpy()
        Y = ctx.empty((nrhs, LL))
        g.receivers_noalloc(k, L, Xd, Y, nrhs=nrhs)
        got = Y.numpy()
        print("LL=%d nrhs=%d forward: rel diff %.2e, bit-identical %s" % (LL, nrhs, _rel(got, want),
                                                                          np.array_equal(got, want)), flush=True)
        assert _rel(got, want) <= 1e-13
        VF = rng.standard_normal((nrhs, LL)) + 1j * rng.standard_normal((nrhs, LL))
        VFd = ctx.to_device(VF)
        want_t = g.receivers(k, L, VFd, nrhs=nrhs, transpose=True).numpy()
        Yt = ctx.empty((nrhs, g.N))
        g.receivers_noalloc(k, L, VFd, Yt, nrhs=nrhs, transpose=True)
        got_t = Yt.numpy()
        print("LL=%d nrhs=%d transposed: rel diff %.2e, bit-identical %s" % (
            LL, nrhs, _rel(got_t, want_t), np.array_equal(got_t, want_t)), flush=True)
        assert _rel(got_t, want_t) <= 1e-13
    # more receivers than the static array holds: a clear error, nothing launched
    Lmax = np.zeros((16385, 3))
    Lmax[:, 1] = 10.0
    Y = ctx.empty((1, 16385))
    with pytest.raises(RuntimeError, match="at most 16384 receivers"):
        g.receivers_noalloc(k, Lmax, ctx.to_device(np.ones((1, g.N), np.complex128)), Y)
