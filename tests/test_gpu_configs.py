"""The gradient, large-relief and serving workloads on the GPU path.

* Adjoint gradient on the benchmark relief (0.6 m, 64 x 64 top, N = 9,828, all 4,096 top vertices
  differentiable, 13 bands x 37 plane waves, 181 receivers): loss = mean over bands of (1 - d_b) with d_b the
  band's mean over the incidences, backward through the torch.autograd.Function. Checked (a) against central
  differences (h = 1e-5 m) on 16 top vertices drawn with default_rng(1), in FP64 and in mixed precision, on the
  exact operator (every leaf pair dense: ACA pivots move discontinuously with the vertices), with the stated
  bound |fd - adjoint| <= 2e-3 max|fd| + 1e-6 (solver tolerance 1e-10 -> FD noise ~1e-10 / 1e-5); (b) against
  the oracle's gradient (tests/golden/relief64_grad_golden.npz) on the 1000 Hz band for three plane waves, with
  the H-matrix at ACA 1e-8: relative L2 <= 1e-3 (ACA pivots at near-ties), FP64 and mixed.
* The 0.9 m relief with a 160 x 160 top (N = 58,724), one low band (250 Hz), 37 plane waves, mixed precision,
  with stored H-matrices and with recompute-on-backward: identical values and gradients bit for bit, and the
  device memory each keeps between forward and backward.
* A seeded slice of the serving stream: the first 50 requests of the 2000-request stream (bench.make_stream,
  seed 0, 64 reliefs; it holds reliefs of 54,468 to 57,218 elements), submitted together to a 4-worker server:
  every request's diffusion coefficients equal its solo run bit for bit and every true residual is <= 1e-6.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C0 = 343.0


@pytest.fixture(scope="module")
def relief64():
    from acoustic_reliefs_amd import meshes
    return meshes.relief_config(0.6, 64, 0.1, seed=0)


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
def test_gradient_vs_central_differences(relief64, ctx, capi, storage):
    import torch
    from acoustic_reliefs_amd import autograd, meshes
    P, E, top = relief64
    st_id = capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED
    eng = autograd.ReliefBEM(P, E, top, meshes.THIRD_OCTAVE_BANDS, meshes.incidence_directions(37),
                             meshes.semicircle_receivers(5.0, 181), c0=C0, storage=st_id, tol=1e-10, ctx=ctx,
                             all_dense=True)
    h = torch.zeros(len(top), dtype=torch.float64, requires_grad=True)
    (1.0 - autograd.diffusion_coefficients(h, eng)).mean().backward()
    g = h.grad.numpy()
    pick = np.random.default_rng(1).choice(len(top), 16, replace=False)
    step = 1e-5
    fds = []
    for t in pick:
        hp, hm = np.zeros(len(top)), np.zeros(len(top))
        hp[t], hm[t] = step, -step
        fds.append((np.mean(1.0 - eng.forward(hp)[0]) - np.mean(1.0 - eng.forward(hm)[0])) / (2 * step))
    fds, ga = np.array(fds), g[pick]
    err = np.abs(fds - ga)
    print("%s: max |fd| %.3e, max |fd - adjoint| %.3e (%.2e of max |fd|), median rel %.2e" % (
        storage, np.abs(fds).max(), err.max(), err.max() / np.abs(fds).max(),
        np.median(err / np.maximum(np.abs(fds), 1e-300))))
    assert np.all(err <= 2e-3 * np.abs(fds).max() + 1e-6), (fds, ga)


@pytest.mark.parametrize("storage", ["fp64", "mixed"])
def test_gradient_vs_oracle_one_band(relief64, ctx, capi, storage):
    from acoustic_reliefs_amd import autograd, meshes
    z = np.load(os.path.join(GOLDEN, "relief64_grad_golden.npz"))
    P, E, top = relief64
    assert np.array_equal(z["top"], top)
    st_id = capi.ARH_FP64 if storage == "fp64" else capi.ARH_MIXED
    eng = autograd.ReliefBEM(P, E, top, [float(z["frequency"])], meshes.incidence_directions(37)[z["pick"]],
                             meshes.semicircle_receivers(5.0, 181), c0=float(z["c0"]), aca_tol=float(z["aca_tol"]),
                             tol=float(z["bicgstab_tol"]), storage=st_id, ctx=ctx)
    d_b, d_all = eng.forward(np.zeros(len(top)))
    g = eng.backward(np.array([-1.0]))  # loss = mean over the directions of (1 - d)
    rel = np.linalg.norm(g - z["grad"]) / np.linalg.norm(z["grad"])
    print("%s: |d - d_oracle| %.2e, gradient relative L2 %.2e" % (storage, np.abs(d_all[0] - z["d"]).max(), rel))
    assert np.abs(d_all[0] - z["d"]).max() <= 1e-6
    assert rel <= 1e-3


def _held_bytes(state):
    """Device bytes a forward's state keeps for its backward: kept H-matrices (after compaction) and the
    per-band solutions and adjoint weights."""
    n = 0
    for s in state["bands"]:
        if s["H"] is not None:
            n += s["H"].info()["bytes_total"]
        n += s["X"].nbytes + s["VF"].nbytes
    return n


def test_large_relief_stored_vs_recompute(ctx, capi):
    import torch
    from acoustic_reliefs_amd import autograd, meshes
    P, E, top = meshes.relief_config(size=0.9, grid=160, hmax=0.15, seed=0)
    assert len(E) == 58724
    res = {}
    for recompute in (False, True):
        eng = autograd.ReliefBEM(P, E, top, [250.0], meshes.incidence_directions(37), meshes.semicircle_receivers(5.0, 181),
                                 c0=C0, storage=capi.ARH_MIXED, recompute=recompute, ctx=ctx)
        h = torch.full((len(top),), 0.01, dtype=torch.float64, requires_grad=True)
        d = autograd.diffusion_coefficients(h, eng)
        held = _held_bytes(eng.last_state)
        (1.0 - d).mean().backward()
        res[recompute] = (d.detach().numpy().copy(), h.grad.numpy().copy(), held)
        print("0.9 m relief, 250 Hz, %s: %.3f GB held between forward and backward, %s" % (
            "recompute" if recompute else "stored", held / 1e9, eng.timing[0]))
        del eng, d, h
    assert np.array_equal(res[False][0], res[True][0]) and np.array_equal(res[False][1], res[True][1])
    assert res[True][2] < 0.1 * res[False][2]  # recompute keeps the H-matrix out of HBM between the passes


def test_serving_stream_slice_same_as_solo(capi):
    import bench
    from acoustic_reliefs_amd import meshes
    reliefs, reqs = bench.make_stream(50, 64, seed=0, rate=5.0)
    assert max(len(reliefs[q["relief"]]["E"]) for q in reqs) >= 50000
    L = meshes.semicircle_receivers(5.0, 181)
    srv = capi.Server(workers=4, maxit=10000)
    mids = {}
    for q in reqs:
        if q["relief"] not in mids:
            r = reliefs[q["relief"]]
            mids[q["relief"]] = srv.add_mesh(r["P"], r["E"])
    rids = [srv.submit(mids[q["relief"]], 2 * np.pi * q["band"] / C0, q["dirs"], L) for q in reqs]
    mixed = [srv.wait(r) for r in rids]
    srv.close()
    solo_srv = capi.Server(workers=1, maxit=10000)
    smids = {k: solo_srv.add_mesh(reliefs[k]["P"], reliefs[k]["E"]) for k in mids}
    worst = 0.0
    for i, q in enumerate(reqs):
        d_solo, info_solo = solo_srv.wait(solo_srv.submit(smids[q["relief"]], 2 * np.pi * q["band"] / C0, q["dirs"], L))
        d_mix, info_mix = mixed[i]
        worst = max(worst, info_mix["max_true_res"])
        assert info_mix["status"] == 2 and info_mix["max_true_res"] <= 1e-6, (i, info_mix)
        assert np.array_equal(d_mix, d_solo), (i, d_mix, d_solo)
    solo_srv.close()
    print("50 requests (%d solves): every result equals its solo run; worst true residual %.2e" % (
        sum(len(q["dirs"]) for q in reqs), worst))
