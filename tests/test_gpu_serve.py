"""Serving: requests admitted as workers and HBM free up, overlapping on the device, each result
bit-identical to the same request run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_same_alone_and_in_the_mix(capi):
    from acoustic_reliefs_amd import meshes
    L = meshes.semicircle_receivers(5.0, 181)
    dirs = meshes.incidence_directions(37)
    srv = capi.Server(workers=4)
    bodies = [meshes.relief_config(0.3, 16, 0.05, seed=s)[:2] for s in (1, 2)] + [meshes.relief_config(0.6, 32, 0.1, seed=3)[:2]]
    ids = [srv.add_mesh(P, E) for P, E in bodies]
    k = 2 * np.pi * 800.0 / 343.0
    # alone
    rid = srv.submit(ids[2], k, dirs[:9], L)
    alone, info = srv.wait(rid)
    assert info["status"] == 2 and info["max_true_res"] <= 1e-6
    # in a mix of uneven meshes, bands and batch sizes, submitted together
    rng = np.random.default_rng(0)
    mix = []
    for q in range(10):
        m = int(rng.integers(0, 3))
        f = float(rng.choice(meshes.THIRD_OCTAVE_BANDS[:8]))
        nr = int(rng.integers(1, 38))
        mix.append(srv.submit(ids[m], 2 * np.pi * f / 343.0, dirs[:nr], L))
        if q == 4:
            target = srv.submit(ids[2], k, dirs[:9], L)
    res = {r: srv.wait(r) for r in mix + [target]}
    d_mix, info_mix = res[target]
    assert np.array_equal(d_mix, alone)
    for r, (d, inf) in res.items():
        assert inf["status"] == 2 and inf["max_true_res"] <= 1e-6 and inf["t_done"] >= inf["t_start"] >= inf["t_submit"]
    # the device ran requests concurrently: some start before another one finished
    spans = sorted((inf["t_start"], inf["t_done"]) for _, inf in res.values())
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    srv.close()


def test_admission_by_footprint(capi):
    """With a budget that fits one request at a time, requests never overlap and still complete."""
    from acoustic_reliefs_amd import meshes
    L = meshes.semicircle_receivers(5.0, 61)
    dirs = meshes.incidence_directions(5)
    P, E, _ = meshes.relief_config(0.3, 16, 0.05, seed=1)
    srv = capi.Server(workers=4, hbm_budget=1)  # every request exceeds it: each runs alone
    mid = srv.add_mesh(P, E)
    rids = [srv.submit(mid, 2 * np.pi * f / 343.0, dirs, L) for f in (250.0, 500.0, 1000.0, 2000.0)]
    infos = [srv.wait(r)[1] for r in rids]
    spans = sorted((i["t_start"], i["t_done"]) for i in infos)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))
    assert all(i["footprint"] > 1 for i in infos)
    srv.close()
