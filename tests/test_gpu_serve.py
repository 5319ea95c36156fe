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


def test_cancel_queued_requests(capi):
    """arh_server_cancel drops what is queued (ARH_ECANCEL from wait) and lets the running request
    finish; close() right after submitting returns once the running work is done."""
    from acoustic_reliefs_amd import meshes
    L = meshes.semicircle_receivers(5.0, 61)
    dirs = meshes.incidence_directions(5)
    P, E, _ = meshes.relief_config(0.3, 16, 0.05, seed=1)
    srv = capi.Server(workers=1)
    mid = srv.add_mesh(P, E)
    rids = [srv.submit(mid, 2 * np.pi * 500.0 / 343.0, dirs, L) for _ in range(6)]
    n = srv.cancel()
    assert 4 <= n <= 6  # the worker may have admitted the first one (or two) already
    done, cancelled = 0, 0
    for r in rids:
        try:
            d, info = srv.wait(r)
            assert info["status"] == 2 and np.all(np.isfinite(d))
            done += 1
        except RuntimeError as e:
            assert "ECANCEL" in str(e)
            assert srv.poll(r)["status"] == capi.ARH_ECANCEL
            cancelled += 1
    assert cancelled == n and done + cancelled == 6
    # new requests are still served after a cancel
    d, info = srv.wait(srv.submit(mid, 2 * np.pi * 500.0 / 343.0, dirs, L))
    assert info["status"] == 2
    # destroy with requests queued and one running: returns, queued ones cancelled
    for _ in range(5):
        srv.submit(mid, 2 * np.pi * 1000.0 / 343.0, dirs, L)
    srv.close()
    assert srv.h is None


def test_exit_with_requests_in_flight(tmp_path):
    """A process that exits (normally or by an exception) while requests are queued and running ends
    cleanly: the exit hook destroys the live server before the HIP runtime's teardown."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "child.py"
    script.write_text(
        "import sys\n"
        "sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from acoustic_reliefs_amd import capi, meshes\n"
        "L = meshes.semicircle_receivers(5.0, 61)\n"
        "dirs = meshes.incidence_directions(9)\n"
        "P, E, _ = meshes.relief_config(0.3, 16, 0.05, seed=1)\n"
        "srv = capi.Server(workers=2)\n"
        "mid = srv.add_mesh(P, E)\n"
        "ids = [srv.submit(mid, 2 * np.pi * f / 343.0, dirs, L) for f in (250.0, 500.0, 1000.0, 2000.0) * 3]\n"
        "print('submitted', len(ids), flush=True)\n"
        "if sys.argv[1] == 'raise':\n"
        "    raise RuntimeError('caller failed with requests in flight')\n" % root)
    for mode, want in (("exit", 0), ("raise", 1)):
        p = subprocess.run([sys.executable, str(script), mode], capture_output=True, text=True, timeout=120)
        assert p.returncode == want, (mode, p.returncode, p.stdout[-2000:], p.stderr[-2000:])
        assert "submitted 12" in p.stdout
