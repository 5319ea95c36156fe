"""Host-side logic without a GPU: the serving-stream generator and the one-at-a-time latency model of
bench.py, the measured-traffic lookup, the mesher's relief set-up, and argument errors of the entry points
added for the listener ACA, the relief parameterisation and the device memory cache."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def test_fifo_latencies_closed_form():
    import bench
    # one server, arrivals 0, 1, 2; service 2, 2, 0.5: completions 2, 4, 4.5
    lat = bench.fifo_latencies([0.0, 1.0, 2.0], [2.0, 2.0, 0.5])
    assert np.allclose(lat, [2.0, 3.0, 2.5])
    # idle server: latency = service
    assert np.allclose(bench.fifo_latencies([0.0, 10.0], [1.0, 1.0]), [1.0, 1.0])


def test_stream_is_seeded_and_in_range():
    import bench
    r1, q1 = bench.make_stream(60, 4, seed=0, rate=5.0)
    r2, q2 = bench.make_stream(60, 4, seed=0, rate=5.0)
    assert [r["grid"] for r in r1] == [r["grid"] for r in r2]
    assert all(np.array_equal(a["dirs"], b["dirs"]) and a["t"] == b["t"] for a, b in zip(q1, q2))
    assert all(0.3 <= r["size"] <= 0.9 and 16 <= r["grid"] <= 160 for r in r1)
    t = np.array([q["t"] for q in q1])
    assert np.all(np.diff(t) > 0)
    assert all(1 <= len(q["dirs"]) <= 37 for q in q1)
    from acoustic_reliefs_amd import meshes
    assert all(q["band"] in meshes.THIRD_OCTAVE_BANDS for q in q1)
    # the first requests do not depend on how many follow
    _, q3 = bench.make_stream(10, 4, seed=0, rate=5.0)
    assert all(a["t"] == b["t"] and a["relief"] == b["relief"] for a, b in zip(q3, q1))


def test_measured_traffic_lookup():
    import bench
    t = bench.measured_traffic()
    assert t is None or 1e8 < t < 1e9
    assert bench.measured_traffic("no such kernel") is None


def test_relief_setup_reference_counts():
    from acoustic_reliefs_amd import mesher
    # configs/base.toml: generate_base_mesh(w, h, b, 30, 7, 30), subdiv_top 1
    P, E, diff, uvs = mesher.relief_setup(0.6, 0.6, 0.15, subdiv_top=1)
    P0, E0 = mesher.generate_base_mesh(0.6, 0.15, 0.6, 30, 7, 30)
    # 1800 top triangles become 7200; their 31 x 31 grid gains 30 x 31 x 2 edge and 30 x 30 diagonal
    # midpoints. The side faces keep their own copies of the top edge's vertices (the box is not welded),
    # so none of them is split, and those copies stay differentiable points (y = h)
    assert len(E) == len(E0) + 3 * 1800
    assert len(P) == len(P0) + 2 * 30 * 31 + 30 * 30
    assert len(diff) == 61 * 61 + 2 * 31 + 2 * 31


def test_new_entry_points_reject_bad_arguments(capi):
    lib = capi.lib()
    out = capi.c_void_p()
    uv = np.zeros(4)
    assert lib.arh_relief_create(None, capi._dp(uv), 2, 4, 4, capi.ctypes.byref(out)) != 0
    L = np.zeros(6)
    assert lib.arh_qaca_create(None, capi.c_double(1.0), capi._dp(L), 2, capi.c_double(1e-3), 10,
                               capi.ctypes.byref(out)) != 0
    assert lib.arh_qaca_apply(None, 0, 1, None, None) != 0
    assert lib.arh_relief_eval(None, None, None) != 0
    b = capi.c_size_t()
    assert lib.arh_device_cached(0, capi.ctypes.byref(b)) == 0 and b.value == 0
    assert lib.arh_device_cached(0, None) != 0


def _line(xs):
    return np.column_stack([np.asarray(xs, dtype=np.float64), np.zeros(len(xs)), np.zeros(len(xs))])


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_partition_by_hand(impl, capi):
    """Cluster tree + dual traversal (clustertree.cpp:60-137) traced by hand on centroids on the x axis,
    input order x = (3, 10, 0, 2, 1), leaf 2, factor 1.5:

    root [0,5): mean 3.2, radius 6.8; x < 3.2 left in input order: ids (0, 2, 3, 4), right (1).
    L [0,4): mean 1.5, radius 1.5 -> LL = ids (2, 4), LR = ids (0, 3); R [4,5): id 1, radius 0.
    perm = (2, 4, 0, 3, 1). Traversal from (root, root): equal sizes split the target; a leaf target
    splits the source; admissible when dist > 1.5 (r_s + r_t):
      (LL,LL) direct; (LR,LL) dist 2 > 1.5 approx; (LL,LR) approx; (LR,LR) direct; (R,L) dist 8.5 > 2.25
      approx; (L,R) approx; (R,R) direct.
    The product's host partition (arh_partition) and the oracle must both give exactly this."""
    if impl == "oracle":
        from oracle import oracle as O
        perm, d, a, _ = O.partition(_line([3, 10, 0, 2, 1]), 2, 1.5)
    else:
        perm, d, a = capi.partition(_line([3, 10, 0, 2, 1]), 2, 1.5)
    assert perm.tolist() == [2, 4, 0, 3, 1]
    assert d.tolist() == [[0, 2, 0, 2], [2, 4, 2, 4], [4, 5, 4, 5]]
    assert a.tolist() == [[2, 4, 0, 2], [0, 2, 2, 4], [4, 5, 0, 4], [0, 4, 4, 5]]
    # a centroid exactly at the mean goes right ('<' goes left): x = (0, 1, 2), leaf 2 -> {0} | {1, 2}
    if impl == "oracle":
        perm, d, a, _ = O.partition(_line([2, 1, 0]), 2, 1.5)
    else:
        perm, d, a = capi.partition(_line([2, 1, 0]), 2, 1.5)
    assert perm.tolist() == [2, 0, 1]
    # (root,root) splits the target; (root,{0}) splits the source: ({0},{0}) direct, ({1,2},{0}) centres
    # 1.5 apart > 1.5 (0.5 + 0) = 0.75: approx; (root,{1,2}) splits the source: ({0},{1,2}) approx,
    # ({1,2},{1,2}) direct
    assert d.tolist() == [[0, 1, 0, 1], [1, 3, 1, 3]] and a.tolist() == [[1, 3, 0, 1], [0, 1, 1, 3]]
