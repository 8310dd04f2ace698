"""Synthetic workloads: random streams, Kuhn cube, gmsh round trip."""
import json
import os

import numpy as np
import pytest

from morganapublic_amd import workloads as W

HERE = os.path.dirname(__file__)


def test_rng_streams_fixture():
    ref = json.load(open(os.path.join(HERE, "golden", "rng_streams.json")))
    for sid, d in ref.items():
        sid = int(sid)
        assert [int(v) for v in W.stream_raw(0, sid, 16)] == d["raw"]
        np.testing.assert_array_equal(W.stream_uniform(0, sid, 16), d["uniform"])
        np.testing.assert_allclose(W.stream_normal(0, sid, 8), d["normal"], rtol=0, atol=1e-15)


def test_rng_counter_based():
    a = W.stream_uniform(3, 2, 100)
    b = W.stream_uniform(3, 2, 40, start=60)
    np.testing.assert_array_equal(a[60:], b)
    u = W.stream_uniform(0, 1, 200000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 5e-3
    z = W.stream_normal(0, 2, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02


@pytest.mark.parametrize("n,jit", [(1, 0.0), (3, 0.0), (4, 0.2), (12, 0.2)])
def test_kuhn_cube(n, jit):
    m = W.kuhn_cube(n, jitter=jit, seed=1)
    c = W.kuhn_counts(n)
    assert m.nT == c["T"] and m.nV == c["V"]
    P = m.xyz[m.tets - 1]
    det = np.linalg.det(np.stack([P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]], axis=2))
    assert det.min() > 0
    assert abs(det.sum() / 6 - 1.0) < 1e-12  # the tets tile the unit cube
    assert m.btri.shape[0] == 12 * n * n
    # boundary triangles are faces of tets
    faces = set()
    for t in m.tets:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            faces.add(tuple(sorted(t[list(f)])))
    assert all(tuple(sorted(b)) in faces for b in m.btri)
    # geoIds match the face they lie on
    for g in range(1, 7):
        ax, side = (g - 1) // 2, (g - 1) % 2
        pts = m.xyz[m.btri[m.bgeo == g] - 1][..., ax]
        assert np.all(np.abs(pts - side) < 1e-14)


def test_gmsh_round_trip(tmp_path):
    m = W.kuhn_cube(3, jitter=0.2)
    p = str(tmp_path / "cube.msh")
    W.write_gmsh22(p, m)
    r = W.read_gmsh22(p)
    np.testing.assert_array_equal(r.tets, m.tets)
    np.testing.assert_array_equal(r.btri, m.btri)
    np.testing.assert_array_equal(r.bgeo, m.bgeo)
    np.testing.assert_array_equal(r.xyz, m.xyz)
