"""BodyMeshObj::precompute_contact_points (simulation/redmax/Body/BodyMeshObj.cpp:184-238)
restated in the host layer: without adaptive sampling the points are the vertices;
with it, the vertices the C library's rand() keeps after srand(1000) (the same
generator, called through ctypes here), then on each face the same draw keeps and
whose area exceeds 0.5, floor(area / 0.5) points of the R2 sequence folded into the
triangle (each inside its triangle, with barycentric weights from the sequence)."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from atac import workloads as W


@pytest.fixture(scope="module")
def R():
    from atac import _redmax
    return _redmax


def _big_mesh():
    # a 3,000-vertex sphere scaled so that its faces straddle the 0.5 area threshold
    V, F = W.icosphere(1.0, levels=4)
    V = V * np.array([22.0, 17.0, 13.0])
    return V, F


def test_vertices_without_adaptive_sampling(R):
    V, F = _big_mesh()
    P, rate = R.precompute_contact_points(V, F, False)
    assert rate == 1
    np.testing.assert_array_equal(P, V)


def test_adaptive_sampling_uses_the_c_library_sequence(R):
    V, F = _big_mesh()
    P, rate = R.precompute_contact_points(V, F, True)
    assert rate == max(int(np.floor(len(V) / 1000.0)), 1) == 2
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(1000)
    keep_v = [i for i in range(len(V)) if libc.rand() % rate == 0]
    keep_f = [t for t in range(len(F)) if libc.rand() % rate == 0]
    nv = len(keep_v)
    np.testing.assert_array_equal(P[:nv], V[keep_v])
    # face samples: count per kept face from Heron's area, each point inside its face
    rest = P[nv:]
    pos = 0
    n_faces_sampled = 0
    for t in keep_f:
        A, B, C = V[F[t]]
        a, b, c = np.linalg.norm(B - C), np.linalg.norm(C - A), np.linalg.norm(A - B)
        s = 0.5 * (a + b + c)
        area = np.sqrt(s * (s - a) * (s - b) * (s - c))
        if not area > 0.5:
            continue
        n = int(np.floor(area / 0.5))
        Q = rest[pos:pos + n]
        pos += n
        n_faces_sampled += 1
        # barycentric coordinates of the samples: inside the triangle
        M = np.stack([B - A, C - A], axis=1)
        uv, *_ = np.linalg.lstsq(M, (Q - A).T, rcond=None)
        assert np.all(uv > -1e-9) and np.all(uv.sum(0) < 1 + 1e-9)
        # the first sample is the R2 point (0.5 + 1/g, 0.5 + 1/g^2) mod 1 folded
        g = np.float32(1.32471795724474602596)
        z1 = np.fmod(np.float32(0.5) + np.float32(1.0) / g, np.float32(1.0))
        z2 = np.fmod(np.float32(0.5) + (np.float32(1.0) / g) ** 2, np.float32(1.0))
        if z1 + z2 > 1:
            z1, z2 = np.float32(1) - z1, np.float32(1) - z2
        np.testing.assert_allclose(uv[:, 0], [z1, z2], atol=1e-6)
    assert pos == len(rest) and n_faces_sampled > 100
