"""The C++ Force/Body layer (atac._redmax) without a GPU: construction, the reference's
.sdf text format round trip (BodySDFObj.cpp:86-116) and the error paths."""
import os

import numpy as np
import pytest

from atac import workloads as W


def _redmax():
    from atac import _redmax
    return _redmax


def test_sdf_text_round_trip(tmp_path):
    R = _redmax()
    g = W.peg_in_hole(n_points=10, n_poses=1).grids[0]
    pts = np.zeros((4, 3))
    b = R.BodySDFObj("hole", g.values.astype(np.float64), g.min_box, g.max_box, g.dxyz, 0.01, pts, 6)
    path = os.path.join(tmp_path, "hole.sdf")
    b.save_sdf(path)
    with open(path) as f:
        head = [f.readline().split() for _ in range(4)]
    assert [int(v) for v in head[0]] == list(g.n)
    assert np.allclose([float(v) for v in head[1]], g.min_box)
    assert np.allclose([float(v) for v in head[3]], g.dxyz)
    b2 = R.BodySDFObj.from_sdf_file("hole2", path, 0.02, pts, 6)
    assert b2.shape == list(g.n)
    # the reference writes values with the stream's default 6 significant digits
    assert np.allclose(b2.values(), b.values(), rtol=1e-5, atol=1e-6)
    assert b2.col_th == 0.02 and b2.index == [6, 7, 8, 9, 10, 11]


def test_missing_file_and_bad_grid():
    R = _redmax()
    with pytest.raises(RuntimeError):
        R.BodySDFObj.from_sdf_file("x", "/nonexistent/file.sdf", 0.0, np.zeros((1, 3)), 0)
    with pytest.raises(ValueError):
        R.Body("x", np.zeros((3, 2)), 0)


def test_force_requires_sdf_body():
    R = _redmax()
    a = R.Body("a", np.zeros((2, 3)), 0)
    c = R.Body("c", np.zeros((2, 3)), 6)
    with pytest.raises(Exception):
        R.ForceGeneralSDFContact(a, c, 1e4, 0.0, 0.0, 0.0)


def test_pose_and_twist_properties():
    R = _redmax()
    a = R.Body("a", np.zeros((2, 3)), 0)
    E = W.pose(W.rotation_exp([0.1, 0.2, 0.3]), [1.0, 2.0, 3.0])
    a.E_0i = E
    a.phi = np.arange(6.0)
    assert np.allclose(a.E_0i, E) and np.allclose(a.phi.ravel(), np.arange(6.0))
