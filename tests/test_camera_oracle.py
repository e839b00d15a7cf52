"""The oracle's per-view constant buffer (the reference's camera set-up and update(),
main.c:1422-1562) against the mathematics (CPU)."""
from __future__ import annotations

import numpy as np

PI, TAU, HALF_PI = np.float32(3.141592654), np.float32(6.283185307), np.float32(1.570796326)


def test_det_and_inverse_tables(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        m = rng.normal(size=(4, 4)).astype(np.float32)
        m64 = m.astype(np.float64)
        d = oracle.det4(m)
        assert abs(d - np.linalg.det(m64)) <= 1e-4 * max(1.0, abs(np.linalg.det(m64)))
        inv = oracle.inv4(m).astype(np.float64)
        np.testing.assert_allclose(inv @ m64, np.eye(4), atol=2e-4)
    # a permutation (every term's sign matters) and a diagonal
    p = np.eye(4, dtype=np.float32)[[2, 0, 3, 1]]
    np.testing.assert_array_equal(oracle.inv4(p), p.T)
    np.testing.assert_allclose(oracle.inv4(np.diag([2, 4, 8, 16]).astype(np.float32)),
                               np.diag([0.5, 0.25, 0.125, 0.0625]))


def _mats(cb):
    return cb[:16].reshape(4, 4).astype(np.float64), cb[16:32].reshape(4, 4).astype(np.float64), \
        cb[32:].reshape(4, 4).astype(np.float64)


def test_default_camera_looks_down_minus_x(oracle):
    # the reference's start-up camera (main.c:1423-1430) on a 1920x1080 target
    pos, near = (3.5, 1.0, 1.0), 0.01
    cb = oracle.camera_cb(pos, 0.0, 0.0, 75.0, near, 1920, 1080)
    clip_from_world, view_from_clip, world_from_view = _mats(cb)
    sy = 1.0 / np.tan(np.radians(75.0) / 2.0)
    # view z maps to world -x (change of basis, main.c:1461-1466); the position is the translation
    np.testing.assert_allclose(world_from_view[:3, 2], (-1.0, 0.0, 0.0), atol=1e-7)
    np.testing.assert_allclose(world_from_view[:3, 3], pos)
    for d in (0.5, 3.0, 40.0):
        c = clip_from_world @ np.array([pos[0] - d, pos[1], pos[2], 1.0])
        ndc = c[:3] / c[3]
        np.testing.assert_allclose(ndc, (0.0, 0.0, near / d), atol=1e-6)
    # a point up and to the side lands where the projection puts it
    c = clip_from_world @ np.array([pos[0] - 2.0, pos[1] - 0.5, pos[2] + 0.25, 1.0])
    # view +x is world +y and view +y is world +z (the change of basis)
    np.testing.assert_allclose(c[:2] / c[3], (-0.5 / 2.0 * sy / (1920 / 1080), 0.25 / 2.0 * sy), rtol=1e-5)
    proj = np.linalg.inv(view_from_clip)
    np.testing.assert_allclose(proj[:2, :2], np.diag([sy / (1920 / 1080), sy]), rtol=1e-6)


def test_update_wraps_yaw_and_clamps_pitch(oracle):
    a = oracle.camera_cb((1.0, 2.0, 3.0), np.float32(4.0), 0.3, 60.0, 0.05, 640, 360)
    b = oracle.camera_cb((1.0, 2.0, 3.0), np.float32(4.0) - TAU, 0.3, 60.0, 0.05, 640, 360)
    np.testing.assert_array_equal(a, b)
    c = oracle.camera_cb((1.0, 2.0, 3.0), 0.7, 2.5, 60.0, 0.05, 640, 360)
    d = oracle.camera_cb((1.0, 2.0, 3.0), 0.7, HALF_PI, 60.0, 0.05, 640, 360)
    np.testing.assert_array_equal(c, d)


def test_matrices_are_consistent(oracle):
    rng = np.random.default_rng(5)
    for _ in range(20):
        pos = rng.uniform(-20, 20, 3)
        fov, near = rng.uniform(30, 100), 0.01
        w, h = int(rng.integers(64, 2048)), int(rng.integers(64, 2048))
        cb = oracle.camera_cb(tuple(pos), rng.uniform(-3, 3), rng.uniform(-1.5, 1.5), fov, near, w, h)
        clip_from_world, view_from_clip, world_from_view = _mats(cb)
        r = world_from_view[:3, :3]
        np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-6)  # a rotation
        sy = 1.0 / np.tan(np.radians(fov) / 2.0)
        proj = np.array([[sy * h / w, 0, 0, 0], [0, sy, 0, 0], [0, 0, 0, near], [0, 0, 1, 0]])
        np.testing.assert_allclose(view_from_clip, np.linalg.inv(proj), rtol=1e-5, atol=1e-6)
        want = proj @ np.linalg.inv(world_from_view)
        np.testing.assert_allclose(clip_from_world, want, atol=1e-5 * np.abs(want).max())
