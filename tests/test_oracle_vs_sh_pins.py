"""CPU pins of the oracle's vertex shaders and SH basis against the mathematics.

The GPU vertex stage is checked against orc_vertex_shader, and the GPU SH light against the
oracle's basis; both sides were written from the same reading of the reference, so these tests
fix them independently:
  - basic_vs / vertex_lighting_vs (basic_vs.c:22-33, vertex_lighting_vs.c:22-39): clip =
    clip_from_world * (p, 1) with clip_from_world the first matrix of the constant buffer
    (basic_vs.c:16-20), row-major (math.h:174-177), each row a pairwise dot (math.h:142-146);
  - fullscreen_vs (fullscreen_vs.c:22-39): dir = world_from_view * (view_from_clip * pos).xyz0
    with the second and third matrices of the buffer;
  - passthrough_vs (passthrough_vs.c:16-25): xy*2-1, the rest copied.
Each continuous result is compared with an FP64 matrix product within a rigorous bound for the
f32 evaluation (a transposed matrix or a swapped constant-buffer slot misses it by orders of
magnitude, which the tests also show); the summation order is fixed by a known answer where the
pairwise and sequential orders differ.
"""
import math

import numpy as np
import pytest

U = 2.0 ** -24  # unit roundoff of f32


def _draw(ob, vertices, vs, cb, vs_tex=None):
    V = vertices.shape[0]
    idx = np.arange(V - V % 3, dtype=np.uint32)
    return ob.Draw(idx, vertices, vs, ob.PS_PASSTHROUGH, cb=cb, vs_tex=vs_tex, viewport=(0, 0, 64, 64, 0, 1))


def _rand_cb(rng):
    cb = rng.uniform(-2.0, 2.0, 48).astype(np.float32)
    M = cb[:16].reshape(4, 4)
    assert np.abs(M - M.T).max() > 0.5  # not symmetric: a transposition is visible
    return cb


def _mat_vec_bound(M, v):
    """Error bound of an f32 4-term pairwise dot per row: two levels of additions after the
    products, so |err| <= 3u * sum|m_k v_k| (plus u^2 terms); 4u leaves margin."""
    return 4.0 * U * (np.abs(M) @ np.abs(v)) + 1e-30


def test_basic_vs_position_is_clip_from_world_times_position(oracle):
    ob = oracle
    rng = np.random.default_rng(11)
    cb = _rand_cb(rng)
    verts = rng.uniform(-3, 3, (300, 8)).astype(np.float32)
    d = _draw(ob, verts, ob.VS_BASIC, cb)
    M = cb[:16].astype(np.float64).reshape(4, 4)  # row-major: clip[r] = sum_c M[r][c] v[c]
    worst = 0.0
    transposed_hits = 0
    for i in range(verts.shape[0]):
        o = ob.vertex_shader(d, i)
        v = np.array([verts[i, 0], verts[i, 1], verts[i, 2], 1.0], np.float64)
        exact = M @ v
        bound = _mat_vec_bound(M, v)
        err = np.abs(o[:4].astype(np.float64) - exact)
        assert np.all(err <= bound), (i, o[:4], exact, bound)
        worst = max(worst, float((err / bound).max()))
        if np.any(np.abs(o[:4] - M.T @ v) > 1e3 * bound):
            transposed_hits += 1
        # NORMAL = normalize(n) (basic_vs.c:31), UV copied (basic_vs.c:32)
        n = verts[i, 3:6].astype(np.float64)
        assert np.allclose(o[4:7], n / np.linalg.norm(n), rtol=0, atol=4e-7)
        assert o[7] == verts[i, 6] and o[8] == verts[i, 7]
        assert np.all(o[9:] == 0.0)
    assert transposed_hits > 290  # the bound separates M from its transpose on these inputs


def test_basic_vs_dot_is_pairwise(oracle):
    """math.h:142-146 sums (x*x' + y*y') + (z*z' + w*w'). With the row (1e8, 1, -1e8, 1) and the
    position (1, 1, 1, 1) the pairwise order gives 0 and the sequential order of math.h:137-140 gives 1."""
    ob = oracle
    cb = np.zeros(48, np.float32)
    cb[0:4] = (1e8, 1.0, -1e8, 1.0)
    cb[5] = cb[10] = cb[15] = 1.0
    verts = np.zeros((3, 8), np.float32)
    verts[:, 0:3] = 1.0
    verts[:, 3] = 1.0
    o = ob.vertex_shader(_draw(ob, verts, ob.VS_BASIC, cb), 0)
    assert o[0] == 0.0
    seq = np.float32(np.float32(np.float32(1e8) + np.float32(1.0)) + np.float32(-1e8)) + np.float32(1.0)
    assert seq == 1.0  # the other order would be visible


def test_vertex_lighting_vs_colour_is_tonemapped_env_lookup(oracle):
    """vertex_lighting_vs.c:30-36 on a constant environment map: COLOR = (1 - e^-c)^(1/2.2) for
    every normal, position as basic_vs, UV copied."""
    ob = oracle
    rng = np.random.default_rng(12)
    cb = _rand_cb(rng)
    c = np.array([0.3, 0.7, 1.9, 1.0], np.float32)
    env = np.broadcast_to(c, (16, 32, 4)).copy()
    verts = rng.uniform(-2, 2, (60, 8)).astype(np.float32)
    d = _draw(ob, verts, ob.VS_VERTEX_LIGHTING, cb, vs_tex=env)
    M = cb[:16].astype(np.float64).reshape(4, 4)
    want = (1.0 - np.exp(-c[:3].astype(np.float64))) ** (1.0 / 2.2)
    for i in range(verts.shape[0]):
        o = ob.vertex_shader(d, i)
        v = np.array([verts[i, 0], verts[i, 1], verts[i, 2], 1.0])
        assert np.all(np.abs(o[:4] - M @ v) <= _mat_vec_bound(M, v))
        assert np.allclose(o[4:7], want, rtol=0, atol=3e-6), (o[4:7], want)
        assert o[7] == verts[i, 6] and o[8] == verts[i, 7]


def test_fullscreen_vs_direction_uses_view_from_clip_then_world_from_view(oracle):
    ob = oracle
    rng = np.random.default_rng(13)
    cb = _rand_cb(rng)
    Vm = cb[16:32].astype(np.float64).reshape(4, 4)
    Wm = cb[32:48].astype(np.float64).reshape(4, 4)
    assert np.abs(Vm @ Wm - Wm @ Vm).max() > 0.5  # swapping the two slots is visible
    verts = rng.uniform(0, 1, (90, 8)).astype(np.float32)
    d = _draw(ob, verts, ob.VS_FULLSCREEN, cb)
    swapped_hits = 0
    for i in range(verts.shape[0]):
        o = ob.vertex_shader(d, i)
        p32 = np.array([np.float32(verts[i, 0] * np.float32(2)) - np.float32(1),
                        np.float32(verts[i, 1] * np.float32(2)) - np.float32(1), verts[i, 2], verts[i, 3]],
                       np.float32)
        # 2x is exact, so x*2-1 has one rounding: the f64 value rounded once
        assert np.array_equal(o[:4], p32)
        p = p32.astype(np.float64)
        dv = Vm @ p
        dv_err = _mat_vec_bound(Vm, p)
        d3 = np.array([dv[0], dv[1], dv[2], 0.0])
        exact = (Wm @ d3)[:3]
        bound = (np.abs(Wm[:3]) @ np.append(dv_err[:3], 0.0)) * (1 + 8 * U) + _mat_vec_bound(Wm[:3], np.abs(d3) +
                                                                                            np.append(dv_err[:3], 0))
        assert np.all(np.abs(o[4:7] - exact) <= bound), (i, o[4:7], exact, bound)
        dsw = Wm @ p
        exact_sw = (Vm @ np.array([dsw[0], dsw[1], dsw[2], 0.0]))[:3]
        if np.any(np.abs(o[4:7] - exact_sw) > 1e3 * bound):
            swapped_hits += 1
        assert o[7] == 0.0
    assert swapped_hits > 85


def test_passthrough_vs_is_exact(oracle):
    ob = oracle
    rng = np.random.default_rng(14)
    verts = rng.uniform(-1, 2, (30, 8)).astype(np.float32)
    d = _draw(ob, verts, ob.VS_PASSTHROUGH, None)
    for i in range(verts.shape[0]):
        o = ob.vertex_shader(d, i)
        x2 = np.float32(np.float64(verts[i, 0]) * 2 - 1)  # exact product, one rounding
        y2 = np.float32(np.float64(verts[i, 1]) * 2 - 1)
        assert o[0] == x2 and o[1] == y2 and o[2] == verts[i, 2] and o[3] == verts[i, 3]
        assert np.array_equal(o[4:7], verts[i, 4:7])


# ------------------------------------------------------------------ SH basis


def _sh_closed_form(x, y, z):
    """Real spherical harmonics of bands 0..2 from their definition, index l(l+1)+m."""
    s = math.sqrt(math.pi)
    return np.array([
        1.0 / (2.0 * s),
        math.sqrt(3.0) / (2.0 * s) * y,
        math.sqrt(3.0) / (2.0 * s) * z,
        math.sqrt(3.0) / (2.0 * s) * x,
        math.sqrt(15.0) / (2.0 * s) * x * y,
        math.sqrt(15.0) / (2.0 * s) * y * z,
        math.sqrt(5.0) / (4.0 * s) * (3.0 * z * z - 1.0),
        math.sqrt(15.0) / (2.0 * s) * x * z,
        math.sqrt(15.0) / (4.0 * s) * (x * x - y * y),
    ])


def test_sh9_matches_the_definition(oracle):
    rng = np.random.default_rng(15)
    dirs = rng.normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = np.vstack([dirs, np.eye(3), -np.eye(3)])
    for n in dirs:
        got = oracle.sh9(n)
        want = _sh_closed_form(*n)
        # the constants are f32-rounded (relative error < 2^-24)
        assert np.allclose(got, want, rtol=1.2e-7, atol=1e-15), (n, got, want)
    # +z: only Y_1^0 and Y_2^0 are non-zero
    gz = oracle.sh9((0.0, 0.0, 1.0))
    assert np.count_nonzero(np.abs(gz[[1, 3, 4, 5, 7, 8]]) > 0) == 0
    assert gz[2] == pytest.approx(math.sqrt(3 / (4 * math.pi)), rel=1e-7)
    assert gz[6] == pytest.approx(math.sqrt(5 / math.pi) / 2, rel=1e-7)


def test_sh9_is_orthonormal_on_the_sphere(oracle):
    """Gauss-Legendre in cos(theta) x uniform phi integrates the products (degree <= 4) exactly."""
    mu, wmu = np.polynomial.legendre.leggauss(8)
    nphi = 16
    G = np.zeros((9, 9))
    for m, wm in zip(mu, wmu):
        st = math.sqrt(1.0 - m * m)
        for k in range(nphi):
            ph = 2 * math.pi * k / nphi
            Y = oracle.sh9((st * math.cos(ph), st * math.sin(ph), m))
            G += wm * (2 * math.pi / nphi) * np.outer(Y, Y)
    assert np.abs(G - np.eye(9)).max() < 1e-6, G
