"""CPU pins of the oracle's extension shading (mip chains, the level nearest the footprint, the
Lambert light, the textured environment shader) against closed forms and against the
reference's own shaders where the extensions reduce to them.

The reference has no mips, lights or second texture (common_shader_core.h:20-24, basic_ps.c:21-24,
env_lighting_ps.c:13-22); the extensions are defined by include/malevich.h and restated by the
oracle. These checks fix the restatement independently:
  - mip generation (oracle/mips.py): constant textures stay constant, 2x2 windows average, odd
    sides drop their last row or column, integer rounding halves up, float16 rounding;
  - the level rule against log2 of the footprint, and at its thresholds;
  - footprints against closed forms (chords of unit vectors at known angles);
  - whole frames: a screen-filling quad whose uv advance 2^k texels per pixel samples level k;
    the direction footprint against an FP64 evaluation of the interpolated direction field;
  - reductions: the Lambert shader with no light and ambient 1 is basic_ps bit for bit; the
    textured environment shader with a white 1x1 albedo and a one-level map is env_lighting_ps
    bit for bit;
  - the Lambert light against its FP64 closed form on a flat quad.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from malevich_amd import scenes as S
from oracle import mips as M
from tests.render_utils import channel_diff

W = H = 64


def test_mip_chain_closed_forms():
    # constant textures stay constant at every level, in every format
    for arr in (np.full((37, 20), 0x80402010, np.uint32), np.full((16, 9, 4), 0.375, np.float32),
                np.full((5, 32, 4), 1.25, np.float16)):
        for lvl in M.chain(arr):
            assert np.all(lvl == arr.reshape(-1)[0]) if arr.dtype == np.uint32 else np.all(lvl == arr[0, 0])
    # sizes: max(w >> l, 1) x max(h >> l, 1) down to 1x1
    assert M.level_sizes(37, 20) == [(37, 20), (18, 10), (9, 5), (4, 2), (2, 1), (1, 1)]
    assert M.level_sizes(1, 9) == [(1, 9), (1, 4), (1, 2), (1, 1)]
    # RGBA8: bytes averaged in integers, halves rounded up
    t = np.array([[0x000000FF | (1 << 8), 0x00000000], [0x00000001 | (3 << 8), 0x00000001]], np.uint32)
    l1 = M.next_level(t)
    assert l1.shape == (1, 1)
    r = (255 + 0 + 1 + 1 + 2) // 4  # 64.25 -> 64
    g = (1 + 0 + 3 + 0 + 2) // 4    # 1.0 -> 1
    assert int(l1[0, 0]) == r | (g << 8)
    assert int(M.next_level(np.array([[1, 2]], np.uint32))[0, 0]) == 2  # (1 + 2 + 1) // 2: 1.5 rounds up
    # an odd side drops its last row or column
    f = np.arange(3 * 5 * 4, dtype=np.float32).reshape(3, 5, 4)
    g1 = M.next_level(f)
    assert g1.shape == (1, 2, 4)
    assert np.array_equal(g1[0, 0], (f[0, 0] + f[0, 1] + f[1, 0] + f[1, 1]) / 4)
    assert np.array_equal(g1[0, 1], (f[0, 2] + f[0, 3] + f[1, 2] + f[1, 3]) / 4)
    # a side of 1 filters with a 1x2 window
    col = np.arange(4 * 1 * 4, dtype=np.float32).reshape(4, 1, 4)
    assert np.array_equal(M.next_level(col)[:, 0], (col[0::2, 0] + col[1::2, 0]) / 2)
    # float16: the f32 average rounded to half, nearest-even (1 + 2^-11 is a tie between 1 and 1 + 2^-10)
    h = np.array([[[1.0] * 4, [1.0 + 2 ** -10] * 4], [[1.0] * 4, [1.0] * 4]], np.float16)
    assert M.next_level(h)[0, 0, 0] == np.float16(1.0)
    h2 = np.array([[[1.0 + 2 ** -10] * 4, [1.0 + 2 ** -10] * 4], [[1.0 + 2 ** -10] * 4, [1.0] * 4]], np.float16)
    assert M.next_level(h2)[0, 0, 0] == np.float16(1.0 + 2 ** -10)  # 1 + 3/4 * 2^-10 rounds up


def test_nearest_level_against_log2(oracle):
    rng = np.random.default_rng(1)
    for rho2 in np.exp(rng.uniform(-3, 25, 2000)).astype(np.float32):
        want = min(max(int(math.floor(0.5 * math.log2(float(rho2)) + 0.5)), 0), 11)
        lo = 2.0 ** (2 * want - 1)
        if abs(float(rho2) - lo) < 1e-4 * lo or abs(float(rho2) - 4 * lo) < 1e-4 * lo:
            continue  # the log2 rounding at the thresholds is checked exactly below
        assert oracle.nearest_level(float(rho2), 12) == want, (rho2, want)
    for k in range(1, 12):
        t = np.float32(2.0 ** (2 * k - 1))
        assert oracle.nearest_level(float(t), 12) == k
        assert oracle.nearest_level(float(np.nextafter(t, np.float32(0))), 12) == k - 1
    assert oracle.nearest_level(1e30, 5) == 4  # clamped to the chain
    assert oracle.nearest_level(float("nan"), 5) == 0


def _quad(ob, ps, u_span, v_span, normals=None, ps_tex=None, ps_tex1=None, ps_cb=None):
    """A screen-filling quad at z = 0.5 drawn with basic_vs under an identity clip_from_world
    (w = 1, so interpolation is affine): uv from (0, 0) at the bottom-left to (u_span, v_span) at
    the top-right. Both windings are drawn; the back-face rule keeps one of each pair."""
    p = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (4, 1)) if normals is None else np.asarray(normals, np.float32)
    uv = np.array([[0, 0], [u_span, 0], [u_span, v_span], [0, v_span]], np.float32)
    verts = np.concatenate([p, np.full((4, 1), 0.5, np.float32), n, uv], 1)
    idx = np.array([0, 1, 2, 2, 3, 0, 0, 2, 1, 2, 0, 3], np.uint32)
    cb = np.zeros(48, np.float32)
    cb[0:16] = np.eye(4, dtype=np.float32).ravel()
    d = ob.Draw(idx, verts, ob.VS_BASIC, ps, cb=cb, ps_tex=ps_tex, ps_tex1=ps_tex1, ps_cb=ps_cb,
                viewport=(0, 0, W, H, 0, 1))
    rt = ob.RenderTarget(W, H)
    rt.clear_render_target_view((0, 0, 0, 0))
    rt.clear_depth_stencil_view(0.0)
    ob.draw_indexed(rt, d)
    return rt


def _srgb8(c):
    s = 1.055 * np.power(np.maximum(np.asarray(c, np.float64), 0), 1 / 2.4) - 0.055
    return np.rint(np.maximum(s, 0) * 255).astype(np.int64)


def _rgb(colors):
    return np.stack([(colors >> 16) & 255, (colors >> 8) & 255, colors & 255], -1).astype(np.int64)


def _level_colours(n):
    return [(17 + 23 * j) | ((200 - 11 * j) << 8) | ((60 + 13 * j) << 16) for j in range(n)]


def test_uv_footprint_picks_the_level_it_covers(oracle):
    """uv advance 2^k level-0 texels per pixel in x and y: every pixel samples level k (each level
    of the chain is a distinct constant colour); Lambert with no light and ambient 1."""
    tw = 256
    sizes = M.level_sizes(tw, tw)
    cols = _level_colours(len(sizes))
    tex = oracle.MipTexture([np.full((h, w), c, np.uint32) for (w, h), c in zip(sizes, cols)])
    cb = np.array([0, 0, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0], np.float32)
    for k in range(5):
        span = W * 2 ** k / tw
        rt = _quad(oracle, oracle.PS_BASIC_LAMBERT, span, span, ps_tex=tex, ps_cb=cb)
        got = _rgb(rt.colors[1:-1, 1:-1])
        c = cols[k]
        want = _srgb8(np.array([c & 255, (c >> 8) & 255, (c >> 16) & 255]) / 255.0)
        assert np.abs(got - want).max() <= 1, (k, got[0, 0], want)


def test_direction_footprint_matches_fp64_direction_field(oracle):
    """ENV_TEXTURED on a quad whose vertex normals sweep an arc: the level of every pixel equals
    the one computed in FP64 from the same direction field (normalised affine interpolation of the
    normals, chords to the next corners, level by log2), wherever the FP64 footprint is not at a
    threshold. The map's level j is the constant 0.04 (j + 1), the albedo a white 1x1 texture."""
    mw = 4096
    sizes = M.level_sizes(mw, mw // 2)
    chain = [np.full((h, w, 4), 0.04 * (j + 1), np.float32) for j, (w, h) in enumerate(sizes)]
    env = oracle.MipTexture(chain)
    white = np.full((1, 1), 0xFFFFFFFF, np.uint32)
    seen = set()
    for half_angle in (0.04, 0.08, 0.15, 0.3, 0.8, 1.2):
        a = np.array([-half_angle, half_angle, half_angle, -half_angle])
        b = np.array([-0.3, -0.3, 0.3, 0.3]) * half_angle
        normals = np.stack([np.cos(a) * np.cos(b), np.sin(a) * np.cos(b), np.sin(b)], 1)
        rt = _quad(oracle, oracle.PS_ENV_TEXTURED, 1, 1, normals=normals, ps_tex=env, ps_tex1=white)
        got = _rgb(rt.colors)[..., 0]
        n32 = normals.astype(np.float32).astype(np.float64)

        ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
        in_a = (1 - ys / H) <= xs / W  # triangle (0, 1, 2) below the diagonal, (2, 3, 0) above

        def field(x, y):  # pixel corner -> the normalised direction on the pixel's own triangle
            s, t = (x / W)[..., None], (1 - y / H)[..., None]
            da = n32[0] + (n32[1] - n32[0]) * (s - t) + (n32[2] - n32[0]) * t
            db = n32[2] + (n32[3] - n32[2]) * (t - s) + (n32[0] - n32[2]) * (1 - t)
            d = np.where(in_a[..., None], da, db)
            return d / np.linalg.norm(d, axis=-1, keepdims=True)

        d0, dx, dy = field(xs, ys), field(xs + 1, ys), field(xs, ys + 1)
        k = mw / (2 * math.pi)
        rho2 = np.maximum(((dx - d0) ** 2).sum(-1), ((dy - d0) ** 2).sum(-1)) * k * k
        lvl = np.clip(np.floor(0.5 * np.log2(rho2) + 0.5), 0, len(sizes) - 1).astype(np.int64)
        frac = 0.5 * np.log2(rho2) + 0.5 - np.floor(0.5 * np.log2(rho2) + 0.5)
        safe = (np.abs(frac) > 0.01) & (np.abs(frac - 1) > 0.01)
        want = _srgb8_tone(np.float32(0.04) * (lvl + 1.0))
        # pixels near the diagonal are left out: both windings of both triangles meet there
        band = np.abs(xs / W - (1 - ys / H)) > 0.1
        m = safe & band
        assert m.sum() > 0.5 * W * H
        bad = np.abs(got - want)[m]
        assert bad.max() <= 1, (half_angle, np.unique(lvl[m]), np.unique(got[m]))
        seen |= set(np.unique(lvl[m]).tolist())
    assert seen >= {0, 1, 2, 3, 4, 5}, seen  # the sweeps reach levels 0..5


def _srgb8_tone(c):
    t = np.power(1.0 - np.exp(-np.asarray(c, np.float64)), 1 / 2.2)
    return np.rint(t * 255).astype(np.int64)


def test_lambert_reduces_to_basic_ps_and_matches_its_closed_form(oracle):
    mesh, cb, tex = S.c1_scene(W, H)
    base = oracle.Draw(mesh.indices, mesh.vertices, oracle.VS_BASIC, oracle.PS_BASIC, cb=cb, ps_tex=tex,
                       viewport=(0, 0, W, H, 0, 1))
    off = oracle.Draw(mesh.indices, mesh.vertices, oracle.VS_BASIC, oracle.PS_BASIC_LAMBERT, cb=cb, ps_tex=tex,
                      ps_cb=np.array([0.3, 0.5, 0.8, 0, 0, 0, 0, 0, 1, 1, 1, 0], np.float32),
                      viewport=(0, 0, W, H, 0, 1))
    frames = []
    for d in (base, off):
        rt = oracle.RenderTarget(W, H)
        rt.clear_render_target_view(S.CLEAR_COLOR)
        rt.clear_depth_stencil_view(0.0)
        oracle.draw_indexed(rt, d)
        frames.append(rt.colors.copy())
    assert np.array_equal(frames[0], frames[1])
    assert (frames[0] != frames[0][0, 0]).sum() > 500
    # closed form on a flat quad facing +z, constant albedo, light at 60 degrees
    c = 0x00408020
    tex1 = np.full((8, 8), c, np.uint32)
    L = np.array([0.0, math.sin(math.pi / 3), math.cos(math.pi / 3)], np.float32)
    cbl = np.array([L[0], L[1], L[2], 0, 0.9, 0.7, 0.5, 0, 0.05, 0.1, 0.15, 0], np.float32)
    rt = _quad(oracle, oracle.PS_BASIC_LAMBERT, 1, 1, ps_tex=tex1, ps_cb=cbl)
    alb = np.array([c & 255, (c >> 8) & 255, (c >> 16) & 255]) / 255.0
    light = np.array([0.9, 0.7, 0.5]) * max(float(np.dot([0, 0, 1], L.astype(np.float64))), 0.0) + \
        np.array([0.05, 0.1, 0.15])
    want = _srgb8(alb * light)
    assert np.abs(_rgb(rt.colors) - want).max() <= 1
    # a light from behind leaves the ambient term
    cbb = cbl.copy()
    cbb[2] = -cbb[2]
    rt = _quad(oracle, oracle.PS_BASIC_LAMBERT, 1, 1, ps_tex=tex1, ps_cb=cbb)
    assert np.abs(_rgb(rt.colors) - _srgb8(alb * np.array([0.05, 0.1, 0.15]))).max() <= 1


def test_env_textured_reduces_to_env_lighting_ps(oracle):
    """A white 1x1 albedo decodes to exactly 1 and bilinear keeps it 1, and a one-level map has no
    level to choose: the frame is env_lighting_ps's bit for bit (vertex_lighting_vs feeding it,
    as the 1080p workload does)."""
    sc = S.c2_scene(W * 2, H, n_instances=4, seed=3, texture_size=0)
    irr = S.smooth_env_map(64, 32, seed=1)
    rad = S.smooth_env_map(128, 64, seed=2)
    white = np.full((1, 1), 0xFFFFFFFF, np.uint32)
    frames = []
    for ps, extra in ((oracle.PS_ENV_LIGHTING, {}), (oracle.PS_ENV_TEXTURED, {"ps_tex1": white})):
        rt = oracle.RenderTarget(W * 2, H)
        rt.clear_render_target_view(S.CLEAR_COLOR)
        rt.clear_depth_stencil_view(0.0)
        for m in sc.instances:
            oracle.draw_indexed(rt, oracle.Draw(m.indices, m.vertices, oracle.VS_VERTEX_LIGHTING, ps, cb=sc.cb,
                                                vs_tex=irr, ps_tex=rad, viewport=(0, 0, W * 2, H, 0, 1), **extra))
        frames.append(rt.colors.copy())
    assert np.array_equal(frames[0], frames[1])
    assert (frames[0] != frames[0][0, 0]).sum() > 300


def test_env_textured_multiplies_albedo_on_the_uv_level(oracle):
    """A one-level map of constant radiance e and an albedo chain of distinct constant levels:
    every pixel is tone(albedo_k * e) with k from the uv footprint (2^k texels per pixel)."""
    e = np.float32(0.8)
    env = np.full((8, 16, 4), e, np.float32)
    tw = 128
    sizes = M.level_sizes(tw, tw)
    cols = _level_colours(len(sizes))
    alb = oracle.MipTexture([np.full((h, w), c, np.uint32) for (w, h), c in zip(sizes, cols)])
    for k in (0, 2, 3):
        span = W * 2 ** k / tw
        rt = _quad(oracle, oracle.PS_ENV_TEXTURED, span, span, ps_tex=env, ps_tex1=alb)
        c = cols[k]
        a = np.array([c & 255, (c >> 8) & 255, (c >> 16) & 255]) / 255.0
        got = _rgb(rt.colors[1:-1, 1:-1])
        assert np.abs(got - _srgb8_tone(a * float(e))).max() <= 1, k


@pytest.mark.parametrize("fmt", ["rgba8", "rgba32f"])
def test_oracle_reads_the_chain_levels(oracle, fmt):
    """tex_level offsets: sampling a chain's level k through the oracle equals sampling the level's
    own array (the levels are random, so a wrong offset or size is visible)."""
    rng = np.random.default_rng(7)
    l0 = rng.integers(0, 2 ** 32, (40, 24), dtype=np.uint64).astype(np.uint32) if fmt == "rgba8" else \
        rng.uniform(0, 1, (40, 24, 4)).astype(np.float32)
    ch = M.chain(l0)
    t = oracle.MipTexture(ch)
    for k in range(len(ch)):
        for u, v in rng.uniform(-0.1, 1.1, (20, 2)):
            assert np.array_equal(oracle.bilinear_level(t, k, u, v), oracle.bilinear(ch[k], u, v))
