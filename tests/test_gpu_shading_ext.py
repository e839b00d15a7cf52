"""The extension shading on the GPU against the oracle: mip chains built on the device, the level
nearest the pixel's footprint, the Lambert pixel shader (the 640x360 workload's shading) and the
textured environment shader (the 1080p workload's shading).

Bar (written here): device mip chains bit-identical to oracle/mips.py's in every format;
visibility, depth and stats bit-exact against the oracle; colour within 1/255 (the shaders'
transcendentals are the hardware's, the oracle's libm's); where an extension shader reduces to a
reference shader (no light and ambient 1; a white albedo on a one-level map) the GPU frames equal
the reference shader's GPU frames bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from oracle import mips as M
from tests.render_utils import DrawSpec, assert_frames_match, channel_diff, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(20, 37), (64, 64), (9, 1), (32, 128), (128, 256)])
def test_device_mip_chains_equal_the_restatement(device, shape):
    h, w = shape
    rng = np.random.default_rng(h * 1000 + w)
    print(f"case: generate_mips {w}x{h}", flush=True)
    for level0 in (rng.integers(0, 2 ** 32, (h, w), dtype=np.uint32),
                   rng.uniform(0, 4, (h, w, 4)).astype(np.float32),
                   rng.uniform(0, 4, (h, w, 4)).astype(np.float16)):
        want = M.chain(level0)
        t = device.create_texture2d(level0, mip_levels=len(want), generate_mips=True)
        try:
            got = t.read_levels()
        finally:
            t.release()
        assert len(got) == len(want)
        for k, (a, b) in enumerate(zip(got, want)):
            assert a.shape == b.shape, (k, a.shape, b.shape)
            np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8), err_msg=f"{level0.dtype} level {k}")


def test_c1_lambert_frame_against_oracle(device, oracle):
    """The 640x360 workload as configured: a 10,242-vertex / 20,480-triangle sphere, a 512x512 RGBA8
    texture with its mip chain, one directional light."""
    W, H = 640, 360
    mesh, cb, tex = S.c1_scene(W, H)
    assert mesh.vertices.shape[0] == 10242 and mesh.indices.size == 3 * 20480
    chain = M.chain(tex)
    print(f"case: c1 lambert {W}x{H}, {len(chain)} texture levels", flush=True)
    d = DrawSpec(mesh.indices, mesh.vertices, "basic", "basic_lambert", cb=cb, ps_tex=chain, ps_cb=S.c1_light())
    g = render_gpu(device, W, H, [d], S.CLEAR_COLOR)
    o = render_oracle(oracle, W, H, [d], S.CLEAR_COLOR)
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 20000
    print(f"  colour max diff {channel_diff(g.colors, o.colors)}", flush=True)
    assert_frames_match(g, o, color_tol=1)


def test_extension_shaders_reduce_to_the_reference_shaders(device):
    W, H = 160, 96
    mesh, cb, tex = S.c1_scene(W, H)
    base = render_gpu(device, W, H, [DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb, ps_tex=tex)],
                      S.CLEAR_COLOR)
    no_light = np.array([0.3, 0.5, 0.8, 0, 0, 0, 0, 0, 1, 1, 1, 0], np.float32)
    lam = render_gpu(device, W, H, [DrawSpec(mesh.indices, mesh.vertices, "basic", "basic_lambert", cb=cb,
                                             ps_tex=tex, ps_cb=no_light)], S.CLEAR_COLOR)
    np.testing.assert_array_equal(lam.colors, base.colors)
    sc = S.c2_scene(W, H, n_instances=4, seed=3, texture_size=0)
    irr = S.smooth_env_map(64, 32, seed=1).astype(np.float16)
    rad = S.smooth_env_map(128, 64, seed=2).astype(np.float16)
    white = np.full((1, 1), 0xFFFFFFFF, np.uint32)
    ref = render_gpu(device, W, H, [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_lighting", cb=sc.cb,
                                             vs_tex=irr, ps_tex=rad) for m in sc.instances], S.CLEAR_COLOR)
    ext = render_gpu(device, W, H, [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_textured", cb=sc.cb,
                                             vs_tex=irr, ps_tex=rad, ps_tex1=white) for m in sc.instances],
                     S.CLEAR_COLOR)
    assert (ref.vis_tri != 0xFFFFFFFF).sum() > 1000
    np.testing.assert_array_equal(ext.colors, ref.colors)


def _quad(ps, u_span, normals=None, **kw):
    p = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (4, 1)) if normals is None else np.asarray(normals, np.float32)
    uv = np.array([[0, 0], [u_span, 0], [u_span, u_span], [0, u_span]], np.float32)
    verts = np.concatenate([p, np.full((4, 1), 0.5, np.float32), n, uv], 1)
    idx = np.array([0, 1, 2, 2, 3, 0, 0, 2, 1, 2, 0, 3], np.uint32)
    cb = np.zeros(48, np.float32)
    cb[0:16] = np.eye(4, dtype=np.float32).ravel()
    return DrawSpec(idx, verts, "basic", ps, cb=cb, **kw)


def test_footprint_levels_against_oracle(device, oracle):
    """Screen-filling quads whose uv and direction footprints reach each level of distinct-constant
    chains: every pixel's level (its colour) equals the oracle's."""
    W = H = 64
    sizes = M.level_sizes(256, 256)
    cols = [(17 + 23 * j) | ((200 - 11 * j) << 8) | ((60 + 13 * j) << 16) for j in range(len(sizes))]
    tex = [np.full((h, w), c, np.uint32) for (w, h), c in zip(sizes, cols)]
    light = np.array([0, 0, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0], np.float32)
    env = [np.full((h, w, 4), 0.04 * (j + 1), np.float16) for j, (w, h) in enumerate(M.level_sizes(4096, 2048))]
    white = np.full((1, 1), 0xFFFFFFFF, np.uint32)
    draws = [[_quad("basic_lambert", W * 2 ** k / 256, ps_tex=tex, ps_cb=light)] for k in range(5)]
    for ha in (0.04, 0.15, 0.3, 0.8, 1.2):
        a = np.array([-ha, ha, ha, -ha])
        b = np.array([-0.3, -0.3, 0.3, 0.3]) * ha
        n = np.stack([np.cos(a) * np.cos(b), np.sin(a) * np.cos(b), np.sin(b)], 1)
        draws.append([_quad("env_textured", 1, normals=n, ps_tex=env, ps_tex1=white)])
    print(f"case: footprint quads ({len(draws)} frames)", flush=True)
    for dl in draws:
        g = render_gpu(device, W, H, dl, (0, 0, 0, 0))
        o = render_oracle(oracle, W, H, dl, (0, 0, 0, 0))
        assert_frames_match(g, o, color_tol=1)


def test_env_textured_scene_against_oracle(device, oracle):
    """The 1080p workload's shading on a smaller frame: vertex_lighting_vs with an RGBA16F
    irradiance map, the textured environment shader with an RGBA16F radiance chain and per-mesh
    RGBA8 albedo chains, all chains built on the device and read back for the oracle."""
    W, H = 480, 272
    sc = S.c2_scene(W, H, n_instances=12, seed=23, texture_size=0)
    irr, rad, albedo = S.c2_textures(n_instances=12, seed=23, albedo_size=256)
    rad = rad[::4, ::4].copy()  # 512x256
    print(f"case: env_textured {W}x{H}, {len(sc.instances)} instances", flush=True)
    t_rad = device.create_texture2d(rad, mip_levels=len(M.level_sizes(512, 256)), generate_mips=True)
    rad_chain = t_rad.read_levels()
    t_rad.release()
    chains = []
    for a in albedo:
        t = device.create_texture2d(a, mip_levels=9, generate_mips=True)
        chains.append(t.read_levels())
        t.release()
    draws = [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_textured", cb=sc.cb, vs_tex=irr,
                      ps_tex=rad_chain, ps_tex1=ch) for m, ch in zip(sc.instances, chains)]
    g = render_gpu(device, W, H, draws, S.CLEAR_COLOR)
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 0.3 * W * H
    print(f"  colour max diff {channel_diff(g.colors, o.colors)}", flush=True)
    assert_frames_match(g, o, color_tol=1)
