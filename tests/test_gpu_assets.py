"""The asset path on the device: load_texture's sRGB pre-pass (main.c:546-558) bit-exact against
the oracle for every byte value, and a scene read from .octrn files (written here, in the
reference's layout) drawn as the reference draws its FTM scene (basic_vs + basic_ps, one sRGB
texture per mesh) against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import assets as A
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def test_srgb_pre_pass_every_byte(device, oracle):
    i, j = np.meshgrid(np.arange(256, dtype=np.uint32), np.arange(256, dtype=np.uint32), indexing="ij")
    texels = (i | (j << 8) | ((255 - i) << 16) | (((i + j) & 255) << 24)).astype(np.uint32)
    print("case: srgb pre-pass 256x256", flush=True)
    t = device.create_texture2d(texels)
    try:
        t.srgb_to_linear()
        got = t.read()
    finally:
        t.release()
    np.testing.assert_array_equal(got, oracle.srgb_to_linear_rgba8(texels))


def test_scene_from_octrn_files(device, oracle, tmp_path):
    W, H = 320, 184
    names = []
    for k in range(3):
        T = np.eye(4)
        T[:3, 3] = (0.9 * k - 0.9, 0.2 * k, 0.1 * k)
        m = S.sphere_mesh(6 + k, displace=0.1, seed=40 + k, transform=T)
        flat = m.vertices[m.indices]  # de-indexed, as the shipped meshes are
        A.write_octrn_mesh(tmp_path / f"part{k}_mesh.octrn", flat, np.arange(len(flat), dtype=np.uint32))
        names.append(f"part{k}")
    scene = A.load_scene(tmp_path, names, tex_size=64, seed=3)
    cb = S.camera_cb((3.5, 1.0, 1.0), (0.0, 0.0, 0.0), 75.0, W, H)
    draws_gpu, draws_ref = [], []
    for m, tex in scene:
        draws_gpu.append(DrawSpec(m.indices, m.vertices, "basic", "basic", cb=cb, ps_tex=tex))
        lin = oracle.srgb_to_linear_rgba8(tex)
        draws_ref.append(DrawSpec(m.indices, m.vertices, "basic", "basic", cb=cb, ps_tex=lin))
    print(f"case: octrn scene {W}x{H}, {len(scene)} meshes", flush=True)
    # the device converts each texture in place before drawing, as load_texture does on load
    orig = device.create_texture2d
    made = []

    def create_and_convert(data, *a, **kw):
        t = orig(data, *a, **kw)
        if t.format == L.FORMAT_RGBA8_UNORM:
            t.srgb_to_linear()
        made.append(t)
        return t

    device.create_texture2d = create_and_convert
    try:
        g = render_gpu(device, W, H, draws_gpu, S.CLEAR_COLOR)
    finally:
        device.create_texture2d = orig
    assert len(made) == len(scene)
    o = render_oracle(oracle, W, H, draws_ref, S.CLEAR_COLOR)
    assert_frames_match(g, o, color_tol=1)
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 2000
