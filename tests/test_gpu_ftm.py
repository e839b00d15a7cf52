"""The reference's own FTM scene (main.c:1317-1361) drawn on the GPU against the oracle: the seven
meshes of assets/ftm_*_mesh.octrn (47,936 triangles, the committed fixture
tests/golden/ftm_scene.npz), basic_vs + basic_ps, the start-up camera (main.c:1422-1476) at the
reference's 1200x720 (main.c:21-22), render()'s clears (main.c:1270-1272), reversed-Z GREATER_EQUAL.
The meshes' textures are not shipped with the reference: seeded sRGB stand-ins, converted by
load_texture's pre-pass (on the device for the GPU frame, by the oracle for its frame).

Bar: visibility (triangle id, draw tag), depth and every stat bit-exact; colour within 1/255
(basic_ps's tone curve runs on the hardware exp2/log2).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "ftm_scene.npz"
W, H = 1200, 720


def ftm_meshes():
    z = np.load(FIXTURE)
    return [(str(n), z[f"{n}_vertices"], z[f"{n}_indices"]) for n in z["names"]]


def test_ftm_scene_against_oracle(device, oracle):
    meshes = ftm_meshes()
    cb = oracle.camera_cb((3.5, 1.0, 1.0), 0.0, 0.0, 75.0, 0.01, W, H)
    texs = [S.random_rgba8(512, 7 + k) for k in range(len(meshes))]
    print(f"case: FTM scene {W}x{H}, {sum(i.size for _, _, i in meshes) // 3} triangles", flush=True)
    draws_gpu = [DrawSpec(i, v, "basic", "basic", cb=cb, ps_tex=t) for (_, v, i), t in zip(meshes, texs)]
    draws_ref = [DrawSpec(i, v, "basic", "basic", cb=cb, ps_tex=oracle.srgb_to_linear_rgba8(t))
                 for (_, v, i), t in zip(meshes, texs)]
    orig = device.create_texture2d

    def create_and_convert(data, *a, **kw):  # load_texture(..., true) converts on load
        t = orig(data, *a, **kw)
        if t.format == L.FORMAT_RGBA8_UNORM:
            t.srgb_to_linear()
        return t

    device.create_texture2d = create_and_convert
    try:
        g = render_gpu(device, W, H, draws_gpu, S.CLEAR_COLOR)
    finally:
        device.create_texture2d = orig
    o = render_oracle(oracle, W, H, draws_ref, S.CLEAR_COLOR)
    covered = (o.vis_tri != 0xFFFFFFFF).mean()
    print(f"  covered {covered:.3f}, draws seen {sorted(set(np.unique(o.vis_draw).tolist()))[:8]}", flush=True)
    assert covered > 0.5  # the scene fills most of the start-up view
    assert_frames_match(g, o, color_tol=1)
