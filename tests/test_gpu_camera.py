"""Per-view constant buffers built on the device equal the oracle's (the reference's camera
set-up and update(), main.c:1422-1562) bit for bit, and a frame drawn with one equals the
oracle's frame drawn with the oracle's buffer."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from malevich_amd import _lib as L
from malevich_amd import scenes as S
from tests.render_utils import DrawSpec, assert_frames_match, render_gpu, render_oracle

pytestmark = pytest.mark.gpu


def _cameras(n, seed):
    rng = np.random.default_rng(seed)
    cams = [((3.5, 1.0, 1.0), 0.0, 0.0, 75.0, 0.01, 1920, 1080)]  # the reference's start-up camera
    for _ in range(n - 1):
        cams.append((tuple(float(x) for x in rng.uniform(-15, 15, 3)), float(rng.uniform(-4, 4)),
                     float(rng.uniform(-2, 2)), float(rng.uniform(20, 120)), float(rng.choice([0.01, 0.05, 0.1])),
                     int(rng.integers(16, 4096)), int(rng.integers(16, 4096))))
    return cams


def test_device_buffers_equal_the_oracle(device, oracle):
    cams = _cameras(257, 11)
    print(f"case: {len(cams)} camera buffers", flush=True)
    out = torch.zeros((len(cams), 48), dtype=torch.float32, device="cuda")
    device.camera_constant_buffers(cams, out.data_ptr())
    device.synchronize()
    got = out.cpu().numpy()
    want = np.stack([oracle.camera_cb(*c) for c in cams])
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_frame_with_a_device_buffer(device, oracle):
    W, H = 320, 184
    mesh = S.sphere_mesh(10, displace=0.15, seed=6)
    tex = S.random_rgba8(64, seed=2)
    cam = ((3.0, 0.4, 0.6), 0.15, -0.2, 60.0, 0.01, W, H)
    out = torch.zeros(48, dtype=torch.float32, device="cuda")
    device.camera_constant_buffers([cam], out.data_ptr())
    device.synchronize()
    cb_dev = out.cpu().numpy()
    cb_ref = oracle.camera_cb(*cam)
    np.testing.assert_array_equal(cb_dev.view(np.uint32), cb_ref.view(np.uint32))
    print("case: frame drawn with a device-built constant buffer", flush=True)
    g = render_gpu(device, W, H, [DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb_dev, ps_tex=tex)],
                   S.CLEAR_COLOR)
    o = render_oracle(oracle, W, H, [DrawSpec(mesh.indices, mesh.vertices, "basic", "basic", cb=cb_ref, ps_tex=tex)],
                      S.CLEAR_COLOR)
    assert_frames_match(g, o, color_tol=1)
    assert (g.vis_tri != 0xFFFFFFFF).sum() > 500
