"""Differentiable shading on the GPU against the FP64 oracle.

Bar (written here): primitive ids equal the oracle raster's; the forward image agrees with the
FP64 forward to 1e-3 per value and 1e-5 RMS (f32 shading against double; the worst pixels are
slivers whose barycentrics are ill-conditioned); every gradient group (positions, colours,
UVs, normals, texels, SH) has a relative error ||g - g64|| / ||g64|| <= 1e-4 against the FP64
adjoint; the deterministic mode is bit-for-bit reproducible; the torch.autograd path returns
the same gradients, with the texel gradient carried back through the box-filtered mip chain.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from malevich_amd import diff as D
from malevich_amd import scenes as S
from tests.diff_utils import NONE, oracle_inputs, oracle_prim_ids, rel_err

pytestmark = pytest.mark.gpu

GROUPS = ("positions", "colors", "uvs", "normals", "texture", "sh")


def _tensors(sc):
    m = sc.mesh
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return dict(positions=cuda(m.positions), colors=cuda(m.colors), uvs=cuda(m.uvs), normals=cuda(m.normals),
                indices=cuda(m.indices.astype(np.int32)), chain=cuda(sc.chain), sh=cuda(sc.sh),
                texture=cuda(sc.texture), M=torch.from_numpy(sc.clip_from_world.copy()))


def _gpu_pass(device, sc, grad_out_np, deterministic=False):
    r = D.DiffRenderer(device, sc.width, sc.height, deterministic=deterministic)
    try:
        t = _tensors(sc)
        verts = torch.cat([t["positions"], t["normals"], t["uvs"]], 1).contiguous()
        eye = torch.eye(4).reshape(-1)
        cb = torch.cat([t["M"].reshape(-1), eye, eye]).cuda()
        prim = r.rasterize(verts, t["indices"], cb).clone()
        st = r.scene_struct(t["positions"], t["colors"], t["uvs"], t["normals"], t["indices"], t["chain"],
                            sc.texture.shape[1], sc.texture.shape[0], sc.levels, t["sh"], t["M"])
        out = r.shade(st, prim)
        g = None
        if grad_out_np is not None:
            go = torch.from_numpy(np.ascontiguousarray(grad_out_np, np.float32)).cuda()
            want = {k: (t[k if k != "texture" else "chain"].shape) for k in GROUPS}
            g = r.shade_backward(st, prim, go, want)
            g = {k: v.cpu().numpy() for k, v in g.items()}
        torch.cuda.synchronize()
        return prim.cpu().numpy().view(np.uint32), out.cpu().numpy(), g
    finally:
        r.release()


@pytest.fixture(scope="module")
def small_scene():
    return S.c4_scene(width=96, height=64, rings=16, segments=24, tex_size=64, seed=5)


@pytest.fixture(scope="module")
def full_scene():
    return S.c4_scene()


def _check(device, oracle, sc, deterministic):
    rng = np.random.default_rng(1)
    print(f"case: diff {sc.width}x{sc.height}, {sc.mesh.triangle_count} triangles, det={deterministic}", flush=True)
    prim_o = oracle_prim_ids(oracle, sc)
    d = oracle_inputs(oracle, sc)
    out_o = oracle.diff_forward(d, prim_o)
    target = np.clip(out_o + rng.normal(0, 0.05, out_o.shape), 0, 1)
    g_out = 2.0 * (out_o - target)
    prim_g, out_g, g = _gpu_pass(device, sc, g_out, deterministic)
    np.testing.assert_array_equal(prim_g, prim_o)
    assert (prim_g != NONE).sum() > 0.2 * sc.width * sc.height
    err = np.abs(out_g - out_o)
    emax, erms = float(err.max()), float(np.sqrt((err * err).mean()))
    print(f"  forward: max abs err {emax:.2e}, rms {erms:.2e}", flush=True)
    assert emax <= 1e-3 and erms <= 1e-5, (emax, erms)
    go = oracle.diff_backward(d, prim_o, g_out)
    for k in GROUPS:
        e = rel_err(g[k], go[k])
        print(f"  {k}: rel err {e:.2e}", flush=True)
        assert e <= 1e-4, (k, e)
    return g


def test_small_scene_atomic(device, oracle, small_scene):
    _check(device, oracle, small_scene, deterministic=False)


def test_coarse_mesh_atomic(device, oracle):
    """96 triangles of about 150 pixels each: most waves see one or two triangles, so the vertex
    gradients take the wave-sum path (lanes of one triangle added across the wave first)."""
    sc = S.c4_scene(width=128, height=128, rings=6, segments=8, tex_size=64, seed=7)
    _check(device, oracle, sc, deterministic=False)


def test_small_scene_deterministic(device, oracle, small_scene):
    g1 = _check(device, oracle, small_scene, deterministic=True)
    rng = np.random.default_rng(1)
    # a second run gives bit-identical gradients
    d = oracle_inputs(oracle, small_scene)
    prim_o = oracle_prim_ids(oracle, small_scene)
    out_o = oracle.diff_forward(d, prim_o)
    target = np.clip(out_o + rng.normal(0, 0.05, out_o.shape), 0, 1)
    _, _, g2 = _gpu_pass(device, small_scene, 2.0 * (out_o - target), True)
    for k in GROUPS:
        np.testing.assert_array_equal(g1[k].view(np.uint32), g2[k].view(np.uint32))


def test_c4_full_scene_gradients(device, oracle, full_scene):
    _check(device, oracle, full_scene, deterministic=False)


def test_c4_full_scene_gradients_deterministic(device, oracle, full_scene):
    _check(device, oracle, full_scene, deterministic=True)



@pytest.mark.parametrize("deterministic", [False, True])
def test_diff_step_equals_the_autograd_step(device, small_scene, deterministic):
    """DiffStep runs the kernels of render() + backward() with a few host calls (the raster as a
    graph launch). Bar: its gradients equal the autograd path's, bit for bit in deterministic mode
    and to 1e-6 relative with atomics (they add in another order), on the first call and after the
    parameters change in place."""
    sc = small_scene
    print(f"case: diff step det={deterministic}", flush=True)
    r = D.DiffRenderer(device, sc.width, sc.height, deterministic=deterministic)
    step = None
    try:
        t = _tensors(sc)
        params = {k: t[k].clone().requires_grad_(True) for k in ("positions", "colors", "uvs", "normals", "texture",
                                                                  "sh")}
        gen = torch.Generator(device="cuda").manual_seed(3)
        target = torch.rand((sc.height, sc.width, 4), device="cuda", generator=gen)
        step = D.DiffStep(r, params["positions"], params["colors"], params["uvs"], params["normals"], t["indices"],
                          params["texture"], params["sh"], t["M"], target)
        for rnd in range(2):
            if rnd == 1:
                with torch.no_grad():
                    params["positions"].mul_(1.001)
                    params["texture"].mul_(0.9)
            for v in params.values():
                v.grad = None
            img = D.render(r, params["positions"], params["colors"], params["uvs"], params["normals"], t["indices"],
                           params["texture"], params["sh"], t["M"])
            ((img - target) ** 2).sum().backward()
            step()
            torch.cuda.synchronize()
            for k, v in params.items():
                a, b = step.grads[k], v.grad
                if deterministic:
                    assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (rnd, k)
                else:
                    e = rel_err(a.cpu().numpy(), b.cpu().numpy())
                    assert e <= 1e-6, (rnd, k, e)
    finally:
        if step is not None:
            step.release()
        r.release()

def _box_adjoint_to_level0(chain_grad: np.ndarray, w: int, h: int, levels: int) -> np.ndarray:
    """Adjoint of the 2x2 box mip chain (power-of-two sizes): every level's gradient spread back
    to level 0, each coarse texel's gradient divided equally over its 4^l source texels."""
    g0 = np.zeros((h, w, 4))
    off = 0
    for l in range(levels):
        lw, lh = max(w >> l, 1), max(h >> l, 1)
        gl = chain_grad[off: off + lw * lh].reshape(lh, lw, 4)
        off += lw * lh
        g0 += np.repeat(np.repeat(gl, h // lh, 0), w // lw, 1) / float((h // lh) * (w // lw))
    return g0


def test_autograd_function(device, oracle, small_scene):
    sc = small_scene
    prim_o = oracle_prim_ids(oracle, sc)
    d = oracle_inputs(oracle, sc)
    out_o = oracle.diff_forward(d, prim_o)
    rng = np.random.default_rng(2)
    target = np.clip(out_o + rng.normal(0, 0.05, out_o.shape), 0, 1)
    go = oracle.diff_backward(d, prim_o, 2.0 * (out_o - target))
    r = D.DiffRenderer(device, sc.width, sc.height)
    try:
        t = _tensors(sc)
        params = {k: t[k].clone().requires_grad_(True) for k in ("positions", "colors", "uvs", "normals", "texture",
                                                                  "sh")}
        img = D.render(r, params["positions"], params["colors"], params["uvs"], params["normals"], t["indices"],
                       params["texture"], params["sh"], t["M"])
        loss = ((img - torch.from_numpy(target.astype(np.float32)).cuda()) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
    finally:
        r.release()
    for k in ("positions", "colors", "uvs", "normals", "sh"):
        e = rel_err(params[k].grad.cpu().numpy(), go[k])
        assert e <= 1e-4, (k, e)
    tex0 = _box_adjoint_to_level0(go["texture"], sc.texture.shape[1], sc.texture.shape[0], sc.levels)
    assert rel_err(params["texture"].grad.cpu().numpy(), tex0) <= 1e-4


def test_raster_autotune_keeps_visibility(device, oracle, small_scene):
    """DiffRenderer.autotune picks launch parameters from the candidate set for its raster; the
    visibility afterwards is the oracle's, as before."""
    sc = small_scene
    prim_o = oracle_prim_ids(oracle, sc)
    saved = device.tuning()
    r = D.DiffRenderer(device, sc.width, sc.height)
    try:
        t = _tensors(sc)
        verts, cb = D.raster_inputs(t["positions"], t["normals"], t["uvs"], t["M"])
        best = r.autotune(verts, t["indices"], cb, repeats=1)
        assert best["heavy_tile_pairs"] in (0, 32, 64, 192, 384, 768, 1536, 0xFFFFFFFF) and best["tile_stage"] in (128, 256)
        prim = r.rasterize(verts, t["indices"], cb).clone()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(prim.cpu().numpy().view(np.uint32), prim_o)
    finally:
        r.release()
        device.set_tuning(**saved)


def _f32_chain(tex: np.ndarray) -> np.ndarray:
    """The chain level by level in float32: each texel the sum of its window in row order, then
    divided by the window's area (numpy's float32 adds and divides round like the device's)."""
    levels, t = [tex], tex
    while t.shape[0] > 1 or t.shape[1] > 1:
        kh, kw = (2 if t.shape[0] > 1 else 1), (2 if t.shape[1] > 1 else 1)
        H, W = t.shape[0] // kh, t.shape[1] // kw
        s = np.zeros((H, W, 4), np.float32)
        for dy in range(kh):
            for dx in range(kw):
                s = s + t[dy:kh * H:kh, dx:kw * W:kw]
        t = (s / np.float32(kw * kh)).astype(np.float32)
        levels.append(t)
    return np.concatenate([x.reshape(-1, 4) for x in levels])


@pytest.mark.parametrize("size", [(64, 64), (37, 20), (1, 9), (128, 32), (600, 45), (1024, 1024), (33, 70)])
def test_mip_chain_and_adjoint(device, size):
    """mlv_mip_chain bit for bit against a float32 level-by-level build and to 1e-6 against the
    float64 box-filter chain, and mlv_mip_chain_adjoint against the adjoint PyTorch's autograd
    takes through the same chain of 2x2 average pools (odd edges dropped) to 1e-6 relative; sizes
    with odd and unit sides, several 32x32 tiles and more than five levels."""
    import torch.nn.functional as F

    w, h = size
    rng = np.random.default_rng(w * 100 + h)
    tex_np = rng.uniform(0.0, 1.0, (h, w, 4)).astype(np.float32)
    ref_chain, ref_levels = S.mip_chain(tex_np)
    r = D.DiffRenderer(device, 8, 8)
    try:
        tex = torch.from_numpy(tex_np).cuda().requires_grad_(True)
        chain, levels = D.mip_chain(tex, r)
        torch.cuda.synchronize()
        assert levels == ref_levels and chain.shape == ref_chain.shape
        got = chain.detach().cpu().numpy().reshape(-1, 4)
        np.testing.assert_array_equal(got.view(np.uint32), _f32_chain(tex_np).view(np.uint32))
        assert rel_err(chain.detach().cpu().numpy(), ref_chain) <= 1e-6
        g = torch.from_numpy(rng.normal(0.0, 1.0, ref_chain.shape).astype(np.float32)).cuda()
        (chain * g).sum().backward()
        torch.cuda.synchronize()
        # the same chain in PyTorch ops, differentiated by autograd
        t2 = torch.from_numpy(tex_np).cuda().requires_grad_(True)
        lv, t = [t2], t2.permute(2, 0, 1).unsqueeze(0)
        while t.shape[2] > 1 or t.shape[3] > 1:
            t = F.avg_pool2d(t, (2 if t.shape[2] > 1 else 1, 2 if t.shape[3] > 1 else 1))
            lv.append(t[0].permute(1, 2, 0))
        (torch.cat([x.reshape(-1, 4) for x in lv]) * g).sum().backward()
        assert rel_err(tex.grad.cpu().numpy(), t2.grad.cpu().numpy()) <= 1e-6
    finally:
        r.release()
