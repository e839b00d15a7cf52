"""The benchmark's workloads at their full shape against the oracle.

- The 1080p forward frame: 64 displaced spheres (2.49 M triangles), vertex_lighting_vs with a
  256x128 RGBA16F irradiance map and env_lighting_ps with a 2048x1024 RGBA16F radiance map; and
  the same frame as bench.py draws it, with the textured environment shader on mip chains (the
  radiance map's and a 1024x1024 RGBA8 albedo per instance). Visibility, depth and stats
  bit-exact; colour within 1/255.
- Batched views at the batched workload's shape: the first 8 views (256^2 to 1024^2) of the
  benchmark's 256-view scene, each drawing all 64 meshes (1k-500k triangles, 5.0 M per view),
  with the flush limits lowered (MLV_MAX_BATCH_TRIANGLES, MLV_MAX_BATCH_DRAWS) so the 512 draws
  run in several flushes whose boundaries fall inside views. Every frame against the oracle;
  the batched differentiable shading of the same views against the FP64 oracle (images to 1e-3
  per value and 1e-5 RMS, every mesh's gradients to 1e-4 relative).
"""
from __future__ import annotations

import copy
import os

import numpy as np
import pytest

from malevich_amd import _lib as L
from malevich_amd import pipeline as P
from malevich_amd import scenes as S
from tests.diff_utils import NONE, rel_err
from tests.render_utils import DrawSpec, Frame, assert_frames_match, channel_diff, render_oracle
from tests.test_gpu_batched import render_views

pytestmark = pytest.mark.gpu

C2_SEED = 2          # bench.py --seed default
C5_SEED = 2 + 5      # bench.py's batched scene seed
BATCH_TRIS = 12_000_000
BATCH_DRAWS = 96


def test_c2_full_frame_against_oracle(device, oracle):
    W, H = 1920, 1080
    sc = S.c2_scene(W, H, n_instances=64, seed=C2_SEED, texture_size=0)
    irr = S.smooth_env_map(256, 128, seed=C2_SEED + 16, lo=0.05, hi=2.0).astype(np.float16)
    rad = S.smooth_env_map(2048, 1024, seed=C2_SEED + 17, lo=0.05, hi=2.0).astype(np.float16)
    print(f"case: c2 full frame {W}x{H}, {len(sc.instances)} instances, {sc.triangle_count} triangles", flush=True)
    dev = device
    keep = [dev.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.instances]
    keep += [dev.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.instances]
    cbuf = dev.create_buffer(L.BUFFER_CONSTANT, sc.cb)
    t_irr, t_rad = dev.create_texture2d(irr), dev.create_texture2d(rad)
    rt, ds = dev.create_render_target(W, H), dev.create_depth_stencil(W, H)
    vt = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    vd = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    try:
        dev.reset_stats()
        dev.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        dev.vs_set_shader(L.VS_VERTEX_LIGHTING)
        dev.ps_set_shader(L.PS_ENV_LIGHTING)
        dev.vs_set_constant_buffer(0, cbuf)
        dev.vs_set_shader_resource(0, t_irr)
        dev.ps_set_shader_resource(0, t_rad)
        dev.rs_set_viewport(0, 0, W, H, 0.0, 1.0)
        dev.om_set_render_targets(rt, ds)
        dev.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        dev.om_set_visibility_buffer(vt, vd)
        dev.clear_render_target_view(rt, S.CLEAR_COLOR)
        dev.clear_depth_stencil_view(ds, 0.0)
        n = len(sc.instances)
        for vb, ib, m in zip(keep[:n], keep[n:], sc.instances):
            dev.ia_set_vertex_buffer(vb, 32)
            dev.ia_set_index_buffer(ib)
            dev.draw_indexed(m.indices.size)
        g = Frame(rt.read(), ds.read(), vt.read(np.uint32).reshape(H, W), vd.read(np.uint32).reshape(H, W),
                  dev.stats())
    finally:
        dev.om_set_visibility_buffer(None, None)
        dev.vs_set_shader_resource(0, None)
        dev.ps_set_shader_resource(0, None)
        for h in keep + [cbuf, t_irr, t_rad, rt, ds, vt, vd]:
            h.release()
    irr32, rad32 = irr.astype(np.float32), rad.astype(np.float32)  # RGBA16F texels are exact in float32
    draws = [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_lighting", cb=sc.cb, vs_tex=irr32,
                      ps_tex=rad32) for m in sc.instances]
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    covered = int((g.vis_tri != 0xFFFFFFFF).sum())
    print(f"  covered {covered} px, assembled {g.stats['assembled_triangle_count']}, colour max diff "
          f"{channel_diff(g.colors, o.colors)}", flush=True)
    assert covered > 0.5 * W * H
    assert_frames_match(g, o, color_tol=1)


def test_c2_full_textured_frame_against_oracle(device, oracle):
    """The benchmark's headline frame as configured: the same 64 spheres with the textured
    environment shader, a 2048x1024 RGBA16F radiance chain (12 levels) and a 1024x1024 RGBA8 albedo
    chain (11 levels) per instance, the chains built on the device as bench.py builds them."""
    W, H = 1920, 1080
    sc = S.c2_scene(W, H, n_instances=64, seed=C2_SEED, texture_size=0)
    irr, rad, albedo = S.c2_textures(64, seed=C2_SEED, albedo_size=1024)
    print(f"case: c2 full textured frame {W}x{H}, {len(sc.instances)} instances", flush=True)
    dev = device
    keep = [dev.create_buffer(L.BUFFER_VERTEX, m.vertices) for m in sc.instances]
    keep += [dev.create_buffer(L.BUFFER_INDEX_U32, m.indices) for m in sc.instances]
    cbuf = dev.create_buffer(L.BUFFER_CONSTANT, sc.cb)
    t_irr = dev.create_texture2d(irr)
    t_rad = dev.create_texture2d(rad, mip_levels=12, generate_mips=True)
    t_alb = [dev.create_texture2d(a, mip_levels=11, generate_mips=True) for a in albedo]
    rt, ds = dev.create_render_target(W, H), dev.create_depth_stencil(W, H)
    vt = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    vd = dev.create_buffer(L.BUFFER_VERTEX, nbytes=W * H * 4)
    try:
        rad_chain = t_rad.read_levels()
        alb_chains = [t.read_levels() for t in t_alb]
        dev.reset_stats()
        dev.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        dev.vs_set_shader(L.VS_VERTEX_LIGHTING)
        dev.ps_set_shader(L.PS_ENV_TEXTURED)
        dev.vs_set_constant_buffer(0, cbuf)
        dev.vs_set_shader_resource(0, t_irr)
        dev.ps_set_shader_resource(0, t_rad)
        dev.rs_set_viewport(0, 0, W, H, 0.0, 1.0)
        dev.om_set_render_targets(rt, ds)
        dev.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        dev.om_set_visibility_buffer(vt, vd)
        dev.clear_render_target_view(rt, S.CLEAR_COLOR)
        dev.clear_depth_stencil_view(ds, 0.0)
        n = len(sc.instances)
        for vb, ib, m, ta in zip(keep[:n], keep[n:], sc.instances, t_alb):
            dev.ia_set_vertex_buffer(vb, 32)
            dev.ia_set_index_buffer(ib)
            dev.ps_set_shader_resource(1, ta)
            dev.draw_indexed(m.indices.size)
        g = Frame(rt.read(), ds.read(), vt.read(np.uint32).reshape(H, W), vd.read(np.uint32).reshape(H, W),
                  dev.stats())
    finally:
        dev.om_set_visibility_buffer(None, None)
        dev.vs_set_shader_resource(0, None)
        dev.ps_set_shader_resource(0, None)
        dev.ps_set_shader_resource(1, None)
        for h in keep + t_alb + [cbuf, t_irr, t_rad, rt, ds, vt, vd]:
            h.release()
    draws = [DrawSpec(m.indices, m.vertices, "vertex_lighting", "env_textured", cb=sc.cb, vs_tex=irr,
                      ps_tex=rad_chain, ps_tex1=ch) for m, ch in zip(sc.instances, alb_chains)]
    o = render_oracle(oracle, W, H, draws, S.CLEAR_COLOR)
    print(f"  covered {int((g.vis_tri != 0xFFFFFFFF).sum())} px, colour max diff {channel_diff(g.colors, o.colors)}",
          flush=True)
    assert_frames_match(g, o, color_tol=1)


@pytest.fixture(scope="module")
def c5_scene():
    sc = S.batched_scene(n_views=8, seed=C5_SEED)
    assert {v.width for v in sc.views} >= {256, 1024}, [v.width for v in sc.views]
    return sc


@pytest.fixture(scope="module")
def split_device():
    """A context whose flushes take at most BATCH_TRIS triangles and BATCH_DRAWS draws."""
    old = {k: os.environ.get(k) for k in ("MLV_MAX_BATCH_TRIANGLES", "MLV_MAX_BATCH_DRAWS")}
    os.environ["MLV_MAX_BATCH_TRIANGLES"] = str(BATCH_TRIS)
    os.environ["MLV_MAX_BATCH_DRAWS"] = str(BATCH_DRAWS)
    try:
        dev = P.Device(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    yield dev
    dev.close()


def test_c5_scale_views_in_several_flushes_against_oracle(split_device, oracle, c5_scene):
    sc = c5_scene
    nm = len(sc.meshes)
    tris = sum(m.indices.size // 3 for m in sc.meshes)
    print(f"case: c5-scale raster {len(sc.views)} views {[v.width for v in sc.views]} x {nm} meshes "
          f"({tris} triangles per view)", flush=True)
    split_device.set_profiling(True)
    split_device.profile()
    frames, stats = render_views(split_device, sc, per_draw=False)
    prof = split_device.profile()
    split_device.set_profiling(False)
    # render_views reads back view by view; the draws of all views were recorded first, so the
    # flushes are the limits' splits of the 512 draws
    print(f"  flushes: {prof['flushes']}", flush=True)
    assert prof["flushes"] >= (len(sc.views) * tris) // BATCH_TRIS
    total = {k: 0 for k in ("assembled_triangle_count", "total_triangle_count_in_bins", "active_bin_count")}
    for i, (v, g) in enumerate(zip(sc.views, frames)):
        draws = [DrawSpec(m.indices, m.vertices, "basic", "basic", cb=v.cb, ps_tex=t)
                 for m, t in zip(sc.meshes, sc.textures)]
        o = render_oracle(oracle, v.width, v.height, draws, S.CLEAR_COLOR)
        g.vis_draw = np.where(g.vis_draw == 0xFFFFFFFF, g.vis_draw, g.vis_draw - i * nm).astype(np.uint32)
        assert_frames_match(g, o, color_tol=1, check_stats=False)
        for k in total:
            total[k] += o.stats[k]
        print(f"  view {i} ({v.width}^2): {int((g.vis_tri != 0xFFFFFFFF).sum())} px covered, exact", flush=True)
    for k, val in total.items():
        assert stats[k] == val, (k, stats[k], val)


def test_c5_scale_batched_diff_in_several_flushes_against_fp64(split_device, oracle, c5_scene):
    import torch

    from malevich_amd import batched as B
    sc = c5_scene
    nm, nv = len(sc.meshes), len(sc.views)
    print(f"case: c5-scale batched fwd+bwd {nv} views x {nm} meshes", flush=True)
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    inputs, meshes = [], []
    for k in range(nm):
        dm = sc.diff_mesh(k)
        chain, levels = S.mip_chain(sc.textures[k])
        h, w = sc.textures[k].shape[:2]
        inputs.append(oracle.DiffInputs(dm.positions, dm.colors, dm.uvs, dm.normals, dm.indices, chain, w, h,
                                        levels, sc.sh, sc.views[0].cb[:16].reshape(4, 4), 1, 1))
        meshes.append(B.MeshTensors(cuda(dm.positions), cuda(dm.colors), cuda(dm.uvs), cuda(dm.normals),
                                    cuda(dm.indices.astype(np.int32)), cuda(chain), w, h, levels))
    views = [B.ViewTensors(v.cb, v.width, v.height) for v in sc.views]

    def for_view(d, v):
        e = copy.copy(d)
        e.M_f = np.ascontiguousarray(v.cb[:16], np.float32).reshape(16)
        e.M = e.M_f.astype(np.float64)
        e.width, e.height = v.width, v.height
        return e

    r = B.BatchedDiffRenderer(split_device, meshes, views)
    try:
        sh = cuda(sc.sh)
        r.rasterize()
        out = r.forward(sh)
        torch.cuda.synchronize()
        prim = r.prim.cpu().numpy().view(np.uint32)
        tags = r.tags.cpu().numpy().view(np.uint32)
        out_g = out.cpu().numpy()
        out_o = np.zeros((len(prim), 4))
        masks = []
        for v, w in enumerate(sc.views):
            o0, o1 = int(r.offsets[v]), int(r.offsets[v + 1])
            pv, tv = prim[o0:o1], tags[o0:o1]
            covered = pv != NONE
            mesh_of = np.where(covered, tv.astype(np.int64) - v * nm, -1)
            assert np.all((mesh_of[covered] >= 0) & (mesh_of[covered] < nm))
            row = []
            for k in range(nm):
                sel = mesh_of == k
                masked = np.where(sel, pv, NONE).astype(np.uint32).reshape(w.height, w.width)
                row.append(masked if sel.any() else None)
                if sel.any():
                    out_o[o0:o1] += oracle.diff_forward(for_view(inputs[k], w), masked).reshape(-1, 4)
            masks.append(row)
            print(f"  view {v}: forward oracle done", flush=True)
        err = np.abs(out_g - out_o)
        emax, erms = float(err.max()), float(np.sqrt((err * err).mean()))
        print(f"  forward: max abs err {emax:.2e}, rms {erms:.2e}", flush=True)
        assert emax <= 1e-3 and erms <= 1e-5, (emax, erms)

        rng = np.random.default_rng(5)
        target = np.clip(out_o + rng.normal(0, 0.05, out_o.shape), 0, 1)
        g_out_np = 2.0 * (out_o - target) / out_o.shape[0]
        g_out = torch.from_numpy(g_out_np.astype(np.float32)).cuda()
        grads = r.new_grads()
        grad_sh = torch.empty((9, 3), dtype=torch.float32, device="cuda")
        r.backward(sh, g_out, grads, grad_sh)
        torch.cuda.synchronize()
        ref_sh = np.zeros((9, 3))
        worst = 0.0
        for k in range(nm):
            ref = None
            for v, w in enumerate(sc.views):
                if masks[v][k] is None:
                    continue
                o0, o1 = int(r.offsets[v]), int(r.offsets[v + 1])
                go = oracle.diff_backward(for_view(inputs[k], w), masks[v][k],
                                          g_out_np[o0:o1].reshape(w.height, w.width, 4))
                ref_sh += go.pop("sh")
                ref = go if ref is None else {n: ref[n] + go[n] for n in go}
            gk = {n: t.cpu().numpy() for n, t in grads[k].items()}
            if ref is None:
                for n in gk:
                    assert not np.any(gk[n]), (k, n)
                continue
            for n in ("positions", "colors", "uvs", "normals", "texture"):
                e = rel_err(gk[n], ref[n])
                worst = max(worst, e)
                assert e <= 1e-4, (k, n, e)
            if k % 16 == 15:
                print(f"  meshes up to {k}: worst gradient rel err {worst:.2e}", flush=True)
        e = rel_err(grad_sh.cpu().numpy(), ref_sh)
        print(f"  gradients: worst mesh rel err {worst:.2e}, sh {e:.2e}", flush=True)
        assert e <= 1e-4, e
    finally:
        r.release()
