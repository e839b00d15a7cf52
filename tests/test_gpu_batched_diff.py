"""Batched differentiable shading (many views of many meshes in one launch) against the FP64
oracle, and against the per-draw launch form of the same kernels.

Bar (written here): every view's image agrees with the oracle's FP64 forward to 1e-3 per value
and 1e-5 RMS; every mesh's gradient groups (positions, colours, UVs, normals, texels) and the
shared SH gradient have ||g - g64|| / ||g64|| <= 1e-4, g64 being the oracle's FP64 adjoint summed
over views; the per-draw form gives the identical image and gradients equal to 1e-6 relative
(the atomics add in another order).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from malevich_amd import batched as B
from malevich_amd import scenes as S
from tests.diff_utils import NONE, rel_err

pytestmark = pytest.mark.gpu

GROUPS = ("positions", "colors", "uvs", "normals", "texture")


@pytest.fixture(scope="module", params=["dense", "coarse"])
def scene(request):
    if request.param == "coarse":  # triangles of tens to hundreds of pixels (the wave-sum path)
        return S.batched_scene(n_views=3, n_meshes=8, min_tris=20, max_tris=60, resolutions=(192, 256),
                               tex_sizes=(32,), seed=12)
    return S.batched_scene(n_views=5, n_meshes=4, min_tris=300, max_tris=5000, resolutions=(64, 96, 128),
                           tex_sizes=(32, 64), seed=11)


def _mesh_inputs(sc):
    out = []
    for k in range(len(sc.meshes)):
        dm = sc.diff_mesh(k)
        chain, levels = S.mip_chain(sc.textures[k])
        out.append((dm, chain, levels))
    return out


def _renderer(device, sc, inputs):
    cuda = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    meshes = [B.MeshTensors(cuda(dm.positions), cuda(dm.colors), cuda(dm.uvs), cuda(dm.normals),
                            cuda(dm.indices.astype(np.int32)), cuda(chain), sc.textures[k].shape[1],
                            sc.textures[k].shape[0], levels)
              for k, (dm, chain, levels) in enumerate(inputs)]
    views = [B.ViewTensors(v.cb, v.width, v.height) for v in sc.views]
    return B.BatchedDiffRenderer(device, meshes, views)


def test_batched_diff_against_oracle_and_per_draw(device, oracle, scene):
    sc = scene
    nm, nv = len(sc.meshes), len(sc.views)
    print(f"case: batched diff {nv} views x {nm} meshes", flush=True)
    inputs = _mesh_inputs(sc)
    r = _renderer(device, sc, inputs)
    try:
        sh = torch.from_numpy(sc.sh).cuda()
        r.rasterize()
        out = r.forward(sh)
        torch.cuda.synchronize()
        prim = r.prim.cpu().numpy().view(np.uint32)
        tags = r.tags.cpu().numpy().view(np.uint32)
        out_g = out.cpu().numpy()

        # the oracle, one (view, mesh) pair at a time on that mesh's pixels
        rng = np.random.default_rng(3)
        out_o = np.zeros((len(prim), 4))
        masks = []
        for v, w in enumerate(sc.views):
            o0, o1 = int(r.offsets[v]), int(r.offsets[v + 1])
            pv, tv = prim[o0:o1], tags[o0:o1]
            covered = pv != NONE
            mesh_of = np.where(covered, tv.astype(np.int64) - v * nm, -1)
            assert np.all((mesh_of[covered] >= 0) & (mesh_of[covered] < nm))
            row = []
            for k, (dm, chain, levels) in enumerate(inputs):
                sel = mesh_of == k
                assert np.all(pv[sel] < dm.triangle_count)
                masked = np.where(sel, pv, NONE).astype(np.uint32).reshape(w.height, w.width)
                row.append(masked)
                if sel.any():
                    d = oracle.DiffInputs(dm.positions, dm.colors, dm.uvs, dm.normals, dm.indices, chain,
                                          sc.textures[k].shape[1], sc.textures[k].shape[0], levels, sc.sh,
                                          w.cb[:16].reshape(4, 4), w.width, w.height)
                    out_o[o0:o1] += oracle.diff_forward(d, masked).reshape(-1, 4)
            masks.append(row)
        assert (prim != NONE).sum() > 0.02 * len(prim)  # meshes spread over a 20-unit cube seen from 24-34
        err = np.abs(out_g - out_o)
        emax, erms = float(err.max()), float(np.sqrt((err * err).mean()))
        print(f"  forward: max abs err {emax:.2e}, rms {erms:.2e}", flush=True)
        assert emax <= 1e-3 and erms <= 1e-5, (emax, erms)

        target = np.clip(out_o + rng.normal(0, 0.05, out_o.shape), 0, 1)
        g_out_np = 2.0 * (out_o - target)
        g_out = torch.from_numpy(g_out_np.astype(np.float32)).cuda()
        grads = r.new_grads()
        grad_sh = torch.empty((9, 3), dtype=torch.float32, device="cuda")
        r.backward(sh, g_out, grads, grad_sh)
        torch.cuda.synchronize()
        gb = [{k: t.cpu().numpy() for k, t in g.items()} for g in grads]
        gsh = grad_sh.cpu().numpy()

        ref = [None] * nm
        ref_sh = np.zeros((9, 3))
        for v, w in enumerate(sc.views):
            o0, o1 = int(r.offsets[v]), int(r.offsets[v + 1])
            for k, (dm, chain, levels) in enumerate(inputs):
                if not (masks[v][k] != NONE).any():
                    continue
                d = oracle.DiffInputs(dm.positions, dm.colors, dm.uvs, dm.normals, dm.indices, chain,
                                      sc.textures[k].shape[1], sc.textures[k].shape[0], levels, sc.sh,
                                      w.cb[:16].reshape(4, 4), w.width, w.height)
                go = oracle.diff_backward(d, masks[v][k], g_out_np[o0:o1].reshape(w.height, w.width, 4))
                ref_sh += go.pop("sh")
                ref[k] = go if ref[k] is None else {n: ref[k][n] + go[n] for n in go}
        for k in range(nm):
            if ref[k] is None:
                for n in GROUPS:
                    assert not np.any(gb[k][n]), (k, n)
                continue
            for n in GROUPS:
                e = rel_err(gb[k][n], ref[k][n])
                print(f"  mesh {k} {n}: rel err {e:.2e}", flush=True)
                assert e <= 1e-4, (k, n, e)
        e = rel_err(gsh, ref_sh)
        print(f"  sh: rel err {e:.2e}", flush=True)
        assert e <= 1e-4, e

        # per-draw form: same kernels, one (view, mesh) launch at a time
        out_p = r.forward_per_draw(sh)
        grads_p = r.new_grads()
        grad_sh_p = torch.empty((9, 3), dtype=torch.float32, device="cuda")
        r.backward_per_draw(sh, g_out, grads_p, grad_sh_p, r.new_accum())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_p.cpu().numpy().view(np.uint32), out_g.view(np.uint32))
        for k in range(nm):
            for n in GROUPS:
                e = rel_err(grads_p[k][n].cpu().numpy(), gb[k][n])
                assert e <= 1e-6, ("per-draw", k, n, e)
        assert rel_err(grad_sh_p.cpu().numpy(), gsh) <= 1e-6
    finally:
        r.release()


def test_captured_step_replays_the_eager_step(device, scene):
    """One HIP graph of the whole step (raster, shading, loss gradient, adjoint). Bar: the image and
    the loss gradient of the capture run and of every replay equal the eager step's bit for bit;
    the parameter gradients agree to 1e-6 relative (atomics add in another order); replays are
    separated by eager work and synchronisations."""
    sc = scene
    print(f"case: captured step {len(sc.views)} views x {len(sc.meshes)} meshes", flush=True)
    r = _renderer(device, sc, _mesh_inputs(sc))
    g = None
    try:
        sh = torch.from_numpy(sc.sh).cuda()
        gen = torch.Generator(device="cuda").manual_seed(5)
        r.rasterize()
        out = r.forward(sh)
        target = torch.rand(out.shape, device="cuda", generator=gen)
        scale = 2.0 / out.numel()
        gout = torch.sub(out, target).mul_(scale)
        grads = r.new_grads()
        grad_sh = torch.empty((9, 3), dtype=torch.float32, device="cuda")
        r.backward(sh, gout, grads, grad_sh)
        torch.cuda.synchronize()
        ref = (out.clone(), gout.clone(), [{k: t.clone() for k, t in d.items()} for d in grads], grad_sh.clone())

        out2, gout2 = torch.zeros_like(out), torch.zeros_like(gout)
        grads2 = r.new_grads()
        grad_sh2 = torch.zeros((9, 3), dtype=torch.float32, device="cuda")
        g = r.capture_step(sh, out2, target, scale, gout2, grads2, grad_sh2)

        def check(tag):
            torch.cuda.synchronize()
            assert torch.equal(out2.view(torch.int32), ref[0].view(torch.int32)), tag
            assert torch.equal(gout2.view(torch.int32), ref[1].view(torch.int32)), tag
            for k in range(len(sc.meshes)):
                for n in GROUPS:
                    e = rel_err(grads2[k][n].cpu().numpy(), ref[2][k][n].cpu().numpy())
                    assert e <= 1e-6, (tag, k, n, e)
            assert rel_err(grad_sh2.cpu().numpy(), ref[3].cpu().numpy()) <= 1e-6, tag

        check("capture run")
        for rep in range(3):
            out2.zero_()
            gout2.zero_()
            for d in grads2:
                for t in d.values():
                    t.fill_(7.0)
            r.forward(sh, torch.empty_like(out))  # eager work between replays
            torch.cuda.synchronize()
            r.run_graph(g)
            check(f"replay {rep}")
    finally:
        if g is not None:
            g.release()
        r.release()
