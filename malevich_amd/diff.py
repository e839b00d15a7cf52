"""Differentiable shading as a torch.autograd.Function over the C ABI.

Forward: the raster pipeline draws the mesh (basic_vs, reversed-Z GREATER_EQUAL, exactly the
reference's coverage and visibility) into a primitive-id buffer; mlv_diff_forward then shades
every covered pixel (perspective-correct interpolation of colour, UV and normal, bilinear
texture fetch on the nearest mip level, 9-coefficient SH light, basic_ps's sRGB curve).
Backward: mlv_diff_backward recomputes the shading from the primitive ids alone and returns
gradients for positions (through the barycentrics, visibility fixed), colours, UVs, normals,
every texel of the mip chain and the SH coefficients.

PyTorch only holds the tensors and the tape; all the work runs in the library's kernels on the
library's stream, ordered against torch's current stream with HIP events.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from . import pipeline as P


class _StreamOrder:
    """Makes one HIP stream wait for the work queued so far on another."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipStreamWaitEvent.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]
        self.events = []
        for _ in range(2):
            e = C.c_void_p()
            if self.hip.hipEventCreateWithFlags(C.byref(e), 2) != 0:  # hipEventDisableTiming
                raise RuntimeError("hipEventCreateWithFlags failed")
            self.events.append(e)

    def wait(self, waiter: int, producer: int, slot: int = 0) -> None:
        e = self.events[slot]
        if self.hip.hipEventRecord(e, C.c_void_p(producer)) != 0 or \
                self.hip.hipStreamWaitEvent(C.c_void_p(waiter), e, 0) != 0:
            raise RuntimeError("stream ordering failed")


def mip_chain(tex: torch.Tensor, renderer: "DiffRenderer") -> tuple[torch.Tensor, int]:
    """(h, w, 4) texture -> ((N, 4) chain of all levels by 2x2 box filtering, level count), built
    by the library on the device (mlv_mip_chain). Differentiable: gradients on coarse levels flow
    back to level 0 (mlv_mip_chain_adjoint)."""
    h, w = int(tex.shape[0]), int(tex.shape[1])
    n, levels = L.mip_chain_size(w, h)
    return MipChainFunction.apply(tex, renderer, n), levels


class MipChainFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tex, renderer, n):
        tex = tex.detach().contiguous().float()
        h, w = int(tex.shape[0]), int(tex.shape[1])
        chain = torch.empty((n, 4), dtype=torch.float32, device=tex.device)
        d = renderer.dev
        renderer._enter()
        P._check(d.ctx, d.lib.mlv_mip_chain(d.ctx, tex.data_ptr(), w, h, chain.data_ptr()))
        renderer._leave()
        ctx.renderer, ctx.size = renderer, (w, h)
        return chain

    @staticmethod
    def backward(ctx, grad_chain):
        w, h = ctx.size
        g = grad_chain.contiguous().float()
        out = torch.empty((h, w, 4), dtype=torch.float32, device=g.device)
        r = ctx.renderer
        d = r.dev
        r._enter()
        P._check(d.ctx, d.lib.mlv_mip_chain_adjoint(d.ctx, g.data_ptr(), w, h, out.data_ptr()))
        r._leave()
        return out, None, None


class DiffRenderer:
    """Raster + differentiable shading for one image size on one mlv context."""

    def __init__(self, device: P.Device, width: int, height: int, deterministic: bool = False):
        self.dev = device
        self.width, self.height = width, height
        self.deterministic = deterministic
        self.rt = device.create_render_target(width, height)
        self.ds = device.create_depth_stencil(width, height)
        self.prim = torch.empty((height, width), dtype=torch.int32, device="cuda")
        self.order = _StreamOrder()
        self._wrapped: dict = {}
        self.cb_cache: dict = {}  # the raster's constant buffer per camera matrix

    def _wrap(self, kind: int, t: torch.Tensor) -> P.Buffer:
        key = (kind, t.data_ptr(), t.numel() * t.element_size())
        b = self._wrapped.get(key)
        if b is None:
            for k in [k for k in self._wrapped if k[0] == kind]:
                self._wrapped.pop(k).release()
            b = self.dev.wrap_buffer(kind, t.data_ptr(), t.numel() * t.element_size())
            self._wrapped[key] = b
        return b

    def _enter(self) -> None:
        self.order.wait(self.dev.stream, torch.cuda.current_stream().cuda_stream, 0)

    def _leave(self) -> None:
        self.order.wait(torch.cuda.current_stream().cuda_stream, self.dev.stream, 1)

    def rasterize(self, vertices: torch.Tensor, indices: torch.Tensor, cb: torch.Tensor) -> torch.Tensor:
        """vertices (V, 8) float32 in basic_vs's layout, indices (T, 3) int32, cb (48,) float32."""
        d = self.dev
        self._enter()
        self._record_raster(vertices, indices, cb)
        d.flush()
        d.om_set_primitive_id_buffer(None)
        self._leave()
        return self.prim

    def autotune(self, vertices: torch.Tensor, indices: torch.Tensor, cb: torch.Tensor, repeats: int = 3) -> dict:
        """Chooses the context's launch parameters for this raster (mlv_autotune_begin/end): the
        tile stage's form for few tiles with long lists differs from a full-screen frame's."""
        d = self.dev
        self._enter()
        d.autotune_begin()
        self._record_raster(vertices, indices, cb)
        best = d.autotune_end(repeats=repeats)
        d.om_set_primitive_id_buffer(None)
        self._leave()
        return best

    def _record_raster(self, vertices: torch.Tensor, indices: torch.Tensor, cb: torch.Tensor) -> None:
        d = self.dev
        vb = self._wrap(L.BUFFER_VERTEX, vertices)
        ib = self._wrap(L.BUFFER_INDEX_U32, indices)
        cbuf = self._wrap(L.BUFFER_CONSTANT, cb)
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.ia_set_vertex_buffer(vb, 32)
        d.ia_set_index_buffer(ib)
        d.vs_set_shader(L.VS_BASIC)
        d.vs_set_constant_buffer(0, cbuf)
        d.ps_set_shader(L.PS_NONE)  # visibility only: the shading recomputes from the primitive ids
        d.rs_set_viewport(0, 0, self.width, self.height, 0.0, 1.0)
        d.om_set_render_targets(self.rt, self.ds)
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        d.om_set_primitive_id_buffer(self.prim.data_ptr())
        d.clear_render_target_view(self.rt, (0.0, 0.0, 0.0, 0.0))
        d.clear_depth_stencil_view(self.ds, 0.0)
        d.draw_indexed(indices.numel())

    def scene_struct(self, positions, colors, uvs, normals, indices, chain, tex_w, tex_h, levels, sh,
                     clip_from_world) -> L.DiffScene:
        m = (C.c_float * 16)(*[float(x) for x in clip_from_world.reshape(-1).tolist()])
        return L.DiffScene(positions.data_ptr(), colors.data_ptr(), uvs.data_ptr(), normals.data_ptr(),
                           indices.data_ptr(), positions.shape[0], indices.shape[0], chain.data_ptr(), tex_w, tex_h,
                           levels, sh.data_ptr(), m, self.width, self.height)

    def shade(self, sc: L.DiffScene, prim: torch.Tensor) -> torch.Tensor:
        out = torch.empty((self.height, self.width, 4), dtype=torch.float32, device="cuda")
        self._enter()
        P._check(self.dev.ctx, self.dev.lib.mlv_diff_forward(self.dev.ctx, C.byref(sc), prim.data_ptr(),
                                                             out.data_ptr()))
        self._leave()
        return out

    def shade_backward(self, sc: L.DiffScene, prim: torch.Tensor, grad_out: torch.Tensor, want: dict) -> dict:
        g = {k: (torch.empty(shape, dtype=torch.float32, device="cuda") if shape is not None else None)
             for k, shape in want.items()}
        gs = L.DiffGrads(*[(g[k].data_ptr() if g[k] is not None else None)
                           for k in ("positions", "colors", "uvs", "normals", "texture", "sh")])
        self._enter()
        P._check(self.dev.ctx, self.dev.lib.mlv_diff_backward(self.dev.ctx, C.byref(sc), prim.data_ptr(),
                                                              grad_out.data_ptr(), C.byref(gs),
                                                              int(self.deterministic)))
        self._leave()
        return g

    def release(self) -> None:
        for b in self._wrapped.values():
            b.release()
        self._wrapped.clear()
        self.cb_cache.clear()
        self.rt.release()
        self.ds.release()


def raster_inputs(positions, normals, uvs, clip_from_world, cache: dict | None = None):
    """The raster's vertex buffer (basic_vs's layout: position, normal, uv) and constant buffer
    (clip_from_world, then two identity matrices). With a cache, the constant buffer is uploaded
    once per matrix: a host-to-device copy every step would make the host wait for the GPU."""
    M = clip_from_world.detach().to(torch.float32).reshape(4, 4).cpu()
    key = M.numpy().tobytes()
    cb = cache.get(key) if cache is not None else None
    if cb is None:
        eye = torch.eye(4, dtype=torch.float32)
        cb = torch.cat([M.reshape(-1), eye.reshape(-1), eye.reshape(-1)]).to(positions.device)
        if cache is not None:
            cache.clear()
            cache[key] = cb
    verts = torch.cat([positions.float(), normals.float(), uvs.float()], 1).detach().contiguous()
    return verts, cb


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, colors, uvs, normals, chain, sh, renderer, indices, clip_from_world, tex_w, tex_h,
                levels):
        positions, colors, uvs, normals = (t.contiguous().float() for t in (positions, colors, uvs, normals))
        chain, sh = chain.contiguous().float(), sh.contiguous().float()
        idx = indices.contiguous().to(torch.int32)
        M = clip_from_world.detach().to(torch.float32).reshape(4, 4).cpu()
        verts, cb = raster_inputs(positions, normals, uvs, M, renderer.cb_cache)
        prim = renderer.rasterize(verts, idx, cb).clone()
        sc = renderer.scene_struct(positions, colors, uvs, normals, idx, chain, tex_w, tex_h, levels, sh, M)
        out = renderer.shade(sc, prim)
        ctx.renderer = renderer
        ctx.meta = (tex_w, tex_h, levels, M)
        ctx.keep = (verts, cb)
        ctx.save_for_backward(positions, colors, uvs, normals, chain, sh, idx, prim)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        positions, colors, uvs, normals, chain, sh, idx, prim = ctx.saved_tensors
        tex_w, tex_h, levels, M = ctx.meta
        r = ctx.renderer
        sc = r.scene_struct(positions, colors, uvs, normals, idx, chain, tex_w, tex_h, levels, sh, M)
        need = ctx.needs_input_grad
        want = {"positions": positions.shape if need[0] else None, "colors": colors.shape if need[1] else None,
                "uvs": uvs.shape if need[2] else None, "normals": normals.shape if need[3] else None,
                "texture": chain.shape if need[4] else None, "sh": sh.shape if need[5] else None}
        g = r.shade_backward(sc, prim, grad_out.contiguous().float(), want)
        return (g["positions"], g["colors"], g["uvs"], g["normals"], g["texture"], g["sh"], None, None, None, None,
                None, None)


def render(renderer: DiffRenderer, positions, colors, uvs, normals, indices, texture, sh, clip_from_world):
    """Image (H, W, 4) of the mesh; differentiable in positions, colours, uvs, normals, the
    level-0 texture (through its box-filtered mip chain) and the SH coefficients."""
    chain, levels = mip_chain(texture, renderer)
    return RenderFunction.apply(positions, colors, uvs, normals, chain, sh, renderer, indices, clip_from_world,
                                texture.shape[1], texture.shape[0], levels)


class DiffStep:
    """One differentiable step (raster, shading, the L2 loss gradient, the adjoint) of a fixed
    mesh, camera and target, with the host work of a training loop's inner iteration cut to a few
    library calls: the vertex buffer packed from the current parameters (one torch op), the
    texture's mip chain, the raster frame as one graph launch (captured once), mlv_diff_forward,
    mlv_diff_l2_grad (the gradient of sum((image - target)^2): scale 2), mlv_diff_backward and the
    mip chain's adjoint. The same kernels as render() + backward; the parameters' gradients land
    in self.grads (positions, colors, uvs, normals, texture, sh)."""

    def __init__(self, renderer: DiffRenderer, positions, colors, uvs, normals, indices, texture, sh,
                 clip_from_world, target):
        self.r = r = renderer
        d = r.dev
        self.params = dict(positions=positions, colors=colors, uvs=uvs, normals=normals, texture=texture, sh=sh)
        self.idx = indices.contiguous().to(torch.int32)
        self.target = target.contiguous().float()
        h, w = int(texture.shape[0]), int(texture.shape[1])
        self.tex_w, self.tex_h = w, h
        n, self.levels = L.mip_chain_size(w, h)
        self.chain = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        self.verts, self.cb = raster_inputs(positions, normals, uvs,
                                            clip_from_world.detach().to(torch.float32).reshape(4, 4).cpu(), r.cb_cache)
        self.image = torch.empty((r.height, r.width, 4), dtype=torch.float32, device="cuda")
        self.grad_image = torch.empty_like(self.image)
        self.grad_chain = torch.empty_like(self.chain)
        self.grads = {k: torch.empty_like(v) for k, v in self.params.items()}
        g = self.grads
        self.gstruct = L.DiffGrads(g["positions"].data_ptr(), g["colors"].data_ptr(), g["uvs"].data_ptr(),
                                   g["normals"].data_ptr(), self.grad_chain.data_ptr(), g["sh"].data_ptr())
        self.scene = r.scene_struct(positions, colors, uvs, normals, self.idx, self.chain, w, h, self.levels, sh,
                                    clip_from_world.detach().to(torch.float32).reshape(4, 4).cpu())
        r._enter()
        d.graph_capture_begin()
        r._record_raster(self.verts, self.idx, self.cb)
        self.raster = d.graph_capture_end()
        d.om_set_primitive_id_buffer(None)
        r._leave()

    def __call__(self) -> None:
        r, d, p = self.r, self.r.dev, self.params
        torch.cat([p["positions"].detach(), p["normals"].detach(), p["uvs"].detach()], 1, out=self.verts)
        r._enter()
        lib, ctx = d.lib, d.ctx
        P._check(ctx, lib.mlv_mip_chain(ctx, p["texture"].data_ptr(), self.tex_w, self.tex_h, self.chain.data_ptr()))
        self.raster.launch()
        P._check(ctx, lib.mlv_diff_forward(ctx, C.byref(self.scene), r.prim.data_ptr(), self.image.data_ptr()))
        P._check(ctx, lib.mlv_diff_l2_grad(ctx, self.image.data_ptr(), self.target.data_ptr(), self.image.numel(), 2.0,
                                           self.grad_image.data_ptr()))
        P._check(ctx, lib.mlv_diff_backward(ctx, C.byref(self.scene), r.prim.data_ptr(), self.grad_image.data_ptr(),
                                            C.byref(self.gstruct), int(r.deterministic)))
        P._check(ctx, lib.mlv_mip_chain_adjoint(ctx, self.grad_chain.data_ptr(), self.tex_w, self.tex_h,
                                                self.grads["texture"].data_ptr()))
        r._leave()

    def release(self) -> None:
        self.raster.release()
