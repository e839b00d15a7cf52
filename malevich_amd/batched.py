"""Many views of many meshes: one raster flush, one shading launch, one adjoint launch.

Each view has its own camera and square-or-not image; every view draws the same meshes, each
mesh with its own RGBA32F mip chain; one SH light is shared. The raster pipeline records every
draw of every view and runs them in one flush (primitive-id and draw-tag buffers per view);
mlv_diff_batch_forward shades every covered pixel of every view in one launch, and
mlv_diff_batch_backward returns the gradients of sum(rgba * grad_rgba) over all views in one
launch, summed per mesh with atomics.

The per-draw form runs the same kernels one (view, mesh) pair at a time, each launch covering
that view's pixels, with a flush of the raster after every draw: the plain per-draw launch
pattern that batching replaces.

Tensors are torch CUDA tensors; the library's work runs on its own stream, ordered against
torch's current stream with HIP events.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import pipeline as P
from .diff import _StreamOrder

NONE = 0xFFFFFFFF
GRAD_KEYS = ("positions", "colors", "uvs", "normals", "texture")


class MeshTensors:
    """One mesh on the device: attributes, (T, 3) int32 indices, its RGBA32F mip chain."""

    def __init__(self, positions, colors, uvs, normals, indices, chain, tex_w: int, tex_h: int, levels: int):
        self.positions, self.colors, self.uvs, self.normals = positions, colors, uvs, normals
        self.indices = indices
        self.chain = chain
        self.tex_w, self.tex_h, self.levels = tex_w, tex_h, levels
        # basic_vs's vertex layout for the raster (pos3, normal3, uv2)
        self.vertices = torch.cat([positions, normals, uvs], 1).contiguous()

    @property
    def vertex_count(self) -> int:
        return self.positions.shape[0]

    @property
    def triangle_count(self) -> int:
        return self.indices.shape[0]


class ViewTensors:
    """One view: its PerFrameCB (clip_from_world first) and image size."""

    def __init__(self, cb: np.ndarray, width: int, height: int):
        self.cb_np = np.asarray(cb, np.float32).reshape(-1)
        self.clip_from_world = self.cb_np[:16].copy()
        self.width, self.height = width, height


class BatchedDiffRenderer:
    def __init__(self, device: P.Device, meshes: list, views: list):
        self.dev = device
        self.meshes = meshes
        self.views = views
        self.order = _StreamOrder()
        self.pixels = [v.width * v.height for v in views]
        self.offsets = np.concatenate([[0], np.cumsum(self.pixels)]).astype(np.int64)
        total = int(self.offsets[-1])
        # per-pixel buffers of all views, one allocation each
        self.prim = torch.empty(total, dtype=torch.int32, device="cuda")
        self.tri = torch.empty(total, dtype=torch.int32, device="cuda")
        self.tags = torch.empty(total, dtype=torch.int32, device="cuda")
        self.cb = torch.from_numpy(np.concatenate([v.cb_np for v in views])).cuda()
        self._handles = []
        wrap = lambda kind, t: self._keep(device.wrap_buffer(kind, t.data_ptr(), t.numel() * t.element_size()))
        self.vbs = [wrap(L.BUFFER_VERTEX, m.vertices) for m in meshes]
        self.ibs = [wrap(L.BUFFER_INDEX_U32, m.indices) for m in meshes]
        self.cbs = [device.wrap_buffer(L.BUFFER_CONSTANT, self.cb.data_ptr() + 192 * i, 192) for i in range(len(views))]
        self._handles += self.cbs
        self.targets = []
        for v, o in zip(views, self.offsets[:-1]):
            rt = device.create_render_target(v.width, v.height)
            ds = device.create_depth_stencil(v.width, v.height)
            n = v.width * v.height * 4
            tri = device.wrap_buffer(L.BUFFER_VERTEX, self.tri.data_ptr() + 4 * int(o), n)
            tag = device.wrap_buffer(L.BUFFER_VERTEX, self.tags.data_ptr() + 4 * int(o), n)
            self.targets.append((rt, ds, tri, tag))
        self._draw_map = [(C.c_uint32 * len(meshes))(*range(len(meshes)))]
        self._one_map = {}

    def _keep(self, h):
        self._handles.append(h)
        return h

    def _enter(self) -> None:
        self.order.wait(self.dev.stream, torch.cuda.current_stream().cuda_stream, 0)

    def _leave(self) -> None:
        self.order.wait(torch.cuda.current_stream().cuda_stream, self.dev.stream, 1)

    def image(self, flat: torch.Tensor, v: int) -> torch.Tensor:
        """View v's (H, W, 4) slice of a flat (total_pixels, 4) image."""
        o = int(self.offsets[v])
        return flat[o: o + self.pixels[v]].view(self.views[v].height, self.views[v].width, 4)

    # ---------------------------------------------------------------- raster
    def rasterize(self, per_draw: bool = False, recording: bool = False) -> None:
        """Draws every mesh into every view (basic_vs, no pixel shader: a depth-only pass, reversed-Z
        GREATER_EQUAL); primitive ids and draw tags land in self.prim / self.tags. Draw tags restart at 0: view v's draw of mesh k
        has tag v * len(meshes) + k. recording: inside a graph capture (no stats reset, no flush)."""
        d = self.dev
        self._enter()
        if not recording:
            d.reset_stats()
        d.ia_set_primitive_topology(L.TOPOLOGY_TRIANGLELIST)
        d.vs_set_shader(L.VS_BASIC)
        d.ps_set_shader(L.PS_NONE)  # the shading pass recomputes everything from the primitive ids
        d.om_set_depth_func(L.DEPTH_GREATER_EQUAL)
        for v, o, (rt, ds, tri, tag), cb in zip(self.views, self.offsets[:-1], self.targets, self.cbs):
            d.rs_set_viewport(0, 0, v.width, v.height, 0.0, 1.0)
            d.om_set_render_targets(rt, ds)
            d.om_set_visibility_buffer(tri, tag)
            d.om_set_primitive_id_buffer(self.prim.data_ptr() + 4 * int(o))
            d.clear_render_target_view(rt, (0.0, 0.0, 0.0, 0.0))
            d.clear_depth_stencil_view(ds, 0.0)
            d.vs_set_constant_buffer(0, cb)
            for vb, ib, m in zip(self.vbs, self.ibs, self.meshes):
                d.ia_set_vertex_buffer(vb, 32)
                d.ia_set_index_buffer(ib)
                d.draw_indexed(m.indices.numel())
                if per_draw:
                    d.flush()
        if not recording:
            d.flush()
        d.om_set_primitive_id_buffer(None)
        d.om_set_visibility_buffer(None, None)
        self._leave()

    # ---------------------------------------------------------------- tables
    def _mesh_structs(self, ks, grads: list | None, accum: list | None = None):
        arr = (L.DiffMesh * len(ks))()
        for j, k in enumerate(ks):
            m = self.meshes[k]
            g = grads[k] if grads is not None else {}
            ptr = lambda t: t.data_ptr() if t is not None else None
            arr[j] = L.DiffMesh(m.positions.data_ptr(), m.colors.data_ptr(), m.uvs.data_ptr(), m.normals.data_ptr(),
                                m.indices.data_ptr(), m.vertex_count, m.triangle_count, m.chain.data_ptr(), m.tex_w,
                                m.tex_h, m.levels, ptr(g.get("positions")), ptr(g.get("colors")), ptr(g.get("uvs")),
                                ptr(g.get("normals")), ptr(g.get("texture")),
                                accum[k].data_ptr() if accum is not None else None)
        return arr

    def _view_structs(self, vs, rgba: torch.Tensor | None, grad: torch.Tensor | None, draw_maps):
        arr = (L.DiffView * len(vs))()
        nm = len(self.meshes)
        for j, (v, dm) in enumerate(zip(vs, draw_maps)):
            w = self.views[v]
            o = int(self.offsets[v])
            arr[j] = L.DiffView((C.c_float * 16)(*w.clip_from_world.tolist()), w.width, w.height,
                                self.prim.data_ptr() + 4 * o, self.tags.data_ptr() + 4 * o, v * nm, nm, dm,
                                rgba.data_ptr() + 16 * o if rgba is not None else None,
                                grad.data_ptr() + 16 * o if grad is not None else None)
        return arr

    def _single_map(self, k: int):
        m = self._one_map.get(k)
        if m is None:
            m = (C.c_uint32 * len(self.meshes))(*[0 if i == k else NONE for i in range(len(self.meshes))])
            self._one_map[k] = m
        return m

    def new_grads(self, want=GRAD_KEYS) -> list:
        out = []
        for m in self.meshes:
            shapes = {"positions": m.positions.shape, "colors": m.colors.shape, "uvs": m.uvs.shape,
                      "normals": m.normals.shape, "texture": m.chain.shape}
            out.append({k: torch.empty(shapes[k], dtype=torch.float32, device="cuda") for k in want})
        return out

    # ---------------------------------------------------------------- shading
    # Each plan_* method builds the C structures of its launches once and returns a callable
    # that queues them (ordered after torch's current stream and before its later work).

    def _run(self, calls) -> None:
        self._enter()
        for fn, args in calls:
            P._check(self.dev.ctx, fn(self.dev.ctx, *args))
        self._leave()

    def new_image(self) -> torch.Tensor:
        return torch.empty((int(self.offsets[-1]), 4), dtype=torch.float32, device="cuda")

    def plan_forward(self, sh: torch.Tensor, out: torch.Tensor):
        """Every view's image into the flat (total_pixels, 4) `out`, in one launch."""
        nm, nv = len(self.meshes), len(self.views)
        args = (self._mesh_structs(range(nm), None), nm,
                self._view_structs(range(nv), out, None, [self._draw_map[0]] * nv), nv, sh.data_ptr())
        calls = [(self.dev.lib.mlv_diff_batch_forward, args)]
        return lambda: self._run(calls)

    def plan_backward(self, sh: torch.Tensor, grad_out: torch.Tensor, grads: list, grad_sh: torch.Tensor | None):
        """Overwrites grads (one dict per mesh) and grad_sh with the gradients of
        sum(image * grad_out) over all views, in one launch."""
        nm, nv = len(self.meshes), len(self.views)
        args = (self._mesh_structs(range(nm), grads), nm,
                self._view_structs(range(nv), None, grad_out, [self._draw_map[0]] * nv), nv, sh.data_ptr(),
                grad_sh.data_ptr() if grad_sh is not None else None, 0)
        calls = [(self.dev.lib.mlv_diff_batch_backward, args)]
        return lambda: self._run(calls)

    def plan_forward_per_draw(self, sh: torch.Tensor, out: torch.Tensor):
        """The same image, one launch per (view, mesh)."""
        calls = []
        for k in range(len(self.meshes)):
            meshes = self._mesh_structs([k], None)
            dm = self._single_map(k)
            for v in range(len(self.views)):
                calls.append((self.dev.lib.mlv_diff_batch_forward,
                              (meshes, 1, self._view_structs([v], out, None, [dm]), 1, sh.data_ptr())))
        return lambda: self._run(calls)

    def plan_backward_per_draw(self, sh: torch.Tensor, grad_out: torch.Tensor, grads: list,
                               grad_sh: torch.Tensor | None, accum: list):
        """The same gradients, one launch per (view, mesh): the gradients are cleared once, then
        every launch adds its view's pixels of one mesh. accum: per mesh, (V, 3) float64 scratch
        for the position sums."""
        calls = []
        gsh = grad_sh.data_ptr() if grad_sh is not None else None
        for k in range(len(self.meshes)):
            meshes = self._mesh_structs([k], grads, accum)
            dm = self._single_map(k)
            for v in range(len(self.views)):
                calls.append((self.dev.lib.mlv_diff_batch_backward,
                              (meshes, 1, self._view_structs([v], None, grad_out, [dm]), 1, sh.data_ptr(), gsh, 1)))
        clear = [t for g in grads for t in g.values()] + list(accum) + ([grad_sh] if grad_sh is not None else [])

        def run():
            for t in clear:
                t.zero_()
            self._run(calls)
        return run

    def capture_step(self, sh: torch.Tensor, out: torch.Tensor, target: torch.Tensor, scale: float,
                     grad_out: torch.Tensor, grads: list, grad_sh: torch.Tensor | None) -> P.Graph:
        """One HIP graph of a whole batched step: the raster frame, the shading pass into `out`,
        grad_out = (out - target) * scale (the mean-squared-error gradient for scale = 2 / numel)
        and the adjoint into grads / grad_sh. Capturing runs the step once; replay it with
        run_graph."""
        fwd = self.plan_forward(sh, out)
        bwd = self.plan_backward(sh, grad_out, grads, grad_sh)
        d = self.dev
        d.reset_stats()  # draw tags count from 0, as in rasterize(); the graph keeps them
        d.graph_capture_begin()
        try:
            self.rasterize(recording=True)
            fwd()
            self._run([(d.lib.mlv_diff_l2_grad, (out.data_ptr(), target.data_ptr(), out.numel(), float(scale),
                                                 grad_out.data_ptr()))])
            bwd()
        except BaseException:
            try:  # leave the recording state
                d.graph_capture_end().release()
            except Exception:
                pass
            raise
        self._enter()
        g = d.graph_capture_end()
        self._leave()
        return g

    def run_graph(self, g: P.Graph) -> None:
        """Replays a captured step, ordered with the torch stream."""
        self._enter()
        g.launch()
        self._leave()

    def forward(self, sh: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = self.new_image() if out is None else out
        self.plan_forward(sh, out)()
        return out

    def backward(self, sh: torch.Tensor, grad_out: torch.Tensor, grads: list, grad_sh: torch.Tensor | None) -> None:
        self.plan_backward(sh, grad_out, grads, grad_sh)()

    def forward_per_draw(self, sh: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = self.new_image() if out is None else out
        self.plan_forward_per_draw(sh, out)()
        return out

    def backward_per_draw(self, sh: torch.Tensor, grad_out: torch.Tensor, grads: list,
                          grad_sh: torch.Tensor | None, accum: list) -> None:
        self.plan_backward_per_draw(sh, grad_out, grads, grad_sh, accum)()

    def new_accum(self) -> list:
        return [torch.zeros((m.vertex_count, 3), dtype=torch.float64, device="cuda") for m in self.meshes]

    def release(self) -> None:
        for h in self._handles:
            h.release()
        self._handles.clear()
        for t in self.targets:
            for h in t:
                h.release()
        self.targets.clear()
