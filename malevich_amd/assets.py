"""The reference's asset files (.octrn): meshes and images, and the scenes built from them.

Layout (all little-endian; read from the files themselves, main.c:526-559 loads them):
- bytes 0-7: the magic, "octarine" stored byte-reversed ("eniratco"); bytes 8-11: u32 type
  (1 = mesh, 0 = image); bytes 12-15: padding.
- mesh: at 16, u32 size_of_data, vertex_count, index_count; from 28, vertex_count vertices of
  8 floats (position xyz, normal xyz, uv), then index_count u32 indices. load_mesh
  (main.c:526-536) takes the vertex buffer at the data start and the indices after it.
- image: at 16, u64 size_of_data, u32 format (bits 0-3 type, 4-11 bits per pixel, 12-15 channels,
  16-31 flags: 0x4801 R32G32B32A32_FLOAT, 0x4401 R16G16B16A16_FLOAT, 0x4203 R8G8B8A8_UNORM,
  0x24203 R8G8B8A8_UNORM_SRGB), u16 width, height, depth, array_size, mip_levels, flags; texels
  from 40.

The reference draws its FTM and Toon scenes from these meshes with basic_vs + basic_ps and one
RGBA8 texture per mesh run through load_texture's sRGB pre-pass (mlv_texture2d_srgb_to_linear on
the device). The mesh textures themselves are not in the reference's asset set; load_scene
stands in seeded textures for them.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np

MAGIC = b"eniratco"
TYPE_IMAGE, TYPE_MESH = 0, 1
FORMAT_RGBA32F = 0x4801
FORMAT_RGBA16F = 0x4401
FORMAT_RGBA8 = 0x4203
FORMAT_RGBA8_SRGB = 0x24203

# the reference's scenes (main.c:1319-1371): mesh files in draw order
FTM_MESHES = ("ftm_piedras", "ftm_madera", "ftm_leaves", "ftm_dec", "ftm_roof", "ftm_ground", "ftm_sky")
TOON_MESHES = ("toon_house", "toon_sky")


@dataclass
class OctrnMesh:
    vertices: np.ndarray  # (V, 8) float32: position, normal, uv (basic_vs's input layout)
    indices: np.ndarray   # (I,) uint32


@dataclass
class OctrnImage:
    format: int
    width: int
    height: int
    mip_levels: int
    texels: np.ndarray    # (h, w, 4) float32 / float16 for the float formats, (h, w) uint32 for RGBA8


def read_octrn(path) -> OctrnMesh | OctrnImage:
    data = Path(path).read_bytes()
    if len(data) < 16 or data[:8] != MAGIC:
        raise ValueError(f"{path}: not an .octrn file")
    kind = struct.unpack_from("<I", data, 8)[0]
    if kind == TYPE_MESH:
        size, nv, ni = struct.unpack_from("<III", data, 16)
        if size != nv * 32 + ni * 4 or len(data) < 28 + size:
            raise ValueError(f"{path}: mesh sizes do not match the file")
        v = np.frombuffer(data, np.float32, nv * 8, 28).reshape(nv, 8).copy()
        i = np.frombuffer(data, np.uint32, ni, 28 + nv * 32).copy()
        return OctrnMesh(v, i)
    if kind == TYPE_IMAGE:
        size, fmt, w, h, depth, arr, mips, flags = struct.unpack_from("<QIHHHHHH", data, 16)
        if len(data) < 40 + size:
            raise ValueError(f"{path}: image data past the end of the file")
        if fmt == FORMAT_RGBA32F:
            t = np.frombuffer(data, np.float32, w * h * 4, 40).reshape(h, w, 4).copy()
        elif fmt == FORMAT_RGBA16F:
            t = np.frombuffer(data, np.float16, w * h * 4, 40).reshape(h, w, 4).copy()
        elif fmt in (FORMAT_RGBA8, FORMAT_RGBA8_SRGB):
            t = np.frombuffer(data, np.uint32, w * h, 40).reshape(h, w).copy()
        else:
            raise ValueError(f"{path}: image format {fmt:#x} not supported")
        return OctrnImage(fmt, w, h, mips, t)
    raise ValueError(f"{path}: unknown .octrn type {kind}")


def write_octrn_mesh(path, vertices: np.ndarray, indices: np.ndarray) -> None:
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 8)
    i = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<II", TYPE_MESH, 0))
        f.write(struct.pack("<III", v.nbytes + i.nbytes, len(v), len(i)))
        f.write(v.tobytes())
        f.write(i.tobytes())


def write_octrn_image(path, texels: np.ndarray, fmt: int) -> None:
    dt = {FORMAT_RGBA32F: np.float32, FORMAT_RGBA16F: np.float16}.get(fmt, np.uint32)
    t = np.ascontiguousarray(texels, dt)
    h, w = t.shape[:2]
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<II", TYPE_IMAGE, 0))
        f.write(struct.pack("<QIHHHHHH", t.nbytes, fmt, w, h, 1, 1, 1, 0))
        f.write(t.tobytes())


def load_scene(asset_dir, names=FTM_MESHES, tex_size: int = 512, seed: int = 0) -> list:
    """[(mesh, stand-in sRGB RGBA8 texture)] for the named meshes of an asset folder. The
    textures are seeded smooth noise (the reference's own mesh textures are not shipped)."""
    from . import scenes as S
    out = []
    for k, name in enumerate(names):
        m = read_octrn(Path(asset_dir) / f"{name}_mesh.octrn")
        if not isinstance(m, OctrnMesh):
            raise ValueError(f"{name}: not a mesh")
        out.append((m, S.random_rgba8(tex_size, seed + k)))
    return out
