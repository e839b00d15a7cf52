"""Mip chains of textures -- TEST INFRASTRUCTURE ONLY (the checker, never the product).

numpy restatement of the library's mip generation (mlv_texture2d_generate_mips), an extension:
the reference's Texture2D has one level (source/common_shader_core.h:20-24) and load_texture
ignores its files' mip levels (source/main.c:538-559). Level l + 1 from level l: each texel is
the average of its 2x2 window (2x1 or 1x2 once a side of level l is 1; an odd side drops its last
row or column):
  - RGBA8 (u32, R in the low byte): each byte's window sum in integers, (sum + n/2) // n;
  - RGBA32F / RGBA16F: the window summed in row order in f32 (((a + b) + c) + d), divided by the
    window's area, RGBA16F then rounded to half (nearest-even).
Only tests/ and bench.py's cpu_baseline leg use this module.
"""
from __future__ import annotations

import numpy as np


def level_sizes(width: int, height: int, levels: int | None = None) -> list[tuple[int, int]]:
    """(w, h) of each level: max(w >> l, 1) x max(h >> l, 1), down to 1x1 (or `levels` of them)."""
    out = [(width, height)]
    while (levels is None or len(out) < levels) and out[-1] != (1, 1):
        w, h = out[-1]
        out.append((max(w >> 1, 1), max(h >> 1, 1)))
    if levels is not None and len(out) < levels:
        raise ValueError(f"{levels} levels do not fit a {width}x{height} texture")
    return out


def _windows(src: np.ndarray, w: int, h: int):
    """The window taps of every destination texel, in row order."""
    sh, sw = src.shape[:2]
    kw, kh = (2 if sw > 1 else 1), (2 if sh > 1 else 1)
    taps = []
    for dy in range(kh):
        for dx in range(kw):
            taps.append(src[dy: dy + kh * h: kh, dx: dx + kw * w: kw][:h, :w])
    return taps


def next_level(src: np.ndarray) -> np.ndarray:
    """One level down. src: (h, w) uint32 RGBA8, or (h, w, 4) float32 / float16."""
    sh, sw = src.shape[:2]
    w, h = max(sw >> 1, 1), max(sh >> 1, 1)
    taps = _windows(src, w, h)
    n = len(taps)
    if src.dtype == np.uint32:
        out = np.zeros((h, w), np.uint32)
        for k in range(4):
            acc = sum(((t >> np.uint32(8 * k)) & np.uint32(255)).astype(np.uint32) for t in taps)
            out |= ((acc + np.uint32(n // 2)) // np.uint32(n)) << np.uint32(8 * k)
        return out
    acc = taps[0].astype(np.float32)
    for t in taps[1:]:
        acc = acc + t.astype(np.float32)
    avg = acc / np.float32(n)
    return avg.astype(src.dtype)


def chain(level0: np.ndarray, levels: int | None = None) -> list[np.ndarray]:
    """Every level, level 0 first."""
    sizes = level_sizes(level0.shape[1], level0.shape[0], levels)
    out = [np.ascontiguousarray(level0)]
    for _ in sizes[1:]:
        out.append(next_level(out[-1]))
    return out


def flatten(levels: list[np.ndarray]) -> np.ndarray:
    """The levels concatenated, level 0 first: (N,) uint32 or (N, 4) float."""
    if levels[0].dtype == np.uint32:
        return np.concatenate([l.reshape(-1) for l in levels])
    return np.concatenate([l.reshape(-1, 4) for l in levels])
