"""Binary PPM (P6) files for presented frames: the headless stand-in for the reference's window
blit (main.c:286-311). RGB8, top row first, as mlv_present returns it."""
from __future__ import annotations

from pathlib import Path

import numpy as np


def write_ppm(path, rgb: np.ndarray) -> None:
    rgb = np.ascontiguousarray(rgb, np.uint8)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError("write_ppm wants an (H, W, 3) uint8 image")
    h, w = rgb.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode("ascii"))
        f.write(rgb.tobytes())


def read_ppm(path) -> np.ndarray:
    data = Path(path).read_bytes()
    fields, pos = [], 0
    while len(fields) < 4:  # magic, width, height, maxval; '#' comments skipped
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            pos = data.index(b"\n", pos) + 1
            continue
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        fields.append(data[pos:end])
        pos = end
    if fields[0] != b"P6" or int(fields[3]) != 255:
        raise ValueError("not an 8-bit binary PPM")
    w, h = int(fields[1]), int(fields[2])
    pos += 1  # the single whitespace byte before the raster
    return np.frombuffer(data, np.uint8, count=w * h * 3, offset=pos).reshape(h, w, 3).copy()
