"""Seeded synthetic scenes for the benchmark configurations and parity tests.

Camera conventions follow the reference (source/main.c:1422-1477): view space
is left-handed with +y up and +z forward, world space is right-handed with
+z up, and the projection is reversed-Z and infinite (main.c:1431-1441), so
z_ndc = near / z_view. The constant buffer is PerFrameCB (main.c:169-173):
clip_from_world, view_from_clip, world_from_view, row-major float32.

Vertex layouts are the reference's input layouts:
  basic_vs / vertex_lighting_vs: pos[3], normal[3], uv[2]  (basic_vs.c:3-7), 32 B
  passthrough_vs / fullscreen_vs: pos[4], color[3], pad    (main.c:208-212), 32 B
Triangles are wound so that the reference's back-face rule (cull when the
screen-space signed area is > 0, main.c:856) keeps the faces seen from outside.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ----------------------------------------------------------------- camera


def perspective_reversed_z(fov_y_deg: float, aspect: float, near: float) -> np.ndarray:
    """clip_from_view of main.c:1431-1441 (left-handed, reversed-Z, infinite)."""
    sy = 1.0 / math.tan(math.radians(fov_y_deg) / 2.0)
    sx = sy / aspect
    return np.array([[sx, 0, 0, 0], [0, sy, 0, 0], [0, 0, 0, near], [0, 0, 1, 0]], np.float64)


def look_at_world_from_view(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """world_from_view whose columns are (right, up, forward, eye) in world space.

    With eye looking down -x and z up this equals the reference's
    change_of_basis (main.c:1459-1464): view +z -> world -x, view +x -> world +y,
    view +y -> world +z.
    """
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    if np.linalg.norm(r) < 1e-9:
        r = np.cross(f, np.array([0.0, 1.0, 0.0]))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, f, eye
    return m


def per_frame_cb(world_from_view: np.ndarray, clip_from_view: np.ndarray) -> np.ndarray:
    """PerFrameCB (main.c:169-173) as 48 float32: clip_from_world, view_from_clip, world_from_view."""
    view_from_world = np.linalg.inv(world_from_view)
    clip_from_world = clip_from_view @ view_from_world
    view_from_clip = np.linalg.inv(clip_from_view)
    return np.concatenate([clip_from_world.ravel(), view_from_clip.ravel(), world_from_view.ravel()]).astype(
        np.float32)


def camera_cb(eye, target, fov_y_deg: float, width: int, height: int, near: float = 0.01) -> np.ndarray:
    return per_frame_cb(look_at_world_from_view(eye, target), perspective_reversed_z(fov_y_deg, width / height, near))


# ------------------------------------------------------------------ meshes


_ICO_FACES = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                       (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                       (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
_GEODESIC_CACHE: dict = {}


def geodesic_sphere(freq: int) -> tuple[np.ndarray, np.ndarray]:
    """Class-I geodesic sphere: each icosahedron face split into freq^2 triangles.

    Returns unit positions (V,3) float64 and faces (F,3) int64 wound
    counter-clockwise about the outward normal. V = 10*freq^2 + 2, F = 20*freq^2
    (freq = 32 gives 10,242 vertices and 20,480 triangles).
    """
    n = int(freq)
    if n in _GEODESIC_CACHE:
        p, f = _GEODESIC_CACHE[n]
        return p.copy(), f.copy()
    t = (1.0 + 5 ** 0.5) / 2.0
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                     [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    m = (ii + jj) <= n
    li, lj = ii[m], jj[m]                      # lattice points of one face
    lut = -np.ones((n + 1, n + 1), np.int64)
    lut[li, lj] = np.arange(li.size)
    nf = len(_ICO_FACES)
    # weights on the face corners (a, b, c) = (n-i-j, i, j); a canonical key per point:
    # (corner id, weight) pairs sorted by corner id, so shared edges/corners coincide
    corners = np.repeat(_ICO_FACES[:, None, :], li.size, 1)                  # (F, L, 3)
    w = np.stack([n - li - lj, li, lj], 1)[None].repeat(nf, 0)               # (F, L, 3)
    ckey = np.where(w > 0, corners, 12)           # zero-weight corners sort last
    order = np.argsort(ckey, axis=2)
    cs = np.take_along_axis(ckey, order, 2)
    ws = np.take_along_axis(w, order, 2)
    key = ((cs[..., 0] * (n + 1) + ws[..., 0]) * 13 * (n + 1) + cs[..., 1] * (n + 1) + ws[..., 1]) \
        * 13 * (n + 1) + cs[..., 2] * (n + 1) + ws[..., 2]
    uk, first, inv = np.unique(key.ravel(), return_index=True, return_inverse=True)
    csf, wsf = np.minimum(cs.reshape(-1, 3)[first], 11), ws.reshape(-1, 3)[first]
    pos = (base[csf[:, 0]] * wsf[:, 0:1] + base[csf[:, 1]] * wsf[:, 1:2] + base[csf[:, 2]] * wsf[:, 2:3]) / n
    pos /= np.linalg.norm(pos, axis=1, keepdims=True)
    gid = inv.reshape(nf, li.size)
    # triangles of one face in lattice coordinates
    tri = []
    for i in range(n):
        for j in range(n - i):
            tri.append((lut[i, j], lut[i + 1, j], lut[i, j + 1]))
            if i + j < n - 1:
                tri.append((lut[i + 1, j], lut[i + 1, j + 1], lut[i, j + 1]))
    tri = np.array(tri, np.int64)
    f = gid[:, tri].reshape(-1, 3)
    nrm = np.cross(pos[f[:, 1]] - pos[f[:, 0]], pos[f[:, 2]] - pos[f[:, 0]])
    flip = (nrm * pos[f].mean(axis=1)).sum(1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    _GEODESIC_CACHE[n] = (pos, f)
    return pos.copy(), f.copy()


def value_noise3(p: np.ndarray, seed: int, octaves: int = 3, base_freq: float = 2.0) -> np.ndarray:
    """Smooth seeded value noise in [-1, 1] at points p (N,3)."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)
    total = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        f = f * f * (3 - 2 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = perm[(i0[:, 0] + dx) % 4096]
                    h = perm[(h + i0[:, 1] + dy) % 4096]
                    h = perm[(h + i0[:, 2] + dz) % 4096]
                    wgt = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (
                        f[:, 2] if dz else 1 - f[:, 2])
                    acc += wgt * table[h]
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def sphere_uv(p: np.ndarray) -> np.ndarray:
    u = np.arctan2(p[:, 1], p[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(p[:, 2], -1, 1)) / np.pi
    return np.stack([u, 1.0 - v], 1)


def vertex_normals(pos: np.ndarray, faces: np.ndarray) -> np.ndarray:
    fn = np.cross(pos[faces[:, 1]] - pos[faces[:, 0]], pos[faces[:, 2]] - pos[faces[:, 0]])
    n = np.zeros_like(pos)
    for k in range(3):
        for c in range(3):
            n[:, c] += np.bincount(faces[:, k], weights=fn[:, c], minlength=len(pos))
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)


@dataclass
class Mesh:
    vertices: np.ndarray   # (V, 8) float32: pos3, normal3, uv2
    indices: np.ndarray    # (F*3,) uint32

    @property
    def triangle_count(self) -> int:
        return self.indices.size // 3


def sphere_mesh(freq: int, displace: float = 0.0, seed: int = 0, transform: np.ndarray | None = None) -> Mesh:
    pos, faces = geodesic_sphere(freq)
    uv = sphere_uv(pos)
    if displace:
        pos = pos * (1.0 + displace * value_noise3(pos, seed))[:, None]
    nrm = vertex_normals(pos, faces) if displace else pos.copy()
    if transform is not None:
        pos = pos @ transform[:3, :3].T + transform[:3, 3]
        nrm = nrm @ np.linalg.inv(transform[:3, :3])
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    v = np.concatenate([pos, nrm, uv], 1).astype(np.float32)
    return Mesh(v, faces.astype(np.uint32).ravel())


def reverse_winding(indices: np.ndarray) -> np.ndarray:
    return indices.reshape(-1, 3)[:, [0, 2, 1]].ravel().copy()


# ----------------------------------------------------------------- textures


def random_rgba8(size: int, seed: int) -> np.ndarray:
    """(size, size) uint32 RGBA8 with uniform random bytes (R in the low byte)."""
    return np.random.default_rng(seed).integers(0, 2 ** 32, (size, size), dtype=np.uint64).astype(np.uint32)


def smooth_env_map(width: int, height: int, seed: int, lo: float = 0.05, hi: float = 2.0) -> np.ndarray:
    """(h, w, 4) float32 equirect map of seeded smooth noise, alpha 1, values in [lo, hi]
    (the reference's irradiance asset spans about [0.085, 2.11]). Maps wider than 256
    texels are the bilinear upsampling of a 256-wide noise map (the noise is smooth at
    that scale), which keeps generation fast."""
    if width > 256:
        small = smooth_env_map(256, max(1, 256 * height // width), seed, lo, hi)
        return _resize_bilinear(small, height, width)
    yy, xx = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    theta = (yy + 0.5) / height * np.pi
    phi = (xx + 0.5) / width * 2 * np.pi
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1).reshape(-1, 3)
    out = np.empty((height * width, 4), np.float64)
    for c in range(3):
        out[:, c] = lo + (hi - lo) * 0.5 * (1.0 + value_noise3(d, seed * 7 + c, octaves=4, base_freq=1.5))
    out[:, 3] = 1.0
    return out.reshape(height, width, 4).astype(np.float32)


def _resize_bilinear(img: np.ndarray, h: int, w: int) -> np.ndarray:
    sh, sw = img.shape[:2]
    ys = np.clip((np.arange(h) + 0.5) * sh / h - 0.5, 0, sh - 1)
    xs = np.clip((np.arange(w) + 0.5) * sw / w - 0.5, 0, sw - 1)
    y0, x0 = np.floor(ys).astype(int), np.floor(xs).astype(int)
    y1, x1 = np.minimum(y0 + 1, sh - 1), np.minimum(x0 + 1, sw - 1)
    fy, fx = (ys - y0)[:, None, None], (xs - x0)[None, :, None]
    top = img[y0][:, x0] * (1 - fx) + img[y0][:, x1] * fx
    bot = img[y1][:, x0] * (1 - fx) + img[y1][:, x1] * fx
    return (top * (1 - fy) + bot * fy).astype(np.float32)


# ------------------------------------------------------------------ scenes


# Suprematism scene, the reference's built-in asset-free draw (main.c:232-254):
# pos4 + colour3 + pad per vertex, passthrough_vs / passthrough_ps, 24 indices.
SUPREMATISM_VERTICES = np.array([
    [0.34107, 0.12215, 0.5, 1.0, 0.07500, 0.08200, 0.06300, 0.0],
    [0.95357, 0.12500, 0.5, 1.0, 0.07500, 0.08200, 0.06300, 0.0],
    [0.96250, 0.86931, 0.5, 1.0, 0.07500, 0.08200, 0.06300, 0.0],
    [0.33928, 0.86505, 0.5, 1.0, 0.07500, 0.08200, 0.06300, 0.0],
    [0.09464, 0.12500, 0.75, 1.0, 0.14100, 0.29000, 0.60800, 0.0],
    [0.69285, 0.39772, 0.75, 1.0, 0.14100, 0.29000, 0.60800, 0.0],
    [0.09107, 0.60937, 0.75, 1.0, 0.14100, 0.29000, 0.60800, 0.0],
    [0.00000, 0.00000, 0.25, 1.0, 0.96100, 0.96100, 0.92900, 0.0],
    [1.00000, 0.00000, 0.25, 1.0, 0.96100, 0.96100, 0.92900, 0.0],
    [1.00000, 1.00000, 0.25, 1.0, 0.96100, 0.96100, 0.92900, 0.0],
    [0.00000, 1.00000, 0.25, 1.0, 0.96100, 0.96100, 0.92900, 0.0],
], np.float32)
SUPREMATISM_INDICES = np.array([0, 1, 2, 2, 3, 0, 4, 5, 6, 7, 8, 9, 9, 10, 7, 0, 0, 0, 0, 0, 0, 0, 0, 0], np.uint32)

# Fullscreen quad (main.c:255-270): drawn with fullscreen_vs / env_lighting_ps.
FULLSCREEN_VERTICES = np.array([
    [0.0, 0.0, 0.0, 1.0, 0, 0, 0, 0], [1.0, 0.0, 0.0, 1.0, 0, 0, 0, 0],
    [1.0, 1.0, 0.0, 1.0, 0, 0, 0, 0], [0.0, 1.0, 0.0, 1.0, 0, 0, 0, 0]], np.float32)
FULLSCREEN_INDICES = np.array([0, 1, 2, 2, 3, 0] + [0] * 18, np.uint32)

# render()'s clear colour (main.c:1271)
CLEAR_COLOR = (227 / 255, 223 / 255, 216 / 255, 0.0)


@dataclass
class ScreenTriangles:
    """Passthrough-layout triangle soup (de-indexed): vertices (3N, 8), indices arange."""
    vertices: np.ndarray
    indices: np.ndarray
    big_mask: np.ndarray = field(default_factory=lambda: np.zeros(0, bool))


def _ndc_from_pixels(px: np.ndarray, py: np.ndarray, width: int, height: int) -> tuple[np.ndarray, np.ndarray]:
    # passthrough_vs maps input x to 2x-1; the viewport maps NDC to pixels with y flipped (main.c:826-831)
    return px / width, 1.0 - py / height


def binning_stress(n_tiny: int, n_big: int, width: int, height: int, seed: int,
                   big_every: int = 10000) -> ScreenTriangles:
    """Tiny random triangles (area U[1,4] px^2, uniform centroids and depths) with
    screen-covering triangles interleaved every `big_every` in submission order."""
    rng = np.random.default_rng(seed)
    total = n_tiny + n_big
    big = np.zeros(total, bool)
    if n_big:
        pos_big = np.linspace(big_every // 2, total - 1, n_big).astype(np.int64) if n_tiny else np.arange(n_big)
        pos_big = np.minimum(pos_big, total - 1)
        big[np.unique(pos_big)] = True
        while big.sum() < n_big:  # unique() may merge positions in tiny cases
            big[rng.integers(0, total)] = True
    nt = int((~big).sum())
    # tiny: centroid + random orientation; equilateral-ish with area A
    cx = rng.uniform(0, width, nt)
    cy = rng.uniform(0, height, nt)
    area = rng.uniform(1.0, 4.0, nt)
    r = np.sqrt(4 * area / (3 * np.sqrt(3)))  # circumradius of an equilateral triangle of that area
    ang = rng.uniform(0, 2 * np.pi, nt)
    k = np.arange(3)[None, :] * (2 * np.pi / 3)
    tx = cx[:, None] + r[:, None] * np.cos(ang[:, None] + k)
    ty = cy[:, None] + r[:, None] * np.sin(ang[:, None] + k)
    # screen y is down: reference keeps signed_area <= 0 (main.c:851-856)
    sa = (tx[:, 1] - tx[:, 0]) * (ty[:, 2] - ty[:, 0]) - (tx[:, 2] - tx[:, 0]) * (ty[:, 1] - ty[:, 0])
    sw = sa > 0
    tx[sw] = tx[sw][:, [0, 2, 1]]
    ty[sw] = ty[sw][:, [0, 2, 1]]
    zt = 0.01 / rng.uniform(0.1, 100.0, nt)
    # big: a triangle covering the whole screen with a margin
    nb = int(big.sum())
    bx = np.array([-0.5 * width, 2.5 * width, -0.5 * width])[None].repeat(nb, 0)
    by = np.array([-0.5 * height, -0.5 * height, 2.5 * height])[None].repeat(nb, 0)
    sa = (bx[:, 1] - bx[:, 0]) * (by[:, 2] - by[:, 0]) - (bx[:, 2] - bx[:, 0]) * (by[:, 1] - by[:, 0])
    if (sa > 0).any():
        bx, by = bx[:, [0, 2, 1]], by[:, [0, 2, 1]]
    zb = 0.01 / rng.uniform(0.1, 100.0, nb)
    X = np.empty((total, 3))
    Y = np.empty((total, 3))
    Z = np.empty(total)
    X[~big], Y[~big], Z[~big] = tx, ty, zt
    X[big], Y[big], Z[big] = bx, by, zb
    ix, iy = _ndc_from_pixels(X, Y, width, height)
    col = rng.uniform(0.0, 1.0, (total, 3))
    v = np.zeros((total, 3, 8), np.float32)
    v[:, :, 0] = ix
    v[:, :, 1] = iy
    v[:, :, 2] = Z[:, None]
    v[:, :, 3] = 1.0
    v[:, :, 4:7] = col[:, None, :]
    return ScreenTriangles(v.reshape(-1, 8), np.arange(total * 3, dtype=np.uint32), big)


def c1_scene(width: int = 640, height: int = 360, seed: int = 1):
    """The 640x360 sphere workload: 10,242 vertices / 20,480 triangles, camera at distance 3,
    60 deg vertical FOV, 512x512 RGBA8 random texture, basic_vs + basic_ps."""
    mesh = sphere_mesh(32)
    cb = camera_cb(eye=(3.0, 0.0, 0.0), target=(0.0, 0.0, 0.0), fov_y_deg=60.0, width=width, height=height)
    tex = random_rgba8(512, seed)
    return mesh, cb, tex


def c1_light() -> np.ndarray:
    """The 640x360 workload's directional light for the Lambert pixel shader (12 floats:
    direction towards the light, colour, ambient, each padded to four): white light from the
    camera's side and above, a dim ambient term."""
    d = np.array([1.0, 0.5, 0.8])
    d /= np.linalg.norm(d)
    return np.array([d[0], d[1], d[2], 0.0, 1.0, 1.0, 1.0, 0.0, 0.15, 0.15, 0.15, 0.0], np.float32)


def c2_textures(n_instances: int = 64, seed: int = 2, albedo_size: int = 1024):
    """The 1080p workload's maps: a 256x128 RGBA16F irradiance map (vertex_lighting_vs, one level),
    a 2048x1024 RGBA16F radiance map and one albedo_size^2 RGBA8 texture of uniform random bytes per
    instance (level 0 of each; their mip chains are built by the device)."""
    irr = smooth_env_map(256, 128, seed=seed + 16, lo=0.05, hi=2.0).astype(np.float16)
    rad = smooth_env_map(2048, 1024, seed=seed + 17, lo=0.05, hi=2.0).astype(np.float16)
    rng = np.random.default_rng(seed + 18)
    albedo = [rng.integers(0, 2 ** 32, (albedo_size, albedo_size), dtype=np.uint32) for _ in range(n_instances)]
    return irr, rad, albedo


def random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@dataclass
class C2Scene:
    instances: list            # Mesh per instance, transform baked into the vertices
    textures: list             # (1024, 1024) uint32 RGBA8 per instance
    cb: np.ndarray
    width: int
    height: int
    depth_complexity: float

    @property
    def triangle_count(self) -> int:
        return sum(m.triangle_count for m in self.instances)


def c2_scene(width: int = 1920, height: int = 1080, n_instances: int = 64, seed: int = 2,
             texture_size: int = 1024, target_complexity: float = 4.0) -> C2Scene:
    """The 1080p forward workload: seeded displaced spheres of 16k-64k triangles (geodesic
    frequency 29..56, about 2M triangles for 64 instances) at random positions in a
    20-unit cube, random rotations, scale 0.5-2 times a common factor chosen so the
    summed projected sphere area is about 4 screens (mean depth complexity 4)."""
    rng = np.random.default_rng(seed)
    freqs = rng.integers(29, 57, n_instances)
    centers = rng.uniform(-10.0, 10.0, (n_instances, 3))
    scales = rng.uniform(0.5, 2.0, n_instances)
    eye = np.array([30.0, 0.0, 0.0])
    fov = 60.0
    f_px = (height / 2.0) / math.tan(math.radians(fov) / 2.0)
    z = np.abs(eye[0] - centers[:, 0])
    proj = np.pi * (f_px * scales / z) ** 2
    k = math.sqrt(target_complexity * width * height / proj.sum())
    cb = camera_cb(eye, (0.0, 0.0, 0.0), fov, width, height)
    inst, texs = [], []
    for i in range(n_instances):
        T = np.eye(4)
        T[:3, :3] = random_rotation(rng) * (scales[i] * k)
        T[:3, 3] = centers[i]
        inst.append(sphere_mesh(int(freqs[i]), displace=0.15, seed=seed * 1000 + i, transform=T))
        texs.append(random_rgba8(texture_size, seed * 1000 + i) if texture_size else None)
    return C2Scene(inst, texs, cb, width, height, float((proj * k * k).sum() / (width * height)))


# ------------------------------------------------------------------ differentiable render scene

@dataclass
class DiffMesh:
    positions: np.ndarray  # (V, 3) float32
    colors: np.ndarray     # (V, 3) float32
    uvs: np.ndarray        # (V, 2) float32
    normals: np.ndarray    # (V, 3) float32
    indices: np.ndarray    # (T, 3) uint32

    @property
    def triangle_count(self) -> int:
        return len(self.indices)

    def basic_vertices(self) -> np.ndarray:
        """(V, 8) float32 in basic_vs's input layout (pos3, normal3, uv2; basic_vs.c:3-8)."""
        return np.concatenate([self.positions, self.normals, self.uvs], 1).astype(np.float32)


def uv_sphere(rings: int, segments: int, displace: float = 0.0, seed: int = 0) -> DiffMesh:
    """Latitude/longitude sphere with 2*rings*segments triangles (the pole rows are degenerate
    slivers), wound counter-clockwise about the outward normal, seeded low-frequency radial
    displacement, smooth per-vertex colours in [0.4, 1], uv = (longitude, latitude)."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2 * np.pi, segments + 1)
    T, PH = np.meshgrid(th, ph, indexing="ij")
    unit = np.stack([np.sin(T) * np.cos(PH), np.sin(T) * np.sin(PH), np.cos(T)], -1).reshape(-1, 3)
    uv = np.stack([PH / (2 * np.pi), 1.0 - T / np.pi], -1).reshape(-1, 2)
    r = np.arange(rings)[:, None]
    s = np.arange(segments)[None, :]
    a = r * (segments + 1) + s
    b = (r + 1) * (segments + 1) + s
    c = a + 1
    d = b + 1
    faces = np.stack([np.stack([a, b, c], -1), np.stack([c, b, d], -1)], 2).reshape(-1, 3)
    pos = unit * (1.0 + displace * value_noise3(unit, seed))[:, None] if displace else unit.copy()
    nrm = vertex_normals(pos, faces)
    bad = ~np.isfinite(nrm).all(1) | (np.linalg.norm(nrm, axis=1) < 0.5)
    nrm[bad] = unit[bad]
    col = 0.7 + 0.3 * np.stack([value_noise3(unit, seed + 11 + k, octaves=2) for k in range(3)], 1)
    return DiffMesh(pos.astype(np.float32), col.astype(np.float32), uv.astype(np.float32), nrm.astype(np.float32),
                    faces.astype(np.uint32))


def smooth_texture(size: int, seed: int, lo: float = 0.25, hi: float = 1.0) -> np.ndarray:
    """(size, size, 4) float32 RGBA of seeded smooth noise in [lo, hi], alpha 1."""
    t = smooth_env_map(size, size, seed, lo, hi)
    return np.ascontiguousarray(t, np.float32)


def mip_chain(tex: np.ndarray) -> tuple[np.ndarray, int]:
    """Full mip chain of an (h, w, 4) float texture by 2x2 box filtering (odd edges clamp);
    returns the levels concatenated as (N, 4), level 0 first, and the level count."""
    levels = [np.asarray(tex, np.float64)]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        t = levels[-1]
        h, w = t.shape[:2]
        h2, w2 = max(h // 2, 1), max(w // 2, 1)
        t = t[: h2 * 2 if h > 1 else 1, : w2 * 2 if w > 1 else 1]
        if h > 1:
            t = 0.5 * (t[0::2] + t[1::2])
        if w > 1:
            t = 0.5 * (t[:, 0::2] + t[:, 1::2])
        levels.append(t)
    chain = np.concatenate([l.reshape(-1, 4) for l in levels]).astype(np.float32)
    return chain, len(levels)


def random_sh(seed: int) -> np.ndarray:
    """(9, 3) float32 RGB SH light: a strong ambient band-0 term and small higher bands, so the
    irradiance stays in about [0.5, 2] over the sphere."""
    rng = np.random.default_rng(seed)
    sh = rng.uniform(-0.25, 0.25, (9, 3))
    sh[0] = rng.uniform(3.2, 3.8, 3)
    return sh.astype(np.float32)


@dataclass
class DiffScene:
    mesh: DiffMesh
    texture: np.ndarray      # (h, w, 4) float32, level 0
    chain: np.ndarray        # (N, 4) float32 mip chain
    levels: int
    sh: np.ndarray           # (9, 3) float32
    cb: np.ndarray           # (48,) float32 PerFrameCB
    width: int
    height: int

    @property
    def clip_from_world(self) -> np.ndarray:
        return self.cb[:16].reshape(4, 4)


def c4_scene(width: int = 512, height: int = 512, rings: int = 250, segments: int = 400, tex_size: int = 1024,
             seed: int = 4) -> DiffScene:
    """Differentiable-render workload: a displaced sphere of 2*rings*segments triangles (200k at
    the defaults) with per-vertex colours, UVs and normals, a tex_size^2 RGBA32F texture with its
    full mip chain and a 9-coefficient RGB SH light, seen from distance 2.8 with a 50 deg FOV."""
    mesh = uv_sphere(rings, segments, displace=0.08, seed=seed)
    tex = smooth_texture(tex_size, seed + 1)
    chain, levels = mip_chain(tex)
    cb = camera_cb(eye=(2.2, -1.4, 0.9), target=(0.0, 0.0, 0.0), fov_y_deg=50.0, width=width, height=height)
    return DiffScene(mesh, tex, chain, levels, random_sh(seed + 2), cb, width, height)


# ------------------------------------------------------------------ batched views

@dataclass
class ViewSpec:
    width: int
    height: int
    cb: np.ndarray   # (48,) float32 PerFrameCB of this view's camera


@dataclass
class BatchedScene:
    meshes: list          # Mesh (basic_vs layout, transform baked in)
    textures: list        # (s, s, 4) float32 per mesh
    views: list           # ViewSpec
    colors: list          # (V, 3) float32 per-vertex colours per mesh (the differentiable shading's)
    sh: np.ndarray        # (9, 3) float32 SH light shared by all views

    @property
    def triangles_per_view(self) -> int:
        return sum(m.triangle_count for m in self.meshes)

    def diff_mesh(self, k: int) -> DiffMesh:
        """Mesh k as the differentiable shading takes it (positions, colours, uvs, normals, (T, 3) indices)."""
        v = self.meshes[k].vertices
        return DiffMesh(np.ascontiguousarray(v[:, 0:3]), self.colors[k], np.ascontiguousarray(v[:, 6:8]),
                        np.ascontiguousarray(v[:, 3:6]), self.meshes[k].indices.reshape(-1, 3))


def sphere_of_about(triangles: int, displace: float, seed: int, transform: np.ndarray) -> tuple[Mesh, np.ndarray]:
    """A displaced lat/long sphere with about `triangles` triangles in basic_vs's layout, and its
    per-vertex colours."""
    rings = max(3, int(round(np.sqrt(triangles / 2.0 / 1.6))))
    segments = max(4, int(round(triangles / (2.0 * rings))))
    dm = uv_sphere(rings, segments, displace=displace, seed=seed)
    pos = np.c_[dm.positions.astype(np.float64), np.ones(len(dm.positions))] @ transform.T
    nrm = dm.normals.astype(np.float64) @ np.linalg.inv(transform[:3, :3]).T
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    v = np.concatenate([pos[:, :3], nrm, dm.uvs], 1).astype(np.float32)
    return Mesh(v, dm.indices.reshape(-1).astype(np.uint32)), dm.colors


def batched_scene(n_views: int = 256, n_meshes: int = 64, min_tris: int = 1000, max_tris: int = 500_000,
                  resolutions=(256, 512, 1024), tex_sizes=(256, 512, 1024, 2048), seed: int = 5) -> BatchedScene:
    """Many views of one set of meshes: triangle counts log-uniform in [min_tris, max_tris],
    seeded random rotations/scales/positions in a 20-unit cube, an RGBA32F texture per mesh of a
    size drawn from tex_sizes, and n_views cameras at random directions and distances 24-34 looking
    at the cube's centre, each with a square resolution drawn from `resolutions`."""
    rng = np.random.default_rng(seed)
    meshes, textures, colors = [], [], []
    for k in range(n_meshes):
        t = int(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        T = np.eye(4)
        T[:3, :3] = random_rotation(rng) * rng.uniform(0.8, 2.5)
        T[:3, 3] = rng.uniform(-10.0, 10.0, 3)
        mesh, col = sphere_of_about(t, 0.1, seed * 1000 + k, T)
        meshes.append(mesh)
        colors.append(col)
        s = int(rng.choice(tex_sizes))
        textures.append(smooth_texture(s, seed * 1000 + 500 + k))
    views = []
    for _ in range(n_views):
        r = int(rng.choice(resolutions))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = d * rng.uniform(24.0, 34.0)
        views.append(ViewSpec(r, r, camera_cb(eye=tuple(eye), target=(0.0, 0.0, 0.0), fov_y_deg=55.0, width=r,
                                              height=r)))
    return BatchedScene(meshes, textures, views, colors, random_sh(seed * 1000 + 999))
