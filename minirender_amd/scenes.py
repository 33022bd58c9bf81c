"""Scene description and seeded synthetic workloads.

Data model (host side, mirrors include/minirender/Scene.h of the reference):
  Mesh      TriMesh (Scene.h:73-87): separate position / normal / uv arrays and
            index arrays; uvs are used only when uv_indices is given, as in
            Renderer.cpp:371 (texcoordsI must be non-empty).
  Material  Scene.h:38-45, defaults of Scene.cpp:65-71.
  Draw      a Renderable (Scene.h:47-53): mesh + world transform + material.
  FrameSpec one Renderer::render() call: size, projection, view, light,
            background, flags and the draws in Scene::collectShapes order.

Generators: createCube / createCylinder / createSphere with the reference's
vertex and index order (src/primitives.cpp:7-230), an fBm height-field, random
triangles, a checker + value-noise texture, and the workloads of the
benchmark configurations.  All randomness is numpy PCG64 with explicit seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

F32 = np.float32
PIF = F32(3.14159265358979323846)


# ---------------------------------------------------------------------------
# data model
# ---------------------------------------------------------------------------
@dataclass(eq=False)
class Texture:
    texels: np.ndarray  # (rows, cols, 3) uint8 (RGB8) or float32

    @property
    def rows(self) -> int:
        return int(self.texels.shape[0])

    @property
    def cols(self) -> int:
        return int(self.texels.shape[1])


@dataclass(eq=False)
class Material:
    diffuse: Sequence[float] = (0.7, 0.7, 0.9)
    specular: Sequence[float] = (0.8, 0.8, 0.8)
    emissive: Sequence[float] = (0.0, 0.0, 0.0)
    shininess: float = 12.0
    texture: Optional[Texture] = None


@dataclass(eq=False)
class Mesh:
    positions: np.ndarray  # (V, 3) float32
    normals: np.ndarray    # (N, 3) float32
    indices: np.ndarray    # (T, 3) int32
    normal_indices: Optional[np.ndarray] = None  # None: same as indices
    uvs: Optional[np.ndarray] = None             # (U, 2) float32
    uv_indices: Optional[np.ndarray] = None      # None: uvs unused

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=F32).reshape(-1, 3)
        self.normals = np.ascontiguousarray(self.normals, dtype=F32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.int32).reshape(-1, 3)
        if self.normal_indices is not None:
            self.normal_indices = np.ascontiguousarray(self.normal_indices, dtype=np.int32).reshape(-1, 3)
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, dtype=F32).reshape(-1, 2)
        if self.uv_indices is not None:
            self.uv_indices = np.ascontiguousarray(self.uv_indices, dtype=np.int32).reshape(-1, 3)

    @property
    def ntri(self) -> int:
        return int(self.indices.shape[0])

    @property
    def nidx(self) -> np.ndarray:
        return self.indices if self.normal_indices is None else self.normal_indices

    @property
    def has_uv(self) -> bool:
        return self.uvs is not None and self.uv_indices is not None and len(self.uvs) > 0

    def nbytes(self) -> int:
        n = self.positions.nbytes + self.normals.nbytes + self.indices.nbytes
        if self.normal_indices is not None:
            n += self.normal_indices.nbytes
        if self.has_uv:
            n += self.uvs.nbytes + self.uv_indices.nbytes
        return n


@dataclass(eq=False)
class Draw:
    mesh: Mesh
    material: Material = field(default_factory=Material)
    transform: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=F32))
    # no near-plane test or clip, as Renderer::paintTriangle(a, b, c, false)
    # draws (Renderer.cpp:169)
    no_clip: bool = False


@dataclass(eq=False)
class FrameSpec:
    width: int
    height: int
    projection: np.ndarray
    view: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=F32))
    draws: List[Draw] = field(default_factory=list)
    light: Sequence[float] = (-0.15, 0.6, 1.0)
    light_is_point: bool = False
    background: Sequence[float] = (0.0, 0.0, 0.0)
    ambient: float = 0.1
    lighting: bool = True
    texturing: bool = True
    save_normals: bool = False

    @property
    def ntri(self) -> int:
        return sum(d.mesh.ntri for d in self.draws)


# ---------------------------------------------------------------------------
# host math shared by every consumer of a frame
# ---------------------------------------------------------------------------
def affine_apply(m: np.ndarray, p: Sequence[float]) -> np.ndarray:
    """Matrix4 * Vec3 (affine), ((m0 x + m1 y) + m2 z) + m3 per row, fp32."""
    m = np.asarray(m, dtype=F32)
    x, y, z = (F32(v) for v in p)
    return np.array([((m[r, 0] * x + m[r, 1] * y) + m[r, 2] * z) + m[r, 3] for r in range(3)],
                    dtype=F32)


def normalized(v: Sequence[float]) -> np.ndarray:
    v = np.asarray(v, dtype=F32)
    ln = np.sqrt(F32((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]))
    return (v / ln).astype(F32)


def eye_light(frame: FrameSpec) -> np.ndarray:
    """Renderer.cpp:317-320: view * light (point) or normalized light."""
    if frame.light_is_point:
        return affine_apply(frame.view, frame.light)
    return normalized(frame.light)


def modelview(frame: FrameSpec, draw: Draw) -> np.ndarray:
    return (np.asarray(frame.view, np.float64) @ np.asarray(draw.transform, np.float64)).astype(F32)


def normal_matrix(mv: np.ndarray) -> np.ndarray:
    return np.linalg.inv(np.asarray(mv, np.float64)).T.astype(F32)


def translate(x, y, z) -> np.ndarray:
    m = np.eye(4, dtype=F32)
    m[:3, 3] = (x, y, z)
    return m


def scale(s) -> np.ndarray:
    m = np.eye(4, dtype=F32)
    m[0, 0] = m[1, 1] = m[2, 2] = s
    return m


def rotate(axis, angle) -> np.ndarray:
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    x, y, z = a
    r = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    m = np.eye(4, dtype=F32)
    m[:3, :3] = r
    return m


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m.astype(F32)


# ---------------------------------------------------------------------------
# projections (Renderer.cpp:40-83), fp32 in the reference's expression order
# ---------------------------------------------------------------------------
def projection_ortho(l, r, b, t, n, f) -> np.ndarray:
    l, r, b, t, n, f = (F32(v) for v in (l, r, b, t, n, f))
    two = F32(2)
    return np.array([[two / (r - l), 0, 0, -(r + l) / (r - l)],
                     [0, two / (t - b), 0, -(t + b) / (t - b)],
                     [0, 0, -two / (f - n), -(f + n) / (f - n)],
                     [0, 0, 0, 1]], dtype=F32)


def projection_perspective(l, r, b, t, n, f) -> np.ndarray:
    l, r, b, t, n, f = (F32(v) for v in (l, r, b, t, n, f))
    two = F32(2)
    return np.array([[two * n / (r - l), 0, (r + l) / (r - l), 0],
                     [0, two * n / (t - b), (t + b) / (t - b), 0],
                     [0, 0, -(f + n) / (f - n), -two * f * n / (f - n)],
                     [0, 0, -1, 0]], dtype=F32)


def projection_frustum(fovy, aspect, n, f) -> np.ndarray:
    fovy, aspect, n = F32(fovy), F32(aspect), F32(n)
    t = F32(math.tan(float(fovy / F32(2))))
    return projection_perspective(-n * t * aspect, n * t * aspect, -n * t, n * t, n, f)


# ---------------------------------------------------------------------------
# reference primitives (src/primitives.cpp), same vertex / index order
# ---------------------------------------------------------------------------
def create_cube(size: float = 1.0) -> Mesh:
    """createCube, primitives.cpp:7-48: 24 vertices, 12 triangles, per-face
    normals; texcoords are filled but texcoordsI is not, so uv = 0."""
    h = F32(size) * F32(0.5)
    faces = [
        ((h, -h, -h), (h, -h, h), (h, h, h), (h, h, -h), (1, 0, 0)),
        ((-h, -h, h), (-h, -h, -h), (-h, h, -h), (-h, h, h), (-1, 0, 0)),
        ((-h, h, -h), (h, h, -h), (h, h, h), (-h, h, h), (0, 1, 0)),
        ((-h, -h, h), (h, -h, h), (h, -h, -h), (-h, -h, -h), (0, -1, 0)),
        ((h, -h, h), (-h, -h, h), (-h, h, h), (h, h, h), (0, 0, 1)),
        ((-h, -h, -h), (h, -h, -h), (h, h, -h), (-h, h, -h), (0, 0, -1)),
    ]
    pos, nrm, uv, idx = [], [], [], []
    for f in faces:
        base = len(pos)
        pos += [f[0], f[1], f[2], f[3]]
        nrm += [f[4]] * 4
        uv += [(0, 0), (1, 0), (1, 1), (0, 1)]
        idx += [(base + 0, base + 2, base + 1), (base + 0, base + 3, base + 2)]
    return Mesh(np.array(pos, F32), np.array(nrm, F32), np.array(idx, np.int32),
                uvs=np.array(uv, F32))


def create_cylinder(radius: float, height: float, segments: int = 32,
                    height_segments: int = 1, caps: bool = True) -> Mesh:
    """createCylinder, primitives.cpp:50-148 (texcoordsI unset: uv = 0)."""
    segments = max(segments, 3)
    height_segments = max(height_segments, 1)
    radius, height = F32(radius), F32(height)
    half = height * F32(0.5)
    two_pi = F32(2.0) * PIF
    pos, nrm, uv, idx = [], [], [], []
    for hh in range(height_segments + 1):
        v = F32(hh) / F32(height_segments)
        z = -half + v * height
        for r in range(segments):
            u = F32(r) / F32(segments)
            ang = u * two_pi
            x = radius * F32(math.cos(ang))
            y = radius * F32(math.sin(ang))
            pos.append((x, y, z))
            nrm.append(normalized((x, y, F32(0))))
            uv.append((u, v))
    for hh in range(height_segments):
        for r in range(segments):
            i0 = hh * segments + r
            i1 = hh * segments + (r + 1) % segments
            i2 = (hh + 1) * segments + r
            i3 = (hh + 1) * segments + (r + 1) % segments
            idx += [(i0, i1, i2), (i1, i3, i2)]
    if caps:
        ring = len(pos)
        for r in range(segments):
            p = pos[r]
            pos.append(p)
            nrm.append((0, 0, -1))
            uv.append(((p[0] / radius + F32(1)) * F32(0.5), (p[1] / radius + F32(1)) * F32(0.5)))
        center = len(pos)
        pos.append((0, 0, -half))
        nrm.append((0, 0, -1))
        uv.append((0.5, 0.5))
        for r in range(segments):
            idx.append((center, ring + (r + 1) % segments, ring + r))
        top_orig = height_segments * segments
        ring = len(pos)
        for r in range(segments):
            p = pos[top_orig + r]
            pos.append(p)
            nrm.append((0, 0, 1))
            uv.append(((p[0] / radius + F32(1)) * F32(0.5), (p[1] / radius + F32(1)) * F32(0.5)))
        center = len(pos)
        pos.append((0, 0, half))
        nrm.append((0, 0, 1))
        uv.append((0.5, 0.5))
        for r in range(segments):
            idx.append((center, ring + r, ring + (r + 1) % segments))
    return Mesh(np.array(pos, F32), np.array(nrm, F32), np.array(idx, np.int32),
                uvs=np.array(uv, F32))


def create_sphere(radius: float, lat_segments: int = 16, long_segments: int = 32) -> Mesh:
    """createSphere, primitives.cpp:150-230.  The top fan (0, i1, i2) is wound
    opposite to the bands, so it is back-face culled seen from outside.
    sin/cos are evaluated per ring and per meridian in double and rounded to
    fp32; every product is an fp32 operation in the order x = (r sin phi) cos th."""
    lat_segments = max(lat_segments, 2)
    long_segments = max(long_segments, 3)
    radius = F32(radius)
    L = long_segments
    lats = np.arange(1, lat_segments)
    v = lats.astype(F32) / F32(lat_segments)
    phi = v * PIF
    sp = np.array([math.sin(float(p)) for p in phi], F32)
    cp = np.array([math.cos(float(p)) for p in phi], F32)
    u = np.arange(L).astype(F32) / F32(L)
    th = u * F32(2.0) * PIF
    ct = np.array([math.cos(float(t)) for t in th], F32)
    st = np.array([math.sin(float(t)) for t in th], F32)
    rs = (radius * sp)[:, None]
    x = rs * ct[None, :]
    y = np.broadcast_to((radius * cp)[:, None], x.shape)
    z = rs * st[None, :]
    ring = np.stack([x, y, z], -1).reshape(-1, 3).astype(F32)
    ln = np.sqrt((ring[:, 0] * ring[:, 0] + ring[:, 1] * ring[:, 1]) + ring[:, 2] * ring[:, 2])
    ring_n = (ring / ln[:, None]).astype(F32)
    ring_uv = np.stack([np.broadcast_to(u[None, :], x.shape),
                        np.broadcast_to((F32(1.0) - v)[:, None], x.shape)], -1).reshape(-1, 2)
    pos = np.concatenate([np.array([[0.0, radius, 0.0]], F32), ring,
                          np.array([[0.0, -radius, 0.0]], F32)])
    nrm = np.concatenate([np.array([[0.0, 1.0, 0.0]], F32), ring_n,
                          np.array([[0.0, -1.0, 0.0]], F32)])
    uv = np.concatenate([np.array([[0.5, 0.0]], F32), ring_uv.astype(F32),
                         np.array([[0.5, 1.0]], F32)])
    first, bottom = 1, len(pos) - 1
    lon = np.arange(L)
    nxt = (lon + 1) % L
    parts = [np.stack([np.zeros(L, np.int64), first + lon, first + nxt], -1)]
    rings = lat_segments - 1
    if rings >= 2:
        row1 = first + np.arange(rings - 1)[:, None] * L
        row2 = row1 + L
        a, b = row1 + lon[None, :], row1 + nxt[None, :]
        c, d = row2 + lon[None, :], row2 + nxt[None, :]
        band = np.stack([np.stack([a, b, c], -1), np.stack([b, d, c], -1)], 2)
        parts.append(band.reshape(-1, 3))
    last = first + (rings - 1) * L
    parts.append(np.stack([np.full(L, bottom), last + lon, last + nxt], -1))
    return Mesh(pos, nrm, np.concatenate(parts).astype(np.int32), uvs=uv)


# ---------------------------------------------------------------------------
# synthetic workloads
# ---------------------------------------------------------------------------
def value_noise(n: int, cells: int, rng: np.random.Generator) -> np.ndarray:
    """Smooth value noise on an n x n grid with `cells` lattice cells."""
    lat = rng.random((cells + 1, cells + 1))
    t = np.linspace(0.0, cells, n, endpoint=False) if n > 1 else np.zeros(1)
    i = np.minimum(t.astype(np.int64), cells - 1)
    f = t - i
    f = f * f * (3 - 2 * f)
    a = lat[i][:, i] * (1 - f)[None, :] + lat[i][:, i + 1] * f[None, :]
    b = lat[i + 1][:, i] * (1 - f)[None, :] + lat[i + 1][:, i + 1] * f[None, :]
    return a * (1 - f)[:, None] + b * f[:, None]


def fbm(n: int, seed: int, octaves: int = 6) -> np.ndarray:
    rng = np.random.default_rng(seed)
    h = np.zeros((n, n))
    amp, cells = 1.0, 4
    for _ in range(octaves):
        h += amp * value_noise(n, cells, rng)
        amp *= 0.5
        cells *= 2
    h -= h.min()
    return h / max(h.max(), 1e-12)


def height_field(n: int, seed: int, amp: float = 0.2, extent: float = 1.0,
                 uv: bool = False) -> Mesh:
    """n x n vertices over [-extent, extent]^2 in x/z, height y = fBm in
    [0, amp], normals from central differences, 2 (n-1)^2 triangles wound
    counter-clockwise seen from +y.  With uv, uv = (j/(n-1), i/(n-1)) and
    uv_indices = indices."""
    h = fbm(n, seed) * amp
    g = np.linspace(-extent, extent, n)
    xs, zs = np.meshgrid(g, g)  # [i, j] -> (x = g[j], z = g[i])
    pos = np.stack([xs, h, zs], axis=-1).reshape(-1, 3).astype(F32)
    dx = np.gradient(h, g, axis=1)
    dz = np.gradient(h, g, axis=0)
    nv = np.stack([-dx, np.ones_like(h), -dz], axis=-1)
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    nrm = nv.reshape(-1, 3).astype(F32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).ravel()
    b = a + 1
    c = a + n
    d = c + 1
    idx = np.empty((2 * a.size, 3), np.int32)
    idx[0::2] = np.stack([a, c, b], -1)
    idx[1::2] = np.stack([b, c, d], -1)
    if not uv:
        return Mesh(pos, nrm, idx)
    u = np.linspace(0.0, 1.0, n)
    uu, vv = np.meshgrid(u, u)
    uvs = np.stack([uu, vv], -1).reshape(-1, 2).astype(F32)
    return Mesh(pos, nrm, idx, uvs=uvs, uv_indices=idx)


def random_triangles(count: int, seed: int, lo: float = 1e-3, hi: float = 1e-1,
                     center_lo=(-0.5, 0.0, -0.5), center_hi=(0.5, 1.0, 0.5)) -> Mesh:
    """`count` independent triangles, centres uniform in a box, edge scale
    log-uniform in [lo, hi], flat per-triangle normals, random winding."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(center_lo, center_hi, (count, 1, 3))
    s = np.exp(rng.uniform(math.log(lo), math.log(hi), (count, 1, 1)))
    d = rng.normal(size=(count, 3, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    v = (c + s * d).reshape(-1, 3).astype(F32)
    tri = v.reshape(-1, 3, 3).astype(np.float64)
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=-1, keepdims=True), 1e-30)
    nrm = np.repeat(fn, 3, axis=0).astype(F32)
    idx = np.arange(3 * count, dtype=np.int32).reshape(-1, 3)
    uv = rng.random((3 * count, 2)).astype(F32)
    return Mesh(v, nrm, idx, uvs=uv, uv_indices=idx)


def checker_texture(size: int, seed: int, checks: int = 32) -> Texture:
    """RGB8 checkerboard modulated by value noise."""
    rng = np.random.default_rng(seed)
    noise = value_noise(size, 16, rng)
    i, j = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    chk = ((i * checks // size + j * checks // size) & 1).astype(np.float64)
    base = np.array([[0.9, 0.85, 0.7], [0.25, 0.35, 0.55]])
    col = base[chk.astype(int)] * (0.6 + 0.4 * noise)[..., None]
    return Texture(np.clip(col * 255.0, 0, 255).astype(np.uint8))


# ---------------------------------------------------------------------------
# benchmark configurations
# ---------------------------------------------------------------------------
def config_sphere(width: int = 640, height: int = 480) -> FrameSpec:
    """Config 1: UV sphere (64 longitude x 32 latitude), 640x480, perspective
    fovy 45 deg, camera at distance 3, point light, Blinn-Phong."""
    mesh = create_sphere(1.0, 32, 64)
    mat = Material(diffuse=(0.7, 0.5, 0.3), specular=(0.5, 0.5, 0.5), shininess=32.0)
    return FrameSpec(width, height,
                     projection_frustum(math.radians(45.0), width / height, 0.1, 100.0),
                     view=translate(0.0, 0.0, -3.0), draws=[Draw(mesh, mat)],
                     light=(2.0, 3.0, 4.0), light_is_point=True, ambient=0.1)


def shape_grid(seed: int, k: int = 10, hf_seed: int = 7, hf_amp: float = 0.2) -> List[Draw]:
    """A k x k grid of seeded spheres / cubes / cylinders over the terrain."""
    rng = np.random.default_rng(seed)
    shapes = [create_sphere(1.0, 16, 32), create_cube(1.0), create_cylinder(0.5, 1.0, 32, 1, True)]
    draws = []
    hgrid = fbm(64, hf_seed) * hf_amp
    for a in range(k):
        for b in range(k):
            x = -0.9 + 1.8 * (b + 0.5) / k
            z = -0.9 + 1.8 * (a + 0.5) / k
            gi = int((z + 1) / 2 * 63)
            gj = int((x + 1) / 2 * 63)
            s = float(rng.uniform(0.03, 0.06))
            y = float(hgrid[gi, gj]) + s
            xf = translate(x, y, z) @ rotate(rng.normal(size=3), float(rng.uniform(0, 6.28))) @ scale(s)
            mat = Material(diffuse=tuple(rng.uniform(0.2, 0.9, 3)), specular=(0.5, 0.5, 0.5),
                           shininess=float(rng.choice([8.0, 16.0, 32.0])))
            draws.append(Draw(shapes[int(rng.integers(0, 3))], mat, xf.astype(F32)))
    return draws


def terrain_camera(width: int, height: int) -> dict:
    """Camera inside the scene's bounding box, low over the terrain, so the
    near plane clips (configs 2 and 3)."""
    return dict(projection=projection_frustum(math.radians(45.0), width / height, 0.05, 100.0),
                view=look_at((0.05, 0.16, 0.85), (0.0, 0.0, -0.6)))


def config_terrain(n: int = 1024, width: int = 1920, height: int = 1080, seed: int = 7,
                   shapes: bool = True) -> FrameSpec:
    """Config 2: fBm height-field (n x n vertices) + 10 x 10 shape grid."""
    hf = height_field(n, seed, 0.2)
    draws = [Draw(hf, Material(diffuse=(0.45, 0.6, 0.35), specular=(0.2, 0.2, 0.2), shininess=16.0))]
    if shapes:
        draws += shape_grid(seed + 1, 10, seed)
    return FrameSpec(width, height, draws=draws, light=(1.0, 3.0, 2.0), light_is_point=True,
                     ambient=0.1, **terrain_camera(width, height))


def config_textured(n: int = 1024, width: int = 1920, height: int = 1080, seed: int = 7,
                    ntris: int = 5000, tex_size: int = 2048) -> FrameSpec:
    """Config 3: textured height-field + 2048^2 RGB8 texture + random triangles."""
    hf = height_field(n, seed, 0.2, uv=True)
    tex = checker_texture(tex_size, seed + 2)
    rt = random_triangles(ntris, seed + 3, center_lo=(-0.8, 0.05, -0.8), center_hi=(0.8, 0.5, 0.8))
    draws = [Draw(hf, Material(diffuse=(1, 1, 1), specular=(0.2, 0.2, 0.2), shininess=16.0, texture=tex)),
             Draw(rt, Material(diffuse=(0.8, 0.3, 0.3), specular=(0.4, 0.4, 0.4), shininess=24.0))]
    return FrameSpec(width, height, draws=draws, light=(1.0, 3.0, 2.0), light_is_point=True,
                     ambient=0.1, **terrain_camera(width, height))


def grid_side_for(tris: int) -> int:
    """Vertices per side of a height-field with about `tris` triangles."""
    return max(2, int(round(math.sqrt(tris / 2.0))) + 1)


def config_megatri(width: int = 1920, height: int = 1080, seed: int = 11, side: int = 708) -> FrameSpec:
    """The 1M-triangle grid (708 x 708 vertices, 999,698 triangles) at 1080p,
    Blinn-Phong, whole mesh in view."""
    hf = height_field(side, seed, 0.2)
    mat = Material(diffuse=(0.55, 0.6, 0.7), specular=(0.5, 0.5, 0.5), shininess=32.0)
    return FrameSpec(width, height,
                     projection=projection_frustum(math.radians(45.0), width / height, 0.1, 100.0),
                     view=look_at((0.0, 1.9, 1.9), (0.0, 0.0, 0.0)),
                     draws=[Draw(hf, mat)], light=(1.0, 3.0, 2.0), light_is_point=True, ambient=0.1)


RESOLUTIONS = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def orbit_camera(rng: np.random.Generator, width: int, height: int, rmin=2.0, rmax=5.0) -> dict:
    r = float(rng.uniform(rmin, rmax))
    th = float(rng.uniform(0, 2 * math.pi))
    ph = float(rng.uniform(0.15, 1.3))
    eye = (r * math.cos(ph) * math.cos(th), r * math.sin(ph), r * math.cos(ph) * math.sin(th))
    return dict(projection=projection_frustum(math.radians(45.0), width / height, 0.1, 100.0),
                view=look_at(eye, (0.0, 0.0, 0.0)))


def megatri_requests(n: int, seed: int = 13, width: int = 1920, height: int = 1080) -> List[FrameSpec]:
    """`n` render requests of the 1M-triangle grid (one shared mesh), each with
    its own camera on a sphere of radius 2-5 looking at the origin and its own
    point light."""
    base = config_megatri(width, height)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        light = tuple(float(v) for v in rng.uniform(-4.0, 4.0, 3))
        out.append(FrameSpec(width, height, draws=base.draws, light=light, light_is_point=True,
                             ambient=0.1, **orbit_camera(rng, width, height)))
    return out


def split_materials(mesh: Mesh, parts: int, rng: np.random.Generator) -> List[Draw]:
    """Split a mesh's triangles into `parts` consecutive meshes (one per
    material, like the OBJ loader's per-usemtl meshes, io.cpp:250-260)."""
    cuts = np.linspace(0, mesh.ntri, parts + 1).astype(int)
    draws = []
    for p in range(parts):
        sl = slice(cuts[p], cuts[p + 1])
        m = Mesh(mesh.positions, mesh.normals, mesh.indices[sl],
                 None if mesh.normal_indices is None else mesh.normal_indices[sl],
                 mesh.uvs, None if mesh.uv_indices is None else mesh.uv_indices[sl])
        mat = Material(diffuse=tuple(rng.uniform(0.2, 0.9, 3)), specular=(0.5, 0.5, 0.5),
                       shininess=float(rng.choice([0.0, 8.0, 16.0, 32.0])))
        draws.append(Draw(m, mat))
    return draws


def serving_pool(nmeshes: int = 64, seed: int = 9, tri_lo: int = 10_000,
                 tri_hi: int = 1_000_000) -> List[List[Draw]]:
    """Config 5's resident pool: `nmeshes` seeded meshes (height-fields or
    spheres) of log-uniform triangle count, each split into 1-8 materials.
    Returns one draw list per mesh."""
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(nmeshes):
        t = int(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi))))
        if rng.random() < 0.5:
            mesh = height_field(grid_side_for(t), int(rng.integers(1 << 30)), 0.3)
        else:
            lat = max(3, int(math.sqrt(t / 4)))
            mesh = create_sphere(1.0, lat, 2 * lat)
        pool.append(split_materials(mesh, int(rng.integers(1, 9)), rng))
    return pool


def serving_requests(pool: List[List[Draw]], n: int, seed: int = 10) -> List[FrameSpec]:
    """`n` requests over a resident pool: a uniformly chosen mesh with its own
    camera on a sphere of radius 2-5, point light and resolution."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        draws = pool[int(rng.integers(0, len(pool)))]
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        light = tuple(float(v) for v in rng.uniform(-4.0, 4.0, 3))
        out.append(FrameSpec(w, h, draws=draws, light=light, light_is_point=True, ambient=0.1,
                             **orbit_camera(rng, w, h)))
    return out


def batch_jobs(njobs: int = 256, seed: int = 5, tri_lo: int = 1000, tri_hi: int = 500_000) -> List[FrameSpec]:
    """Config 4: independent jobs, each a seeded mesh of log-uniform triangle
    count, 1-8 materials, resolution from RESOLUTIONS, camera on a sphere of
    radius 2-5 looking at the origin."""
    rng = np.random.default_rng(seed)
    frames = []
    for j in range(njobs):
        t = int(math.exp(rng.uniform(math.log(tri_lo), math.log(tri_hi))))
        if rng.random() < 0.5:
            mesh = height_field(grid_side_for(t), int(rng.integers(1 << 30)), 0.3)
        else:
            lat = max(3, int(math.sqrt(t / 4)))
            mesh = create_sphere(1.0, lat, 2 * lat)
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        draws = split_materials(mesh, int(rng.integers(1, 9)), rng)
        frames.append(FrameSpec(w, h, draws=draws, light=(2.0, 4.0, 3.0), light_is_point=True,
                                ambient=0.1, **orbit_camera(rng, w, h)))
    return frames
