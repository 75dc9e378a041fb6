"""Deterministic, seeded workloads for the frame path.

All randomness comes from one counter-based generator (SplitMix64) so every scene is
a pure function of its seed, independent of numpy's generator versions.

Workloads (sizes as the benchmark and parity tests use them):
  uv_sphere_scene     UV sphere 32 lon x 16 lat (960 faces), 512x512, fovy 60, near 0.1,
                      far 100, three seeded lights — the CPU-runnable case.
  icosphere_scene     icosphere subdivided 6 times (81,920 faces) radially displaced by
                      seeded value noise (amplitude 0.1), spherical uvs with seam
                      duplicates, a seeded 1024x1024 RGBA8 noise texture, three lights,
                      1920x1080, seeded YXZ rotation in [-pi, pi]^3; flat or smooth.
  batch_scenes        256 meshes of log-uniform face counts in [500, 200k]: displaced
                      icospheres (subdivision chosen to hold the count, then
                      truncated) or random triangle soups (edges U[0.01, 0.2] in the
                      unit cube), each with its own square viewport (side 256..1024),
                      transform (scale U[0.5,2], displacement U[-1,1]^3, rotation
                      U[-pi,pi]^3; resampled until every vertex lies beyond the near
                      plane) and 1-8 lights.
  learned_shading_scene  the icosphere workload with 64 lights and a seeded 1024x1024
                      tangent-space normal map whose unit normals lie within 30 degrees
                      of +z (the learned-shader comparison).
  screen_soup_scene   N seeded screen-space triangles (edges 0.5-3 px, eye depth
                      U[0.1, 100] through the fovy-60/0.1/100 projection, about half
                      of them wound backwards) on a square image.

Light intensity is U[0.2, 1] * 255*pi: at a literal U[0.2, 1] every pixel of the
reference shader truncates to 0 (shading.cpp:21,31-33), so colour parity would test
nothing.  The view is the reference's fixed translate(0, 0, -3) (drawing.cpp:224).
"""
from __future__ import annotations

import math

import numpy as np

from .abi import make_args
from .scene import MaterialSpec, Scene, make_faces, make_lights

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    """Counter-based SplitMix64 stream: value k is mix(seed + (k+1) * golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def next_u64(self, n: int) -> np.ndarray:
        k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.state + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def bytes(self, n: int) -> np.ndarray:
        words = self.next_u64((n + 7) // 8)
        return words.view(np.uint8)[:n].copy()


def random_lights(rng: SplitMix64, n: int, scale: float = 255.0 * math.pi) -> np.ndarray:
    """n lights: direction uniform on the sphere, intensity U[0.2,1]*scale, RGB U[0,1]."""
    z = rng.uniform(n, -1.0, 1.0)
    phi = rng.uniform(n, 0.0, 2.0 * math.pi)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    inten = rng.uniform(n, 0.2, 1.0) * scale
    cols = rng.uniform(3 * n, 0.0, 1.0).reshape(n, 3)
    return make_lights(dirs, inten, cols)


def _orient_outward(pos: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """Flip faces so (p1-p0) x (p2-p0) points away from the origin (anticlockwise from outside)."""
    p0, p1, p2 = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    flip = np.einsum("ij,ij->i", nrm, p0 + p1 + p2) < 0
    tri = tri.copy()
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return tri


def uv_sphere(nlon: int = 32, nlat: int = 16):
    """Unit UV sphere: 2 poles + (nlat-1)*nlon ring vertices, 2*nlon*(nlat-1) faces."""
    pos = [(0.0, 1.0, 0.0)]
    for i in range(1, nlat):
        th = math.pi * i / nlat
        for j in range(nlon):
            ph = 2.0 * math.pi * j / nlon
            pos.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
    pos.append((0.0, -1.0, 0.0))
    pos = np.array(pos, np.float64)
    south = len(pos) - 1

    def ring(i, j):
        return 1 + (i - 1) * nlon + (j % nlon)

    tris = []
    for j in range(nlon):
        tris.append((0, ring(1, j), ring(1, j + 1)))
    for i in range(1, nlat - 1):
        for j in range(nlon):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            tris.append((a, c, b))
            tris.append((b, c, d))
    for j in range(nlon):
        tris.append((south, ring(nlat - 1, j + 1), ring(nlat - 1, j)))
    tri = _orient_outward(pos, np.array(tris, np.int64))
    return pos, tri


def icosphere(subdivisions: int):
    """Unit icosphere: 20 * 4^s faces, 10 * 4^s + 2 vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                  (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                  (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                  (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        nf = f.shape[0]
        m01 = inv[:nf] + v.shape[0]
        m12 = inv[nf:2 * nf] + v.shape[0]
        m20 = inv[2 * nf:] + v.shape[0]
        v = np.concatenate([v, mid])
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, _orient_outward(v, f)


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """3D lattice value noise in [-1, 1], smoothstep-trilinear between hashed corners."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    fr = q - i0
    s = fr * fr * (3.0 - 2.0 * fr)

    def corner(dx, dy, dz):
        ix, iy, iz = (i0[:, 0] + dx), (i0[:, 1] + dy), (i0[:, 2] + dz)
        h = ((ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ seed).astype(np.uint64)
        with np.errstate(over="ignore"):
            z = h * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0

    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        wx = s[:, 0] if dx else 1.0 - s[:, 0]
        for dy in (0, 1):
            wy = s[:, 1] if dy else 1.0 - s[:, 1]
            for dz in (0, 1):
                wz = s[:, 2] if dz else 1.0 - s[:, 2]
                out += wx * wy * wz * corner(dx, dy, dz)
    return out


def vertex_normals(pos: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """Area-weighted smooth normals."""
    p0, p1, p2 = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    fn = np.cross(p1 - p0, p2 - p0)
    acc = np.zeros_like(pos)
    for k in range(3):
        np.add.at(acc, tri[:, k], fn)
    ln = np.linalg.norm(acc, axis=1, keepdims=True)
    out = np.tile(np.array([0.0, 0.0, 1.0]), (pos.shape[0], 1))   # vertices no face uses
    np.divide(acc, ln, out=out, where=ln > 0)
    return out


def spherical_uvs(dirs: np.ndarray, tri: np.ndarray):
    """u = 0.5 + atan2(z, x)/2pi, v = 0.5 + asin(y)/pi; faces across the seam get
    duplicated uv entries shifted by +1 in u.  Returns (uvs, uv index per corner)."""
    u = 0.5 + np.arctan2(dirs[:, 2], dirs[:, 0]) / (2.0 * math.pi)
    v = 0.5 + np.arcsin(np.clip(dirs[:, 1], -1.0, 1.0)) / math.pi
    uvs = [np.stack([u, v], 1)]
    tuv = tri.copy()
    cu = u[tri]
    seam = (cu.max(1) - cu.min(1)) > 0.5
    extra = {}
    nxt = len(u)
    rows, cols = np.nonzero(seam[:, None] & (cu < 0.5))
    add = []
    for r, c in zip(rows.tolist(), cols.tolist()):
        vi = int(tri[r, c])
        if vi not in extra:
            extra[vi] = nxt
            nxt += 1
            add.append((u[vi] + 1.0, v[vi]))
        tuv[r, c] = extra[vi]
    if add:
        uvs.append(np.array(add, np.float64))
    return np.concatenate(uvs), tuv


def noise_texture(rng: SplitMix64, width: int = 1024, height: int = 1024, channels: int = 4):
    tex = rng.bytes(width * height * channels).reshape(height, width, channels)
    tex[0, 0, 0] = 0      # the reference normalises by min/max over every channel:
    tex[0, 1, 0] = 255    # pin them at 0 and 255 so texels are exactly v/255
    return tex


def cone_normal_map(rng: SplitMix64, width: int = 1024, height: int = 1024, cone_deg: float = 30.0):
    """Unit normals uniform on the cap within cone_deg of +z, stored as round((n+1)/2*255)."""
    n = width * height
    cz = rng.uniform(n, math.cos(math.radians(cone_deg)), 1.0)
    phi = rng.uniform(n, 0.0, 2.0 * math.pi)
    sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
    nrm = np.stack([sz * np.cos(phi), sz * np.sin(phi), cz], 1)
    return np.clip(np.rint((nrm + 1.0) * 0.5 * 255.0), 0, 255).astype(np.uint8).reshape(height, width, 3)


def random_angles(rng: SplitMix64):
    return tuple(float(np.float32(a)) for a in rng.uniform(3, -math.pi, math.pi))


def uv_sphere_scene(seed: int = 1, width: int = 512, height: int = 512):
    """The 960-face UV sphere, smooth normals, Kd = 1, three lights."""
    rng = SplitMix64(0x5EED0000 + seed)
    pos, tri = uv_sphere(32, 16)
    uvs = np.zeros((1, 2))
    faces = make_faces(tri, n=tri, t=None)
    scene = Scene(pos.astype(np.float32), pos.astype(np.float32), uvs, faces,
                  [MaterialSpec((1.0, 1.0, 1.0))], random_lights(rng, 3))
    args = make_args(width, height, fovy=60.0, z_near=0.1, z_far=100.0)
    return scene, args


def icosphere_scene(seed: int = 2, width: int = 1920, height: int = 1080, flat: bool = False,
                    subdivisions: int = 6, texture_size: int = 1024, lights: int = 3):
    """Displaced, textured icosphere (81,920 faces at 6 subdivisions) at 1080p."""
    rng = SplitMix64(0x5EED0000 + seed)
    unit, tri = icosphere(subdivisions)
    r = 1.0 + 0.1 * value_noise(unit, seed)
    pos = unit * r[:, None]
    nrm = vertex_normals(pos, tri)
    uvs, tuv = spherical_uvs(unit, tri)
    tex = noise_texture(rng, texture_size, texture_size, 4)
    faces = make_faces(tri, n=tri, t=tuv)
    scene = Scene(pos.astype(np.float32), nrm.astype(np.float32), uvs.astype(np.float32), faces,
                  [MaterialSpec((0.8, 0.8, 0.8), tex)], random_lights(rng, lights))
    args = make_args(width, height, fovy=60.0, z_near=0.1, z_far=100.0, flat=flat,
                     angles=random_angles(rng))
    return scene, args


def large_triangle_scene(seed: int = 9, n: int = 3000, size: int = 1024, near: float = 0.1,
                         far: float = 100.0):
    """Screen-space triangles far larger than a tile: a third with edges 20-600 px, a third
    slivers (the third vertex within 1e-3..1 px of the line through the other two), a
    third with a vertex up to 5,000 px off screen; eye depths and winding as in
    screen_soup_scene.  Exercises the raster pass's reduction of large candidate
    rectangles, whose exactness rests on the rounding bounds of the edge functions."""
    rng = SplitMix64(0x5EED0000 + seed)
    kind = (rng.next_u64(n) % np.uint64(3)).astype(np.int64)
    c = rng.uniform(2 * n, 0.0, float(size)).reshape(n, 2)
    ang = rng.uniform(n, 0.0, 2.0 * math.pi)
    edge = rng.uniform(n, 20.0, 600.0)
    jit = rng.uniform(2 * n, -0.5, 0.5).reshape(n, 2)
    rad = edge / math.sqrt(3.0)
    a = np.stack([ang, ang + 2.0 * math.pi / 3.0 + jit[:, 0], ang + 4.0 * math.pi / 3.0 + jit[:, 1]], 1)
    xs = c[:, 0:1] + rad[:, None] * np.cos(a)
    ys = c[:, 1:2] - rad[:, None] * np.sin(a)
    # slivers: vertex 2 near the segment (0, 1), offset along its normal by 1e-3..1 px
    sl = kind == 1
    t = rng.uniform(n, 0.1, 0.9)
    off = np.exp(rng.uniform(n, math.log(1e-3), 0.0)) * np.where(rng.uniform(n) < 0.5, -1.0, 1.0)
    dx, dy = xs[:, 1] - xs[:, 0], ys[:, 1] - ys[:, 0]
    ln = np.maximum(np.hypot(dx, dy), 1e-9)
    xs[sl, 2] = (xs[:, 0] + t * dx - off * dy / ln)[sl]
    ys[sl, 2] = (ys[:, 0] + t * dy + off * dx / ln)[sl]
    # far vertices: vertex 0 pushed up to 5,000 px away
    fa = kind == 2
    far_d = rng.uniform(n, 500.0, 5000.0)
    fang = rng.uniform(n, 0.0, 2.0 * math.pi)
    xs[fa, 0] = (c[:, 0] + far_d * np.cos(fang))[fa]
    ys[fa, 0] = (c[:, 1] + far_d * np.sin(fang))[fa]
    d = rng.uniform(3 * n, 0.1, 100.0).reshape(n, 3)
    z = (far + near) / (far - near) - 2.0 * far * near / ((far - near) * d)
    verts = np.stack([xs, ys, z, 1.0 / d], axis=2).reshape(-1, 4).astype(np.float32)
    tri = np.arange(3 * n, dtype=np.int64).reshape(n, 3)
    back = rng.uniform(n) < 0.5
    tri[back, 1], tri[back, 2] = tri[back, 2].copy(), tri[back, 1].copy()
    faces = make_faces(tri, n=np.zeros_like(tri), t=None)
    scene = Scene(verts, np.array([[0.0, 0.0, 1.0]], np.float32), np.zeros((1, 2), np.float32), faces,
                  [MaterialSpec((1.0, 1.0, 1.0))],
                  make_lights([[0.0, 0.0, -1.0]], [255.0 * math.pi], [[1.0, 1.0, 1.0]]),
                  raster_space=True)
    return scene, make_args(size, size, fovy=60.0, z_near=near, z_far=far)


def screen_soup_scene(seed: int = 4, n: int = 4_000_000, size: int = 4096,
                      near: float = 0.1, far: float = 100.0):
    """n screen-space triangles: centroid uniform, edge 0.5-3 px, random angle, per-vertex
    eye depth d ~ U[0.1, 100] mapped to z = (f+n)/(f-n) - 2fn/((f-n) d), w = 1/d; about
    half are wound backwards."""
    rng = SplitMix64(0x5EED0000 + seed)
    c = rng.uniform(2 * n, 0.0, float(size)).reshape(n, 2)
    edge = rng.uniform(n, 0.5, 3.0)
    ang = rng.uniform(n, 0.0, 2.0 * math.pi)
    jit = rng.uniform(2 * n, -0.3, 0.3).reshape(n, 2)
    rad = edge / math.sqrt(3.0)
    a = np.stack([ang, ang + 2.0 * math.pi / 3.0 + jit[:, 0], ang + 4.0 * math.pi / 3.0 + jit[:, 1]], 1)
    # anticlockwise on screen (y up) = front under the reference's raster y flip
    xs = c[:, 0:1] + rad[:, None] * np.cos(a)
    ys = c[:, 1:2] - rad[:, None] * np.sin(a)
    d = rng.uniform(3 * n, 0.1, 100.0).reshape(n, 3)
    z = (far + near) / (far - near) - 2.0 * far * near / ((far - near) * d)
    verts = np.stack([xs, ys, z, 1.0 / d], axis=2).reshape(-1, 4).astype(np.float32)
    tri = np.arange(3 * n, dtype=np.int64).reshape(n, 3)
    back = rng.uniform(n) < 0.5
    tri[back, 1], tri[back, 2] = tri[back, 2].copy(), tri[back, 1].copy()
    faces = make_faces(tri, n=np.zeros_like(tri), t=None)
    scene = Scene(verts, np.array([[0.0, 0.0, 1.0]], np.float32), np.zeros((1, 2), np.float32), faces,
                  [MaterialSpec((1.0, 1.0, 1.0))],
                  make_lights([[0.0, 0.0, -1.0]], [255.0 * math.pi], [[1.0, 1.0, 1.0]]),
                  raster_space=True)
    args = make_args(size, size, fovy=60.0, z_near=near, z_far=far)
    return scene, args


def learned_shading_scene(seed: int = 5, width: int = 1920, height: int = 1080, lights: int = 64,
                          subdivisions: int = 6, map_size: int = 1024):
    """Icosphere workload with `lights` lights and a cone normal map."""
    scene, args = icosphere_scene(seed=seed, width=width, height=height, subdivisions=subdivisions,
                                  texture_size=map_size, lights=lights)
    rng = SplitMix64(0x5EED1000 + seed)
    scene.materials[0].normal_map = cone_normal_map(rng, map_size, map_size)
    return scene, args


_ICO_CACHE = {}


def _icosphere_cached(subdivisions):
    if subdivisions not in _ICO_CACHE:
        _ICO_CACHE[subdivisions] = icosphere(subdivisions)
    return _ICO_CACHE[subdivisions]


def _eye_depths_ok(pos, args, near):
    """True when every vertex is in front of the near plane (clip w > near) under args."""
    import ctypes
    from .abi import Args
    ax, ay, az = (float(np.float32(t)) for t in args.tait_bryan_angles)
    def rot(axis, t):
        c, s_ = math.cos(t), math.sin(t)
        if axis == 0:
            return np.array([[1, 0, 0], [0, c, -s_], [0, s_, c]])
        if axis == 1:
            return np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]])
        return np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
    R = rot(1, ay) @ rot(0, ax) @ rot(2, az)
    p = (pos.astype(np.float64) * float(args.scale)) @ R.T + np.array(args.displacement[:], np.float64)
    z_eye = p[:, 2] - 3.0
    return bool((-z_eye > near * 1.5).all())


def batch_scenes(n=256, seed=3, near=0.1, far=100.0, max_faces=200_000, min_faces=500):
    """The batched workload: a list of (Scene, Args) pairs (see the module docstring)."""
    rng = SplitMix64(0x5EED0000 + seed)
    out = []
    for i in range(n):
        t = int(round(math.exp(rng.uniform(1, math.log(min_faces), math.log(max_faces))[0])))
        kind = rng.uniform(1)[0] < 0.5
        side = int(256 + (rng.next_u64(1)[0] % np.uint64(769)))
        if kind:
            sub = 0
            while 20 * 4 ** sub < t:
                sub += 1
            unit, tri = _icosphere_cached(sub)
            tri = tri[:t]
            r = 1.0 + 0.1 * value_noise(unit, seed * 1000 + i)
            pos = unit * r[:, None]
            nrm = vertex_normals(pos, tri)
            faces = make_faces(tri, n=tri, t=None)
        else:
            c = rng.uniform(3 * t, 0.0, 1.0).reshape(t, 3)
            e = rng.uniform(t, 0.01, 0.2)
            d1 = rng.uniform(3 * t, -1, 1).reshape(t, 3)
            d2 = rng.uniform(3 * t, -1, 1).reshape(t, 3)
            d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
            d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
            pos = np.concatenate([c, c + e[:, None] * d1, c + e[:, None] * d2], 1).reshape(-1, 3)
            tri = np.arange(3 * t).reshape(t, 3)
            fn = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
            nrm = fn / np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-30)
            faces = make_faces(tri, n=np.repeat(np.arange(t)[:, None], 3, 1), t=None)
        nl = int(1 + (rng.next_u64(1)[0] % np.uint64(8)))
        lights = random_lights(rng, nl)
        col = tuple(float(x) for x in rng.uniform(3, 0.3, 1.0))
        scene = Scene(pos.astype(np.float32), nrm.astype(np.float32), np.zeros((1, 2), np.float32), faces,
                      [MaterialSpec(col)], lights)
        for _ in range(100):
            sc = float(np.float32(rng.uniform(1, 0.5, 2.0)[0]))
            disp = tuple(float(np.float32(x)) for x in rng.uniform(3, -1.0, 1.0))
            args = make_args(side, side, fovy=60.0, z_near=near, z_far=far, scale=sc, displacement=disp,
                             angles=random_angles(rng))
            if _eye_depths_ok(scene.vertices, args, near):
                break
        out.append((scene, args))
    return out


def square_scene():
    """The reference's default scene when no OBJ is given (renderer.cpp:32-50)."""
    pos = np.array([[-0.5, -0.5, 0.0], [0.5, -0.5, 0.0], [-0.5, 0.5, 0.0], [0.5, 0.5, 0.0]], np.float32)
    nrm = np.array([[0.0, 0.0, 1.0]], np.float32)
    uvs = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32)
    faces = make_faces([[0, 1, 2], [2, 1, 3]], n=[[0, 0, 0], [0, 0, 0]], t=[[0, 1, 2], [2, 1, 3]])
    return Scene(pos, nrm, uvs, faces, [MaterialSpec((1.0, 1.0, 1.0))])
