"""Pins of the CPU oracle (oracle/rz_oracle.cpp) — CPU only.

The reference ships no tests, golden images or known-answer vectors, and its path
cannot be compiled here (it needs GLM, which is absent).  Each oracle routine is
therefore pinned by what fixes its result independently of its own code:

* exact rational arithmetic (fractions.Fraction) on dyadic inputs, for which every
  binary32 operation of the edge/barycentric/area/bbox code is exact;
* closed forms (projection of known points, Lambert with one light along the normal,
  rotations by known angles);
* the reference's own vendored CImg code (oracle/_ref, or the vectors it wrote into
  tests/golden/cimg_vectors.npz) for the texture normalisation and sampler;
* invariants the reference semantics fix (shared edges covered twice, ties to the
  lowest face index, strict depth test, buffers not cleared).
"""
import math
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from rasteriser_amd import make_args
from rasteriser_amd.scene import MaterialSpec, Scene, make_faces, make_lights
from rasteriser_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cimg_vectors.npz")
F32 = np.float32


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# edge / barycentric / signed area / bbox against exact rationals

def exact_edge(p, a, b):
    return (Fr(b[0]) - Fr(a[0])) * (Fr(p[1]) - Fr(a[1])) - (Fr(b[1]) - Fr(a[1])) * (Fr(p[0]) - Fr(a[0]))


def dyadic(rng, n, lo, hi, q=4):
    return rng.integers(lo * q, hi * q + 1, n) / q


def test_edge_exact_on_dyadic_inputs():
    rng = np.random.default_rng(1)
    for _ in range(300):
        p, a, b = dyadic(rng, 2, -8, 24), dyadic(rng, 2, -8, 24), dyadic(rng, 2, -8, 24)
        got = oracle.edge(p[0], p[1], [a[0], a[1], 0, 1], [b[0], b[1], 0, 1])
        assert Fr(got) == exact_edge(p, a, b)
    # orientation: p to the left of a->b (y up) is positive
    assert oracle.edge(0, 1, [0, 0, 0, 1], [1, 0, 0, 1]) == 1.0


def test_barycentric_vertex_identity_and_exact_values():
    v = [[1.0, 2.0, 0, 1], [9.0, 2.0, 0, 1], [1.0, 10.0, 0, 1]]   # area 64: quotients are exact
    flat = np.array(v, np.float32).reshape(-1)
    for k in range(3):
        b = oracle.barycentric(v[k][0], v[k][1], flat)
        assert list(b) == [1.0 if i == k else 0.0 for i in range(3)]
    rng = np.random.default_rng(2)
    for _ in range(200):
        p = dyadic(rng, 2, -4, 14)
        b = oracle.barycentric(p[0], p[1], flat)
        area = exact_edge(v[2], v[0], v[1])
        want = [exact_edge(p, v[1], v[2]) / area, exact_edge(p, v[2], v[0]) / area,
                exact_edge(p, v[0], v[1]) / area]
        assert [Fr(float(x)) for x in b] == want
        assert sum(want) == 1


def test_signed_area_closed_form():
    # -0.5 * shoelace; the y-up anticlockwise unit triangle has shoelace +1
    assert oracle.signed_area([0, 0, 0, 1, 1, 0, 0, 1, 0, 1, 0, 1]) == -0.5
    assert oracle.signed_area([0, 0, 0, 1, 0, 1, 0, 1, 1, 0, 0, 1]) == 0.5
    rng = np.random.default_rng(3)
    for _ in range(200):
        x = dyadic(rng, 3, -16, 16)
        y = dyadic(rng, 3, -16, 16)
        sh = sum(Fr(x[i]) * Fr(y[(i + 1) % 3]) - Fr(x[(i + 1) % 3]) * Fr(y[i]) for i in range(3))
        v = [x[0], y[0], 0, 1, x[1], y[1], 0, 1, x[2], y[2], 0, 1]
        assert Fr(oracle.signed_area(v)) == Fr(-1, 2) * sh


@pytest.mark.parametrize("verts,W,H,want", [
    ([2.5, 3.5, 7.0, 1.25, 4.0, 9.0], 20, 20, (2, 1, 7, 9)),       # floor(min), ceil(max); ceil(7.0)=7
    ([-3.0, -1.0, 5.5, 2.2, 1.0, 30.0], 10, 12, (0, 0, 6, 11)),      # clamped to the image
    ([25.0, 5.0, 40.0, 6.0, 30.0, 7.5], 20, 20, (19, 5, 19, 8)),     # off screen to the right
    ([0.0, 0.0, 1.0, 0.0, 0.0, 1.0], 4, 4, (0, 0, 1, 1)),
])
def test_bounding_box(verts, W, H, want):
    v = [verts[0], verts[1], 0, 1, verts[2], verts[3], 0, 1, verts[4], verts[5], 0, 1]
    assert oracle.bounding_box(v, W, H) == want


def test_cast_semantics():
    """The float->unsigned casts are UB in C++; the defined meaning is GCC x86-64's."""
    assert oracle.cast_u32(3.99) == 3
    assert oracle.cast_u32(255.0) == 255
    assert oracle.cast_u32(-0.5) == 0
    assert oracle.cast_u32(-1.5) == 0xFFFFFFFF
    assert oracle.cast_u32(float("nan")) == 0
    assert oracle.cast_u32(1e30) == 0
    assert oracle.cast_u32(4294967296.0 * 2 + 4096) == 4096


# ---------------------------------------------------------------------------
# vertex stage and matrices

def np_rot(axis, t):
    c, s = math.cos(t), math.sin(t)
    if axis == "x":
        return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])
    if axis == "y":
        return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]])
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])


def np_persp(fovy_deg, aspect, n, f):
    t = math.tan(math.radians(fovy_deg) / 2)
    return np.array([[1 / (aspect * t), 0, 0, 0], [0, 1 / t, 0, 0],
                     [0, 0, -(f + n) / (f - n), -2 * f * n / (f - n)], [0, 0, -1, 0]])


def as_rows(m_colmajor):
    """(4,4) stored [col][row] -> conventional row-major matrix."""
    return np.asarray(m_colmajor, np.float64).T


def test_matrices_default_args_exact():
    a = make_args(540, 304)
    m = oracle.matrices(a)
    assert np.array_equal(m["model"], np.eye(4, dtype=np.float32))
    view = np.eye(4, dtype=np.float32)
    view[3, 2] = -3.0          # column 3, row 2
    assert np.array_equal(m["view"], view)
    assert np.array_equal(m["modelview"], view)
    P = np_persp(45.0, 540 / 304, 0.1, 6.0)
    assert np.allclose(as_rows(m["camera"]), P @ as_rows(view), rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("angles,scale,disp", [
    ((0.3, -1.1, 2.0), 1.5, (0.25, -0.5, 0.75)),
    ((math.pi / 2, 0.0, 0.0), 1.0, (0.0, 0.0, 0.0)),
    ((-2.5, 0.7, -0.2), 0.6, (1.0, 0.0, -1.0)),
])
def test_model_matrix_is_T_Ry_Rx_Rz_S(angles, scale, disp):
    """transformation_matrix composes YXZ (geometry.cpp:22-25); a different order fails."""
    a = make_args(640, 480, angles=angles, scale=scale, displacement=disp, fovy=60, z_far=100)
    m = oracle.matrices(a)
    ax, ay, az = (f32(t) for t in angles)
    T = np.eye(4)
    T[:3, 3] = [f32(d) for d in disp]
    S = np.diag([f32(scale)] * 3 + [1.0])
    want = T @ np_rot("y", ay) @ np_rot("x", ax) @ np_rot("z", az) @ S
    assert np.allclose(as_rows(m["model"]), want, atol=3e-6)
    view = np.eye(4)
    view[2, 3] = -3
    mv = view @ want
    assert np.allclose(as_rows(m["modelview"]), mv, atol=5e-6)
    assert np.allclose(as_rows(m["camera"]), np_persp(60, 640 / 480, 0.1, 100) @ mv, atol=2e-5)
    # normal matrix = inverse-transpose of the model-view
    assert np.allclose(as_rows(m["normal"]), np.linalg.inv(mv).T, atol=2e-5)


def test_product_matrices_match_oracle_bit_for_bit():
    """librz builds the same matrices on the host (no GPU needed for this call)."""
    rz = pytest.importorskip("rasteriser_amd")
    try:
        rz.load_library()
    except (RuntimeError, OSError) as e:
        pytest.skip(f"librz not loadable: {e}")
    rng = np.random.default_rng(4)
    for _ in range(50):
        a = make_args(int(rng.integers(16, 4000)), int(rng.integers(16, 3000)),
                      angles=tuple(rng.uniform(-4, 4, 3)), scale=float(rng.uniform(0.2, 3)),
                      displacement=tuple(rng.uniform(-2, 2, 3)), fovy=float(rng.uniform(20, 90)),
                      z_near=float(rng.uniform(0.05, 1)), z_far=float(rng.uniform(5, 200)))
        mo, mp = oracle.matrices(a), rz.frame_matrices(a)
        for k in mo:
            assert np.array_equal(mo[k].view(np.uint32), mp[k].view(np.uint32)), k


def test_vertex_to_raster_closed_forms():
    I = np.eye(4, dtype=np.float32).reshape(-1)
    W, H = 640, 480
    assert list(oracle.vertex_to_raster(I, [0, 0, 0.25], W, H)) == [320.0, 240.0, 0.25, 1.0]
    assert list(oracle.vertex_to_raster(I, [1, 1, 0], W, H)) == [640.0, 0.0, 0.0, 1.0]
    assert list(oracle.vertex_to_raster(I, [-1, -1, 0], W, H)) == [0.0, 480.0, 0.0, 1.0]
    # a point straight ahead at eye distance d maps to ndc z = (f+n)/(f-n) - 2fn/((f-n) d), w = 1/d
    a = make_args(W, H, fovy=60, z_near=0.1, z_far=100)
    cam = oracle.matrices(a)["camera"].reshape(-1)
    for d in (0.5, 2.0, 3.0, 10.0, 90.0):
        r = oracle.vertex_to_raster(cam, [0, 0, 3.0 - d], W, H)
        zc = (100.1 / 99.9) - 2 * 100 * 0.1 / (99.9 * d)
        assert abs(r[0] - 320) < 1e-3 and abs(r[1] - 240) < 1e-3
        assert abs(r[2] - zc) < 4e-7 * max(1, abs(zc)) + 2e-7
        assert abs(r[3] - 1 / d) <= 2 * np.spacing(np.float32(1 / d))


# ---------------------------------------------------------------------------
# shading and texture

def light(direction, intensity, colour):
    L = make_lights([direction], [intensity], [colour])
    L[0, 7:10] = L[0, 0:3]   # trans_dir (oracle.shade uses it as given)
    return L


def test_shade_one_light_along_normal():
    n = [0.0, 0.0, 1.0]
    for I, col, alb in [(512.0, (1, 0.9, 0.5), (0.64, 0.5, 0.25)), (100.0, (0.3, 1, 0.7), (1, 1, 1)),
                        (2000.0, (1, 1, 1), (0.8, 0.8, 0.8))]:
        L = light([0, 0, -1], I, col)          # travels against the normal: dot(n, -l) = 1
        got = oracle.shade(n, alb, L)
        want = [min(math.floor(I * c * a / math.pi + 1e-9), 255) for c, a in zip(col, alb)]
        assert list(got) == want


def test_shade_lights_add_before_clamp_and_backlight_is_zero():
    n = [0.0, 1.0, 0.0]
    L = np.concatenate([light([0, -1, 0], 100.25 * math.pi, (1, 1, 1)),
                        light([0, -1, 0], 100.25 * math.pi, (1, 1, 1)),
                        light([0, 1, 0], 900 * math.pi, (1, 1, 1))])   # from behind: max(0, .) = 0
    # sums 200.5, 100.25, 401 -> truncated, clamped after the sum
    assert list(oracle.shade(n, [1, 0.5, 2], L)) == [200, 100, 255]


def test_texture_normalisation_matches_cimg_vectors():
    g = np.load(GOLDEN)
    for name in ("noise4", "noise3", "const", "unit"):
        got = oracle.normalized_texture(g[f"{name}_texels"])
        assert np.array_equal(got.view(np.uint32), g[f"{name}_normalized"].view(np.uint32)), name


def test_texture_sampler_matches_cimg_vectors():
    g = np.load(GOLDEN)
    for name in ("noise4", "noise3", "const", "unit"):
        tex, fx, fy, want = g[f"{name}_texels"], g[f"{name}_fx"], g[f"{name}_fy"], g[f"{name}_samples"]
        for c in range(3):
            got = np.array([oracle.linear_atXY(tex, x, y, c) for x, y in zip(fx, fy)], np.float32)
            assert np.array_equal(got.view(np.uint32), want[c].view(np.uint32)), (name, c)


def test_texture_sampler_matches_live_cimg_when_built():
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(5)
    for shape in [(31, 29, 4), (8, 64, 3), (1, 7, 4)]:
        tex = rng.integers(0, 256, shape, dtype=np.uint8)
        H, W, C = shape
        planar = np.ascontiguousarray(np.transpose(tex, (2, 0, 1)).astype(np.float32))
        norm = np.zeros_like(planar)
        ref.cimg_ref_normalize(planar.ctypes.data, W, H, C, norm.ctypes.data)
        assert np.array_equal(oracle.normalized_texture(tex), norm)
        fx = rng.uniform(-3, W + 3, 300).astype(np.float32)
        fy = rng.uniform(-3, H + 3, 300).astype(np.float32)
        out = np.zeros(300, np.float32)
        for c in range(3):
            ref.cimg_ref_linear_atXY(norm.ctypes.data, W, H, C, fx.ctypes.data, fy.ctypes.data, c, 300,
                                     out.ctypes.data)
            got = np.array([oracle.linear_atXY(tex, x, y, c) for x, y in zip(fx, fy)], np.float32)
            assert np.array_equal(got, out)


def test_sample_uv_convention():
    """u*W, (1-v)*H (material.cpp:19-20): uv (0,1) is texel (0,0); flat colour without texture."""
    tex = np.zeros((4, 8, 4), np.uint8)
    tex[0, 0] = [255, 0, 0, 0]
    tex[3, 7] = [0, 255, 0, 0]
    assert list(oracle.sample(tex, 0.0, 1.0)) == [1.0, 0.0, 0.0]
    assert list(oracle.sample(tex, 7 / 8, 1 / 4)) == [0.0, 1.0, 0.0]
    assert list(oracle.sample(None, 0.3, 0.3, diffuse=(0.25, 0.5, 0.75))) == [0.25, 0.5, 0.75]
    # bilinear weights on a linear ramp are exact: value at fx = x + 0.25 is 1/4 of the way
    ramp = np.zeros((2, 2, 3), np.uint8)
    ramp[:, 1, 0] = 255
    assert oracle.linear_atXY(ramp, 0.25, 0.0, 0) == 0.25


# ---------------------------------------------------------------------------
# whole frames on raster-space dyadic triangles, against exact coverage

def raster_scene(tris, zs=None, ws=None, wind=None):
    """Raster-space scene from a list of 3-vertex triangles [(x,y),...]."""
    n = len(tris)
    verts = np.zeros((3 * n, 4), np.float32)
    for i, t in enumerate(tris):
        for k in range(3):
            verts[3 * i + k, 0:2] = t[k]
            verts[3 * i + k, 2] = 0.5 if zs is None else zs[i][k]
            verts[3 * i + k, 3] = 1.0 if ws is None else ws[i][k]
    faces = make_faces(np.arange(3 * n).reshape(n, 3), n=np.zeros((n, 3), np.int32), t=None)
    return Scene(verts, np.array([[0, 0, 1]], np.float32), np.zeros((1, 2), np.float32), faces,
                 [MaterialSpec((1.0, 1.0, 1.0))],
                 make_lights([[0, 0, -1]], [255 * math.pi], [[1, 1, 1]]), raster_space=True)


def exact_cover(tri, W, H, wind_cw=False):
    """Pixels the reference covers for one raster triangle, in exact arithmetic
    (valid because every binary32 operation on these dyadic inputs is exact)."""
    v = [(Fr(x), Fr(y)) for x, y in tri]
    sh = sum(v[i][0] * v[(i + 1) % 3][1] - v[(i + 1) % 3][0] * v[i][1] for i in range(3))
    s = Fr(-1, 2) * sh
    if not ((s > 0) != wind_cw):
        return set()
    xs, ys = [p[0] for p in v], [p[1] for p in v]
    x0 = max(0, min(W - 1, math.floor(min(xs)))); x1 = max(0, min(W - 1, math.ceil(max(xs))))
    y0 = max(0, min(H - 1, math.floor(min(ys)))); y1 = max(0, min(H - 1, math.ceil(max(ys))))
    area = exact_edge(v[2], v[0], v[1])
    out = set()
    for y in range(y0, y1 + 1):
        for x in range(x0, x1 + 1):
            e = [exact_edge((x, y), v[1], v[2]), exact_edge((x, y), v[2], v[0]), exact_edge((x, y), v[0], v[1])]
            if area != 0:
                ok = all(ei == 0 or (ei > 0) == (area > 0) for ei in e)
            else:   # e/+0: +inf passes, -inf and NaN (0/0) fail
                ok = all(ei > 0 for ei in e)
            if ok:
                out.add((x, y))
    return out


def random_dyadic_triangles(rng, n, W, H):
    tris = []
    for _ in range(n):
        c = rng.integers(0, 4 * W, 2) / 4
        t = [(float(c[0] + d[0]), float(c[1] + d[1])) for d in (rng.integers(-24, 25, (3, 2)) / 4)]
        tris.append(t)
    return tris


@pytest.mark.parametrize("seed,wind", [(10, False), (11, False), (12, True)])
def test_coverage_matches_exact_rationals(seed, wind):
    rng = np.random.default_rng(seed)
    W, H = 24, 18
    tris = random_dyadic_triangles(rng, 14, W, H)
    # add exact edge-sharing pairs and an off-screen and a degenerate face
    tris += [[(2.0, 2.0), (2.0, 10.0), (10.0, 2.0)], [(10.0, 2.0), (2.0, 10.0), (10.0, 10.0)]]
    tris += [[(-5.0, -5.0), (-1.0, -5.0), (-5.0, -1.0)], [(3.0, 3.0), (6.0, 6.0), (9.0, 9.0)]]
    scene = raster_scene(tris)
    args = make_args(W, H, wind_clockwise=wind)
    frame, depth, ids, counts, _ = oracle.draw_frame(scene, args)
    covers = [exact_cover(t, W, H, wind) for t in tris]
    assert counts["fragments"] == sum(len(c) for c in covers)
    covered = set().union(*covers)
    got = {(int(x), int(y)) for y, x in zip(*np.nonzero(ids >= 0))}
    assert got == covered
    # every face lies at z = 0.5: the first face (lowest index) covering a pixel keeps it
    for (x, y) in covered:
        assert ids[y, x] == min(i for i, c in enumerate(covers) if (x, y) in c)


def test_shared_edge_is_covered_by_both_faces():
    """No fill convention: bary >= 0 covers pixels on a shared edge twice (drawing.cpp:111)."""
    W = H = 12
    a = [(1.0, 1.0), (1.0, 9.0), (9.0, 1.0)]      # anticlockwise on screen: front facing
    b = [(9.0, 1.0), (1.0, 9.0), (9.0, 9.0)]
    scene = raster_scene([a, b])
    _, _, ids, counts, _ = oracle.draw_frame(scene, make_args(W, H))
    ca, cb = exact_cover(a, W, H), exact_cover(b, W, H)
    shared = ca & cb
    assert len(shared) == 9                       # the diagonal x + y = 10, x in 1..9
    assert counts["fragments"] == len(ca) + len(cb)
    assert all(ids[y, x] == 0 for x, y in shared)


def test_strict_depth_test_and_ties():
    W = H = 8
    t = [(0.0, 0.0), (0.0, 8.0), (8.0, 0.0)]
    near = raster_scene([t, t], zs=[(0.3,) * 3, (0.3,) * 3])            # equal depth: face 0 wins
    _, d, ids, c, _ = oracle.draw_frame(near, make_args(W, H))
    cov = ids >= 0
    assert (ids[cov] == 0).all() and c["depth_passes"] == cov.sum()
    behind = raster_scene([t, t], zs=[(0.625,) * 3, (0.25,) * 3])      # later face in front wins
    _, d, ids, _, _ = oracle.draw_frame(behind, make_args(W, H))         # (dyadic depths: exact sums)
    assert (ids[ids >= 0] == 1).all() and (d[ids >= 0] == np.float32(0.25)).all()
    at_one = raster_scene([t], zs=[(1.0,) * 3])                         # depth 1.0 never passes
    _, d, ids, c, _ = oracle.draw_frame(at_one, make_args(W, H))
    assert (ids < 0).all() and c["fragments"] > 0
    # -0 then +0 (the sign of each product follows the barycentric's sign): the second
    # face never replaces the first (+0 < -0 and -0 < +0 are both false), so the depth
    # bits are those of the first face drawn alone
    neg = raster_scene([t, t], zs=[(-0.0,) * 3, (0.0,) * 3])
    _, d, ids, _, _ = oracle.draw_frame(neg, make_args(W, H))
    _, d1, ids1, _, _ = oracle.draw_frame(raster_scene([t], zs=[(-0.0,) * 3]), make_args(W, H))
    cov = ids >= 0
    assert (ids[cov] == 0).all() and np.array_equal(d.view(np.uint32), d1.view(np.uint32))
    assert (d[cov].view(np.uint32) == np.uint32(0x80000000)).any()


def test_buffers_are_not_cleared():
    W = H = 10
    t = [(0.0, 0.0), (0.0, 10.0), (10.0, 0.0)]
    scene = raster_scene([t], zs=[(0.5,) * 3])
    frame0 = np.full((3, H, W), 17, np.uint8)
    depth0 = np.full((H, W), 0.75, np.float32)
    depth0[:, :3] = 0.25                              # nearer than the face: keeps old colour
    f, d, ids, _, _ = oracle.draw_frame(scene, make_args(W, H), frame=frame0, depth=depth0)
    cov = exact_cover(t, W, H)
    for y in range(H):
        for x in range(W):
            if (x, y) in cov and x >= 3:
                assert d[y, x] == np.float32(0.5) and ids[y, x] == 0
            else:
                assert d[y, x] == depth0[y, x] and (f[:, y, x] == 17).all() and ids[y, x] == -1


def test_perspective_correct_interpolation():
    """uv follows (sum w_i b_i uv_i) / (sum w_i b_i), not screen-linear interpolation: a
    red ramp texture (bilinear on a linear ramp is exact) turns the interpolated u into
    the red channel of a one-light Lambert frame."""
    W, H = 64, 8
    verts = np.array([[0, 0, 0.5, 1.0], [64, 0, 0.5, 0.125], [0, 8, 0.5, 1.0], [64, 8, 0.5, 0.125]], np.float32)
    tex = np.zeros((2, 256, 3), np.uint8)
    tex[:, :, 0] = np.arange(256, dtype=np.uint8)[None, :]
    uvs = np.array([[0, 1], [1, 1], [0, 0], [1, 0]], np.float32)
    faces = make_faces([[0, 2, 1], [1, 2, 3]], n=[[0, 0, 0], [0, 0, 0]], t=[[0, 2, 1], [1, 2, 3]])
    scene = Scene(verts, np.array([[0, 0, 1]], np.float32), uvs, faces, [MaterialSpec((1, 1, 1), tex)],
                  make_lights([[0, 0, -1]], [254.999 * math.pi], [[1, 1, 1]]), raster_space=True)
    f, d, ids, _, _ = oracle.draw_frame(scene, make_args(W, H))
    for x in range(1, 63):
        b1 = x / 64.0                   # screen weight of the right-hand edge
        u = (0.125 * b1) / ((1 - b1) * 1.0 + 0.125 * b1)
        fx = min(max(u * 256, 0), 255)
        want = 254.999 * (math.floor(fx) + (fx - math.floor(fx))) / 255.0
        assert abs(int(f[0, 4, x]) - want) <= 1.0, (x, f[0, 4, x], want)
        assert abs(int(f[0, 4, x]) - 254.999 * min(b1 * 256, 255) / 255) > 1.5 or x < 3


def test_flat_equals_per_face_normals():
    """--flat shades with normalize(cross(p1-p0, p2-p0)); identical to the reference run
    on the same mesh with one vn per face (the flag is a no-op in the reference)."""
    scene, args = scenes.uv_sphere_scene(width=160, height=120)
    args.flat = 1
    f1, d1, i1, _, _ = oracle.draw_frame(scene, args)
    p = scene.vertices.astype(np.float32)
    t = scene.faces[:, 0:3]
    e1, e2 = p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]]
    cr = np.stack([e1[:, 1] * e2[:, 2] - e2[:, 1] * e1[:, 2], e1[:, 2] * e2[:, 0] - e2[:, 2] * e1[:, 0],
                   e1[:, 0] * e2[:, 1] - e2[:, 0] * e1[:, 1]], 1).astype(np.float32)
    ln = (cr[:, 0] * cr[:, 0] + cr[:, 1] * cr[:, 1]) + cr[:, 2] * cr[:, 2]
    nrm = (cr * (np.float32(1) / np.sqrt(ln.astype(np.float32)))[:, None]).astype(np.float32)
    faces = scene.faces.copy()
    faces[:, 3:6] = np.arange(len(faces))[:, None]
    per_face = Scene(scene.vertices, nrm, scene.uvs, faces, scene.materials, scene.lights)
    args.flat = 0
    f2, d2, i2, _, _ = oracle.draw_frame(per_face, args)
    assert np.array_equal(f1, f2) and np.array_equal(i1, i2)


def test_square_scene_closed_form():
    """The reference's default scene (renderer.cpp:32-50) at its default 540x304 view:
    the square's projected corners are known in closed form."""
    scene = scenes.square_scene()
    scene.lights = make_lights([[0, 0, -1]], [255 * math.pi], [[1, 1, 1]])
    W, H = 540, 304
    f, d, ids, counts, _ = oracle.draw_frame(scene, make_args(W, H))
    t = math.tan(math.radians(45) / 2)
    sx = 0.5 / (3 * t * (W / H)) * W / 2            # half-width of the square on screen
    sy = 0.5 / (3 * t) * H / 2
    cx, cy = W / 2, H / 2
    inside = [(x, y) for x in range(W) for y in range(H)
              if abs(x - cx) < sx - 1 and abs(y - cy) < sy - 1]
    outside = [(x, y) for x in range(0, W, 7) for y in range(0, H, 5)
               if abs(x - cx) > sx + 1 or abs(y - cy) > sy + 1]
    assert all(ids[y, x] >= 0 for x, y in inside) and all(ids[y, x] < 0 for x, y in outside)
    # normal (0,0,1) lit head-on by intensity 255*pi: every covered pixel is 255 * 1 * 1
    assert (f[:, ids >= 0] == 255).all()
    zc = (6.1 / 5.9) - 2 * 6 * 0.1 / (5.9 * 3)
    assert abs(float(d[int(cy), int(cx)]) - zc) < 1e-6


def test_invalid_indices_are_rejected():
    scene, args = scenes.uv_sphere_scene(width=32, height=32)
    bad = Scene(scene.vertices, scene.normals, scene.uvs, scene.faces.copy(), scene.materials, scene.lights)
    bad.faces[5, 3] = len(scene.normals)
    assert oracle.draw_frame_status(bad, args) == 2
    bad.faces[5, 3] = 0
    bad.faces[7, 9] = 3
    assert oracle.draw_frame_status(bad, args) == 2


def test_flat_normal_map_matches_smooth_shading():
    """A normal map of +z texels (128,128,255) bends the normal by under 0.4 degrees:
    every covered pixel stays within 1/255 of the unmapped frame (the definition's
    invariant), and ids and depth do not move."""
    scene, args = scenes.icosphere_scene(subdivisions=4, width=320, height=240, texture_size=64)
    f0, d0, i0, _, _ = oracle.draw_frame(scene, args)
    scene.materials[0].normal_map = np.tile(np.array([128, 128, 255], np.uint8), (32, 32, 1))
    f1, d1, i1, _, _ = oracle.draw_frame(scene, args)
    assert np.array_equal(i0, i1) and np.array_equal(d0, d1)
    cov = i0 >= 0
    diff = np.abs(f1.astype(int) - f0.astype(int))[:, cov]
    assert diff.max() <= 1 and cov.sum() > 10000


def test_normal_map_tilts_toward_the_tangent():
    """A map of normals tilted toward +x (tangent) brightens faces whose tangent points
    at a light and darkens the rest, compared with +z; both are Lambert closed forms
    through the same definition."""
    scene, args = scenes.icosphere_scene(subdivisions=4, width=320, height=240, texture_size=64)
    tilt = np.tile(np.array([255, 128, 128], np.uint8), (16, 16, 1))       # t = (1, ~0, ~0)
    flat = np.tile(np.array([128, 128, 255], np.uint8), (16, 16, 1))
    scene.materials[0].normal_map = flat
    fa, _, ia, _, _ = oracle.draw_frame(scene, args)
    scene.materials[0].normal_map = tilt
    fb, _, ib, _, _ = oracle.draw_frame(scene, args)
    assert np.array_equal(ia, ib)
    assert np.abs(fa.astype(int) - fb.astype(int))[:, ia >= 0].mean() > 5


def test_interpolated_normal_is_renormalised():
    """drawing.cpp:131 normalises the interpolated normal: vertex normals of lengths 3,
    0.5 and 7 along +z shade the square exactly like unit ones (Lambert with one light
    along the normal is I*c*a/pi, a constant); without the normalisation the colour would
    follow the interpolated length across each face."""
    scene = scenes.square_scene()
    scene.normals = np.array([[0, 0, 3.0], [0, 0, 0.5], [0, 0, 7.0]], np.float32)
    scene.faces = make_faces([[0, 1, 2], [2, 1, 3]], n=[[0, 1, 2], [2, 1, 0]], t=[[0, 1, 2], [2, 1, 3]])
    I = 60.5 * math.pi
    scene.lights = make_lights([[0, 0, -1]], [I], [[1, 1, 1]])
    f, d, ids, _, _ = oracle.draw_frame(scene, make_args(96, 64))
    cov = ids >= 0
    assert cov.sum() > 500
    assert (f[:, cov] == 60).all()          # trunc(60.5 * dot), dot within an ulp or two of 1


def test_light_direction_is_renormalised():
    """geometry.cpp:124-133: trans_dir = normalize(view * (direction, 0)).  Directions of
    length 4 and 15 come back as unit vectors (the view matrix is a translation, so the
    direction is only rescaled), and shading uses the unit vector: a light at 2/3 of the
    way to head-on gives trunc(I*c*a/pi * 2/3)."""
    scene = scenes.square_scene()
    I = 150.75 * math.pi
    for direction, cos in (([0, 0, -4.0], 1.0), ([5.0, 10.0, -10.0], 2.0 / 3.0)):
        scene.lights = make_lights([direction], [I], [[1, 1, 1]])
        f, d, ids, _, lights = oracle.draw_frame(scene, make_args(96, 64))
        t = lights[0, 7:10].astype(np.float64)
        want = np.array(direction) / np.linalg.norm(direction)
        assert np.abs(t - want).max() <= 2 * np.spacing(np.float32(1)), (t, want)
        cov = ids >= 0
        assert cov.sum() > 500 and (f[:, cov] == math.floor(150.75 * cos)).all(), direction
