"""GPU parity of the frame path: librz (through its C ABI) against the CPU oracle on
the same seeded inputs.  Face ids, depth bits and colour bytes must match exactly."""
import numpy as np
import pytest

from rasteriser_amd import scenes
from tests.parity import check

pytestmark = pytest.mark.gpu


def test_uv_sphere_512(ctx):
    scene, args = scenes.uv_sphere_scene()
    gpu, ref = check(ctx, scene, args, label="uv sphere 512")
    assert (gpu[2] >= 0).sum() > 50000


@pytest.mark.parametrize("flat", [False, True])
def test_icosphere_1080p(ctx, flat):
    scene, args = scenes.icosphere_scene(flat=flat)
    gpu, ref = check(ctx, scene, args, label=f"icosphere 1080p flat={flat}")
    assert (gpu[2] >= 0).sum() > 300000


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_icosphere_1080p_more_rotations(ctx, seed):
    """The 1080p workload at further seeded rotations, lights and textures (smooth and,
    for one seed, flat): every face id, depth bit and colour byte as the oracle's."""
    scene, args = scenes.icosphere_scene(seed=seed, flat=(seed == 4))
    gpu, ref = check(ctx, scene, args, label=f"icosphere 1080p seed {seed}")
    assert (gpu[2] >= 0).sum() > 300000


@pytest.mark.parametrize("tile,cap", [(8, 64), (16, 512), (32, 2048), (16, 1), (16, 4)])
def test_tuning_and_overflow(ctx, tile, cap):
    """Every tile size, and tiny bin capacities that force the overflow path."""
    scene, args = scenes.icosphere_scene(flat=False, subdivisions=4, width=640, height=360)
    gpu, ref = check(ctx, scene, args, tuning=(tile, cap), label=f"tile {tile} cap {cap}")
    if cap <= 4:
        assert gpu[3]["overflow_tiles"] > 0


@pytest.mark.parametrize("tile,cap", [(8, 256), (16, 512), (32, 2048)])
def test_large_triangles_and_slivers(ctx, tile, cap):
    """Triangles much larger than a tile, slivers within 1e-3 px of a line and faces with a
    vertex thousands of pixels off screen: the raster pass shrinks large candidate
    rectangles to the triangle's reach, and the fragment count (every (face, pixel)
    passing the inside test) must still equal the oracle's."""
    scene, args = scenes.large_triangle_scene(n=600)
    gpu, ref = check(ctx, scene, args, tuning=(tile, cap), label=f"large triangles tile {tile}")
    assert ref[3]["fragments"] > 5_000_000


@pytest.mark.parametrize("W,H", [(300, 200), (2048, 2048)])
def test_not_cleared_buffers(ctx, W, H):
    """draw_frame does not clear: the depth test reads the existing depth and pixels it
    does not win keep their old colour (drawing.cpp:119, renderer.cpp:107-108).  2048x2048
    has 16,384 tiles: 64-thread workgroups that compact the covered pixels."""
    scene, args = scenes.uv_sphere_scene(width=W, height=H)
    rng = np.random.default_rng(7)
    frame = rng.integers(0, 256, (3, H, W), dtype=np.uint8)
    depth = rng.uniform(0.85, 1.0, (H, W)).astype(np.float32)
    depth[::7, ::5] = np.float32(2.0)
    check(ctx, scene, args, frame=frame, depth=depth, label="not cleared")


@pytest.mark.parametrize("case", ["all_back", "off_screen", "one_pixel", "huge_coords"])
def test_degenerate_frames(ctx, case):
    """Frames with nothing or almost nothing to draw: (almost) every face turned to face
    away, every face off screen, a 1x1 image, and a third of the vertices at raster
    coordinates of 10^9 pixels."""
    from rasteriser_amd.scene import Scene
    scene, args = scenes.large_triangle_scene(n=300, size=256)
    v = scene.vertices.copy()
    faces = scene.faces.copy()
    if case == "all_back":
        x, y = v[faces[:, 0:3], 0].astype(np.float64), v[faces[:, 0:3], 1].astype(np.float64)
        s2 = (x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0] + x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1]
              + x[:, 2] * y[:, 0] - x[:, 0] * y[:, 2])
        flip = s2 < 0   # front under the raster y flip: wind them the other way
        faces[flip, 1], faces[flip, 2] = faces[flip, 2].copy(), faces[flip, 1].copy()
    elif case == "off_screen":
        v[:, 0] += 100_000.0
    elif case == "one_pixel":
        args = scenes.make_args(1, 1, fovy=60.0, z_near=0.1, z_far=100.0)
        v[:, 0:2] *= 1.0 / 256.0
    elif case == "huge_coords":
        v[::3, 0:2] *= 1e7
    scene = Scene(v, scene.normals, scene.uvs, faces, scene.materials, scene.lights, raster_space=True)
    gpu, ref = check(ctx, scene, args, label=f"degenerate {case}")
    if case == "off_screen":
        assert (gpu[2] < 0).all()
    if case == "all_back":
        assert gpu[3]["front_triangles"] < 10


def test_wind_clockwise(ctx):
    scene, args = scenes.uv_sphere_scene(width=256, height=256)
    args.wind_clockwise = 1
    check(ctx, scene, args, label="wind clockwise")


def test_screen_soup_small(ctx):
    scene, args = scenes.screen_soup_scene(n=200_000, size=1024)
    gpu, ref = check(ctx, scene, args, label="screen soup 200k")
    assert gpu[3]["front_triangles"] > 80_000


def _grid_scene(cells, step, origin, jitter, seed):
    """Screen-space grid of cells x cells quads of `step` pixels at `origin`, two
    triangles each (half wound backwards), vertices moved by up to `jitter` pixels,
    depths random: with no jitter every shared edge passes through pixel centres, whose
    edge values are exactly zero."""
    import math
    from rasteriser_amd.abi import make_args
    from rasteriser_amd.scene import MaterialSpec, Scene, make_faces, make_lights
    rng = np.random.default_rng(seed)
    g = np.arange(cells + 1, dtype=np.float64) * step + origin
    xs, ys = np.meshgrid(g, g, indexing="xy")
    xs = xs + rng.uniform(-jitter, jitter, xs.shape)
    ys = ys + rng.uniform(-jitter, jitter, ys.shape)
    n = (cells + 1) ** 2
    d = rng.uniform(0.5, 50.0, n)
    z = 1.0 - 2.0 / d
    verts = np.stack([xs.ravel(), ys.ravel(), z, 1.0 / d], 1).astype(np.float32)
    i = np.arange(cells)
    a = (i[:, None] * (cells + 1) + i[None, :]).ravel()
    tri = np.concatenate([np.stack([a, a + cells + 1, a + 1], 1), np.stack([a + 1, a + cells + 1, a + cells + 2], 1)])
    back = rng.uniform(size=len(tri)) < 0.5
    tri[back, 1], tri[back, 2] = tri[back, 2].copy(), tri[back, 1].copy()
    faces = make_faces(tri.astype(np.int64), n=np.zeros_like(tri, dtype=np.int64), t=None)
    size = int(math.ceil(origin + cells * step)) + 2
    scene = Scene(verts, np.array([[0.0, 0.0, 1.0]], np.float32), np.zeros((1, 2), np.float32), faces,
                  [MaterialSpec((1.0, 1.0, 1.0))],
                  make_lights([[0.0, 0.0, -1.0]], [255.0 * math.pi], [[1.0, 1.0, 1.0]]), raster_space=True)
    return scene, make_args(size, size, fovy=60.0, z_near=0.1, z_far=100.0)


@pytest.mark.parametrize("cells,step,origin,jitter", [(64, 8.0, 0.0, 0.0), (96, 5.0, 0.5, 0.0),
                                                      (16, 1.0, 0.0, 2.0 ** -19), (24, 0.75, 1.0, 2.0 ** -21)])
def test_edges_through_pixel_centres(ctx, cells, step, origin, jitter):
    """Candidates whose edge values are zero or near 2^-20 — the edge of the range where
    the raster pass takes its quotients from the face's reciprocal — take the general
    test; the frame must still match the oracle bit for bit (ties between faces sharing
    an edge included)."""
    scene, args = _grid_scene(cells, step, origin, jitter, seed=cells)
    for tuning in ((16, 0), (8, 64)):
        check(ctx, scene, args, tuning=tuning, label=f"grid {cells}x{step} jitter {jitter} tiles {tuning}")


def test_shared_reciprocal_quotients(ctx):
    """The raster pass's quotient from a per-face reciprocal equals the IEEE division
    bit for bit on 2^30 random operand pairs in its range (and is not used outside it)."""
    for seed in (1, 2):
        compared, bad = ctx.check_quotients(1 << 30, seed=seed)
        assert bad == 0, f"{bad} of {compared} quotients differ"
        assert compared > 0.9 * (1 << 30)


def test_screen_soup_full_4m(ctx):
    """The small-triangle depth stress at its full size: 4M screen-space triangles at
    4096x4096, about half back-facing, bit for bit against the oracle — once with the
    default bins and once with bins of 40 so that crowded tiles take the overflow path."""
    import oracle
    from tests.parity import assert_same_frame, render_gpu
    scene, args = scenes.screen_soup_scene()
    assert scene.num_faces == 4_000_000 and args.image_width == 4096
    ref = oracle.draw_frame(scene, args)
    gpu = render_gpu(ctx, scene, args)
    assert_same_frame(gpu, ref, "screen soup 4M")
    assert gpu[3]["overflow_tiles"] == 0
    assert 1_900_000 < gpu[3]["front_triangles"] < 2_100_000
    gpu = render_gpu(ctx, scene, args, tuning=(16, 40))
    assert_same_frame(gpu, ref, "screen soup 4M, bins of 40")
    assert gpu[3]["overflow_tiles"] > 1000


@pytest.mark.parametrize("tuning", [(16, 512), (16, 4), (8, 64), (32, 2048)])
def test_split_shade_pass(ctx, tuning):
    """The deferred shade as its own pass over the HBM visibility buffer (the default
    shades non-overflowed tiles inside the raster pass): same bits."""
    ctx.set_fused_shade(False)
    try:
        scene, args = scenes.icosphere_scene(subdivisions=5, width=800, height=600)
        check(ctx, scene, args, tuning=tuning, label=f"split shade {tuning}")
    finally:
        ctx.set_fused_shade(True)
    assert ctx.fused_shade


def test_repeatable(ctx):
    """Two renders of the same frame give identical visibility (order-free atomics)."""
    scene, args = scenes.icosphere_scene(subdivisions=5, width=800, height=600)
    from tests.parity import render_gpu
    a = render_gpu(ctx, scene, args)
    b = render_gpu(ctx, scene, args)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_frame_graph_spin(ctx):
    """A captured 3-frame spin graph leaves the last frame, rotated by 2*ry_step."""
    import oracle
    from tests.parity import assert_same_frame
    scene, args = scenes.icosphere_scene(subdivisions=4, width=640, height=480)
    W, H = 640, 480
    mesh = ctx.upload(scene)
    frame = ctx.empty((3, H, W), np.uint8)
    depth = ctx.empty((H, W), np.float32)
    ids = ctx.empty((H, W), np.int32)
    lights = scene.lights_copy()
    step = 0.25
    g = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, 3, ids_ptr=ids.ptr, ry_step=step)
    for _ in range(2):          # replays must give the same result each time
        frame.upload(np.full((3, H, W), 7, np.uint8))
        g.launch()
        ctx.sync()
        last = np.float32(args.tait_bryan_angles[1])
        a2 = type(args).from_buffer_copy(args)
        a2.tait_bryan_angles[1] = float(np.float32(np.float32(last + np.float32(step)) + np.float32(step)))
        ref = oracle.draw_frame(scene, a2)
        gpu = (frame.download(), depth.download(), ids.download(),
               {"fragments": ref[3]["fragments"], "front_triangles": ref[3]["front_triangles"],
                "covered_pixels": int((ref[2] >= 0).sum())}, lights)
        assert_same_frame(gpu, ref, "graph spin")
    g.close()
    mesh.close()


def test_graph_stale_after_realloc():
    """A frame that grows the context buffers makes an older graph refuse to launch (a
    fresh context, whose buffers are sized by this test's frames alone)."""
    import rasteriser_amd as rz
    with rz.Context(0) as ctx:
        scene, args = scenes.uv_sphere_scene(width=64, height=64)
        mesh = ctx.upload(scene)
        frame = ctx.empty((3, 64, 64), np.uint8)
        depth = ctx.empty((64, 64), np.float32)
        g = ctx.graph(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, 1)
        g.launch()
        ctx.sync()
        big, bargs = scenes.uv_sphere_scene(width=2500, height=2500)
        ctx.draw(mesh, scene.lights_copy(), bargs)
        with pytest.raises(rz.RzError):
            g.launch()


def test_context_close_releases_children():
    """Closing a context first releases its meshes, graphs and device arrays; objects
    collected afterwards do not touch the destroyed context."""
    import gc
    import rasteriser_amd as rz
    ctx = rz.Context(0)
    scene, args = scenes.uv_sphere_scene(width=64, height=64)
    mesh = ctx.upload(scene)
    frame = ctx.empty((3, 64, 64), np.uint8)
    depth = ctx.empty((64, 64), np.float32)
    g = ctx.graph(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, 2)
    g.launch()
    ctx.close()
    assert mesh.handle is None and g.handle is None and frame.ptr == 0 and depth.ptr == 0
    del mesh, frame, depth, g
    gc.collect()


def _device_outputs(ctx, args, with_ids=True):
    W, H = args.image_width, args.image_height
    return (ctx.empty((3, H, W), np.uint8), ctx.empty((H, W), np.float32),
            ctx.empty((H, W), np.int32) if with_ids else None)


def test_batch_matches_single_frames_and_oracle(ctx):
    """Every mesh of one batched launch equals its own single-mesh frame bit for bit,
    and a few equal the oracle."""
    import oracle
    from tests.parity import assert_same_frame
    work = scenes.batch_scenes(n=12, seed=21, max_faces=20_000)
    meshes = [ctx.upload(s) for s, a in work]
    outs = [_device_outputs(ctx, a) for s, a in work]
    lights = [s.lights_copy() for s, a in work]
    batch = [(m, l, a, f, d, i) for m, l, (s, a), (f, d, i) in zip(meshes, lights, work, outs)]
    st = ctx.draw_batch(batch, stats=True)
    total = {"front_triangles": 0, "fragments": 0, "covered_pixels": 0}
    for k, ((scene, args), (f, d, i), m) in enumerate(zip(work, outs, meshes)):
        fb, db, ib = f.download(), d.download(), i.download()
        single = ctx.draw(m, scene.lights_copy(), args, ids=True, stats=True)
        assert np.array_equal(fb, single[0]) and np.array_equal(ib, single[2]), k
        assert np.array_equal(db.view(np.uint32), single[1].view(np.uint32)), k
        for key in total:
            total[key] += single[3][key]
        if k < 4:
            ref = oracle.draw_frame(scene, args)
            assert_same_frame((fb, db, ib, single[3], lights[k]), ref, f"batch item {k}")
    for key in total:   # the batch's statistics are the sums of its meshes'
        assert st[key] == total[key], key
    for m in meshes:
        m.close()


def test_batch_graph_and_autotune(ctx):
    """The autotuner's choice renders the same pixels; a batch graph replays them."""
    work = scenes.batch_scenes(n=6, seed=22, max_faces=30_000)
    meshes = [ctx.upload(s) for s, a in work]
    outs = [_device_outputs(ctx, a) for s, a in work]
    batch = [(m, s.lights_copy(), a, f, d, i) for m, (s, a), (f, d, i) in zip(meshes, work, outs)]
    ctx.draw_batch(batch)
    before = [(f.download(), d.download(), i.download()) for f, d, i in outs]
    ts, cap, ms = ctx.autotune(batch)
    assert ts in (8, 16, 32) and cap >= 16 and ms > 0
    assert ctx.tuning() == (ts, cap)
    g = ctx.batch_graph(batch, frames=2)
    for _ in range(2):
        for f, d, i in outs:
            f.upload(np.full(f.shape, 3, np.uint8))
        g.launch()
        ctx.sync()
        for (f, d, i), (fb, db, ib) in zip(outs, before):
            assert np.array_equal(f.download(), fb) and np.array_equal(i.download(), ib)
            assert np.array_equal(d.download().view(np.uint32), db.view(np.uint32))
    g.close()
    ctx.set_tuning(16, 0)
    for m in meshes:
        m.close()


@pytest.mark.parametrize("tile,cap", [(16, 2), (8, 1)])
def test_batch_overflow_path(ctx, tile, cap):
    work = scenes.batch_scenes(n=5, seed=23, max_faces=8_000)
    ctx.set_tuning(tile, cap)
    try:
        meshes = [ctx.upload(s) for s, a in work]
        outs = [_device_outputs(ctx, a) for s, a in work]
        batch = [(m, s.lights_copy(), a, f, d, i) for m, (s, a), (f, d, i) in zip(meshes, work, outs)]
        st = ctx.draw_batch(batch, stats=True)
        assert st["overflow_tiles"] > 0
        import oracle
        for (scene, args), (f, d, i) in list(zip(work, outs))[:3]:
            of, od, oi, _, _ = oracle.draw_frame(scene, args)
            assert np.array_equal(i.download(), oi) and np.array_equal(f.download(), of)
            assert np.array_equal(d.download().view(np.uint32), od.view(np.uint32))
    finally:
        ctx.set_tuning(16, 0)


def test_normal_map_and_64_lights(ctx):
    """The learned-shading workload's plain shader (64 lights, cone normal map) at a
    reduced size, bit for bit against the oracle."""
    scene, args = scenes.learned_shading_scene(width=640, height=360, subdivisions=5, map_size=256)
    gpu, ref = check(ctx, scene, args, label="normal map 64 lights")
    assert (gpu[2] >= 0).sum() > 20000


def test_counter_generations_across_calls():
    """Single frames keep two generations of tile counters: each frame's bin pass zeroes
    the one the next frame uses.  Eager frames, frame graphs with odd and even frame
    counts, a change of image size, the overflow path and pass timing interleaved on one
    context must each still match the oracle."""
    import oracle
    import rasteriser_amd as rz
    from tests.parity import assert_same_frame
    big, small = (640, 480), (320, 200)
    with rz.Context(0) as c:
        assert c.tuning() == (16, 0)   # default: automatic bin capacity
        scenes_by_size = {wh: scenes.icosphere_scene(subdivisions=4, width=wh[0], height=wh[1]) for wh in (big, small)}
        meshes = {wh: c.upload(sa[0]) for wh, sa in scenes_by_size.items()}
        bufs = {wh: (c.empty((3, wh[1], wh[0]), np.uint8), c.empty((wh[1], wh[0]), np.float32),
                     c.empty((wh[1], wh[0]), np.int32)) for wh in (big, small)}
        refs = {wh: oracle.draw_frame(*scenes_by_size[wh]) for wh in (big, small)}

        def compare(wh, label, stats=None):
            f, d, i = bufs[wh]
            ref = refs[wh]
            st = stats or {"fragments": ref[3]["fragments"], "front_triangles": ref[3]["front_triangles"],
                           "covered_pixels": int((ref[2] >= 0).sum())}
            assert_same_frame((f.download(), d.download(), i.download(), st, ref[4]), ref, label)

        def eager(wh, label):
            scene, args = scenes_by_size[wh]
            f, d, i = bufs[wh]
            print("eager", label, flush=True)
            st = c.draw_device(meshes[wh], scene.lights_copy(), args, f.ptr, d.ptr, i.ptr,
                               rz.RZ_FLAG_CLEARED, stats=True)
            compare(wh, label, st)

        def graph(wh, frames, label, launches=1):
            scene, args = scenes_by_size[wh]
            f, d, i = bufs[wh]
            g = c.graph(meshes[wh], scene.lights_copy(), args, f.ptr, d.ptr, frames, ids_ptr=i.ptr)
            try:
                for k in range(launches):
                    print("graph", label, frames, "frames, launch", k, flush=True)
                    i.upload(np.full((wh[1], wh[0]), 5, np.int32))
                    g.launch()
                    c.sync()
                    compare(wh, f"{label} launch {k}")
                    eager(wh, f"{label} eager after launch {k}")
            finally:
                g.close()

        eager(big, "first")
        graph(big, 3, "odd graph", launches=2)
        eager(small, "smaller image")
        graph(big, 2, "even graph", launches=2)
        eager(big, "after even graph")
        c.set_tuning(16, 4)   # overflowing bins: the overflow count lives in the counter lines
        eager(big, "overflow 1")
        eager(small, "overflow smaller")
        graph(big, 3, "overflow odd graph", launches=2)
        c.set_tuning(16, 0)
        scene, args = scenes_by_size[big]
        f, d, i = bufs[big]
        for reps in (3, 4):
            print("time_passes reps", reps, flush=True)
            c.time_passes(meshes[big], scene.lights_copy(), args, f, d, reps=reps)
            eager(big, f"after time_passes reps {reps}")
        for m in meshes.values():
            m.close()
        for b in bufs.values():
            for x in b:
                x.close()


def test_many_lights(ctx):
    """The reference shades any number of lights (shading.cpp:24-29): 200 lights (64 in
    the kernel arguments, 136 from device memory) against the oracle, eager, in a frame
    graph (its own copy of the extra lights, kept across a later frame with other lights)
    and as batch items with 65 and 300 lights."""
    import oracle
    from tests.parity import assert_same_frame
    import math
    scene, args = scenes.icosphere_scene(subdivisions=4, width=640, height=360, texture_size=256, lights=200)
    # dim lights, so that no pixel saturates and the 136 beyond the kernel arguments show
    scene.lights = scenes.random_lights(scenes.SplitMix64(0x5EED7000), 200, scale=255 * math.pi / 60)
    gpu, ref = check(ctx, scene, args, label="200 lights")
    assert (gpu[0][:, gpu[2] >= 0] == 255).mean() < 0.01
    W, H = 640, 360
    mesh = ctx.upload(scene)
    f, d, i = _device_outputs(ctx, args)
    g = ctx.graph(mesh, scene.lights_copy(), args, f.ptr, d.ptr, 2, ids_ptr=i.ptr)
    other, _ = scenes.icosphere_scene(seed=9, subdivisions=4, width=640, height=360, texture_size=256,
                                      lights=150)
    ctx.draw(mesh, other.lights_copy(), args)   # rewrites the context's extra lights
    g.launch()
    ctx.sync()
    ref = oracle.draw_frame(scene, args)
    st = {"fragments": ref[3]["fragments"], "front_triangles": ref[3]["front_triangles"],
          "covered_pixels": int((ref[2] >= 0).sum())}
    assert_same_frame((f.download(), d.download(), i.download(), st, ref[4]), ref, "200-light graph")
    g.close()
    work = [scenes.icosphere_scene(seed=s, subdivisions=3, width=256, height=200, texture_size=64, lights=n)
            for s, n in ((11, 65), (12, 300), (13, 3))]
    meshes = [ctx.upload(s) for s, a in work]
    outs = [_device_outputs(ctx, a) for s, a in work]
    batch = [(m, s.lights_copy(), a, bf, bd, bi) for m, (s, a), (bf, bd, bi) in zip(meshes, work, outs)]
    ctx.draw_batch(batch)
    for (s, a), (bf, bd, bi), item in zip(work, outs, batch):
        ref = oracle.draw_frame(s, a)
        st = {"fragments": ref[3]["fragments"], "front_triangles": ref[3]["front_triangles"],
              "covered_pixels": int((ref[2] >= 0).sum())}
        assert_same_frame((bf.download(), bd.download(), bi.download(), st, item[1]), ref,
                          f"batch item with {s.lights.shape[0]} lights")
    for m in meshes + [mesh]:
        m.close()


@pytest.mark.parametrize("W,H", [(40000, 48), (64, 65535)])
def test_image_sides_above_32767(ctx, W, H):
    """Image sides up to 65,535 pixels (bbox corners packed in 16 bits each): a 40,000-pixel
    wide and a 65,535-pixel tall frame against the oracle."""
    scene, args = scenes.uv_sphere_scene(width=W, height=H)
    gpu, ref = check(ctx, scene, args, label=f"{W}x{H}")
    assert (gpu[2] >= 0).sum() > 500


def test_spin_loop_replays(ctx):
    """rz_spin: one frame captured once, replayed with the rotation the spin loop reaches
    (renderer.cpp:117-124) — each replay equals the oracle at its angles, with eager
    frames of another size on the same context in between (they move the tile-counter
    generations the replays rely on)."""
    import oracle
    from tests.parity import assert_same_frame
    scene, args = scenes.icosphere_scene(subdivisions=4, width=640, height=480, texture_size=128)
    other, oargs = scenes.uv_sphere_scene(width=320, height=200)
    mesh, omesh = ctx.upload(scene), ctx.upload(other)
    f, d, i = _device_outputs(ctx, args)
    spin = ctx.spin(mesh, scene.lights_copy(), args, f.ptr, d.ptr, i.ptr)
    a0 = [float(x) for x in args.tait_bryan_angles]
    for k, dy in enumerate([0.0, 0.37, 1.25, 2.5, -0.6]):
        angles = [a0[0], float(np.float32(np.float32(a0[1]) + np.float32(dy))), a0[2]]
        print("spin frame", k, angles, flush=True)
        i.upload(np.full(i.shape, 3, np.int32))
        spin.frame(angles)
        ctx.sync()
        a = type(args).from_buffer_copy(args)
        a.tait_bryan_angles[:] = angles
        ref = oracle.draw_frame(scene, a)
        st = {"fragments": ref[3]["fragments"], "front_triangles": ref[3]["front_triangles"],
              "covered_pixels": int((ref[2] >= 0).sum())}
        assert_same_frame((f.download(), d.download(), i.download(), st, ref[4]), ref, f"spin frame {k}")
        if k in (1, 2):
            ctx.draw(omesh, other.lights_copy(), oargs)
    spin.close()
    for m in (mesh, omesh):
        m.close()


def test_draw_text(ctx):
    """rz_draw_text: 12x16 cells of white glyph pixels on black at (x, y), nothing else
    touched, clipped at the frame's edges."""
    W, H = 120, 40
    f = ctx.empty((3, H, W), np.uint8)
    f.upload(np.full((3, H, W), 7, np.uint8))
    ctx.draw_text(f.ptr, W, H, 5, 5, "12.5")
    ctx.draw_text(f.ptr, W, H, W - 20, H - 8, "88")   # clipped on two sides
    ctx.sync()
    img = f.download()
    box = np.zeros((H, W), bool)
    box[5:21, 5:53] = True
    box[H - 8:, W - 20:] = True
    assert (img[:, ~box] == 7).all()
    cells = img[:, box]
    assert ((cells == 255).all(axis=0) | (cells == 0).all(axis=0)).all()
    lit = (img[0, 5:21, 5:53] == 255)
    assert 40 < lit.sum() < 48 * 16 / 2
    # the four cells differ from each other (distinct glyphs) and the '.' is the smallest
    per = [lit[:, 12 * c:12 * c + 12].sum() for c in range(4)]
    assert per[2] == min(per) and len(set(per)) >= 3
    f.close()


def test_batch_256_meshes_full_size(ctx):
    """The batched workload at its stated size: 256 meshes (log-uniform 500-200k faces,
    8.5M in all, viewports 256^2-1024^2) autotuned and drawn in one batch graph.  Every
    mesh equals its own single-mesh frame at the tuned tile size and capacity, bit for bit;
    16 seeded meshes, and every mesh whose bins overflow at that capacity, equal the
    oracle (drawing.cpp:205-258 per mesh)."""
    import oracle
    from tests.parity import assert_same_frame
    work = scenes.batch_scenes()
    assert len(work) == 256 and sum(s.num_faces for s, a in work) > 8_000_000
    print("uploading 256 meshes", flush=True)
    meshes = [ctx.upload(s) for s, a in work]
    outs = [_device_outputs(ctx, a) for s, a in work]
    lights = [s.lights_copy() for s, a in work]
    batch = [(m, l, a, f, d, i) for m, l, (s, a), (f, d, i) in zip(meshes, lights, work, outs)]
    try:
        ts, cap, ms = ctx.autotune(batch)
        print("autotuned", ts, cap, ms, flush=True)
        g = ctx.batch_graph(batch, frames=2)
        for f, d, i in outs:
            i.upload(np.full(i.shape, 7, np.int32))
        g.launch()
        ctx.sync()
        g.close()
        rng = np.random.default_rng(2026)
        sample = set(rng.choice(256, 16, replace=False).tolist())
        overflowed = 0
        for k, ((scene, args), (f, d, i), m) in enumerate(zip(work, outs, meshes)):
            fb, db, ib = f.download(), d.download(), i.download()
            single = ctx.draw(m, scene.lights_copy(), args, ids=True, stats=True)
            assert np.array_equal(ib, single[2]) and np.array_equal(fb, single[0]), k
            assert np.array_equal(db.view(np.uint32), single[1].view(np.uint32)), k
            if single[3]["overflow_tiles"] > 0:
                overflowed += 1
                sample.add(k)
            if k in sample:
                print("oracle check of mesh", k, scene.num_faces, "faces", flush=True)
                ref = oracle.draw_frame(scene, args)
                assert_same_frame((fb, db, ib, single[3], lights[k]), ref, f"batch mesh {k}")
        print("meshes with overflowing bins:", overflowed, flush=True)
    finally:
        ctx.set_tuning(16, 0)
        for f, d, i in outs:
            f.close(); d.close(); i.close()
        for m in meshes:
            m.close()


@pytest.mark.parametrize("tile,cap", [(16, 512), (32, 2048), (8, 64), (16, 2)])
def test_batch_one_launch_equals_five(ctx, tile, cap):
    """A batch frame as one persistent launch and as five launches, one per pass (the
    default): the same bits and statistics, at every tile size and with overflowing bins;
    a batch graph of three frames replays the one-launch queue frame after frame."""
    work = scenes.batch_scenes(n=9, seed=41, max_faces=25_000)
    meshes = [ctx.upload(s) for s, a in work]
    outs = [_device_outputs(ctx, a) for s, a in work]
    batch = [(m, s.lights_copy(), a, f, d, i) for m, (s, a), (f, d, i) in zip(meshes, work, outs)]
    ctx.set_tuning(tile, cap)
    try:
        results = {}
        for persistent in (False, True):
            ctx.set_batch_persistent(persistent)
            print("batch", tile, cap, "persistent", persistent, flush=True)
            st = ctx.draw_batch(batch, stats=True)
            results[persistent] = (st, [(f.download(), d.download(), i.download()) for f, d, i in outs])
        (s5, r5), (s1, r1) = results[False], results[True]
        for key in ("front_triangles", "bin_pairs", "fragments", "covered_pixels", "overflow_tiles",
                    "raster_tiles", "raster_pixels"):
            assert s1[key] == s5[key], (key, s1[key], s5[key])
        if cap <= 2:
            assert s1["overflow_tiles"] > 0
        for k, (a, b) in enumerate(zip(r1, r5)):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), k
            assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), k
        g = ctx.batch_graph(batch, frames=3)
        for launch in range(2):
            for f, d, i in outs:
                i.upload(np.full(i.shape, 5, np.int32))
            g.launch()
            ctx.sync()
            for (f, d, i), b in zip(outs, r5):
                assert np.array_equal(i.download(), b[2]) and np.array_equal(f.download(), b[0]), launch
        g.close()
    finally:
        ctx.set_tuning(16, 0)
        ctx.set_batch_persistent(False)
        for m in meshes:
            m.close()
