"""Shared helpers of the GPU parity tests: run one scene through librz and the oracle
and compare with the bar of the frame path — face ids bit-exact, depth bit-exact (the
stated tolerance is 1 ulp; none is used), colour bit-exact (stated tolerance 1/255;
none is used)."""
import numpy as np

import oracle

DEPTH_ULP_TOL = 1      # documented bar; the comparisons below are exact
COLOUR_TOL = 1         # documented bar in 8-bit steps; the comparisons below are exact


def render_gpu(ctx, scene, args, frame=None, depth=None, tuning=None):
    if tuning is not None:
        ctx.set_tuning(*tuning)
    mesh = ctx.upload(scene)
    try:
        lights = scene.lights_copy()
        f = None if frame is None else frame.copy()
        d = None if depth is None else depth.copy()
        out = ctx.draw(mesh, lights, args, frame=f, depth=d, ids=True, stats=True)
    finally:
        mesh.close()
        if tuning is not None:
            ctx.set_tuning(16, 0)
    return out + (lights,)


def assert_same_frame(gpu, ref, label=""):
    frame, depth, ids, stats, lights = gpu
    of, od, oids, counts, olights = ref
    bad_ids = np.argwhere(ids != oids)
    assert bad_ids.size == 0, f"{label}: {len(bad_ids)} face-id mismatches, first at {bad_ids[:5].tolist()}"
    bad_d = np.argwhere(depth.view(np.uint32) != od.view(np.uint32))
    assert bad_d.size == 0, (f"{label}: {len(bad_d)} depth mismatches, first "
                             f"{[(tuple(p), depth[tuple(p)], od[tuple(p)]) for p in bad_d[:3]]}")
    bad_f = np.argwhere(frame != of)
    assert bad_f.size == 0, (f"{label}: {len(bad_f)} colour mismatches, first "
                             f"{[(tuple(p), frame[tuple(p)], of[tuple(p)]) for p in bad_f[:3]]}")
    assert np.array_equal(lights.view(np.uint32), olights.view(np.uint32)), f"{label}: trans_dir differs"
    assert stats["fragments"] == counts["fragments"], f"{label}: fragments {stats} vs {counts}"
    assert stats["front_triangles"] == counts["front_triangles"], f"{label}: front {stats} vs {counts}"
    assert stats["covered_pixels"] == int((oids >= 0).sum()), f"{label}: covered"


def check(ctx, scene, args, frame=None, depth=None, tuning=None, label=""):
    gpu = render_gpu(ctx, scene, args, frame, depth, tuning)
    ref = oracle.draw_frame(scene, args, frame=frame, depth=depth)
    assert_same_frame(gpu, ref, label)
    return gpu, ref
