#!/usr/bin/env python3
"""Frame-path benchmark on one MI355X.

Workload: the 81,920-face displaced, textured icosphere (6 subdivisions, smooth
normals, 1024x1024 RGBA8 texture, three lights) at 1920x1080 with a seeded YXZ
rotation — one step is one whole frame (bin with the vertex transform, raster/depth,
overflow, shade) with the mesh resident in HBM and the frame/depth buffers in device
memory.

Timed region: `steps` frames captured in one HIP graph, launched once between two
device synchronisations.  `value` = faces / frame time in Mtri/s.

`roofline` is for the dominant kernel of the frame: the pass with the longest
average duration.  Each pass is timed as a HIP graph of min(steps, 200) back-to-back
launches of its kernel on the librz stream, with HIP events around the graph
(rz_time_passes); `achieved` is the pass's algorithmic bytes per launch (counts from
the frame itself) divided by that duration; `traffic` is the HBM bytes per launch read
from the committed hardware-counter summary under profiles/.

`cpu_baseline` is the CPU oracle (a single-threaded restatement of the reference's
draw_frame) timed on this host for a bounded number of full frames of the same
workload.

`learned_shading` (skipped with --no-mlp): the same mesh with 64 lights and a
1024x1024 tangent-space normal map at 1920x1080.  The plain shader renders frames at
seeded rotations until 2^22 fragments are collected; the BF16 MLP is trained on them
by the HIP training kernels; then the deferred shade pass (run as its own pass) is
timed with the plain 64-light loop and with the network (rz_time_passes), whole frames
are timed both ways in both pipelines, and the MLP frame's error against the plain
frame is reported per channel over the covered pixels of the workload's own rotation
(not among the training rotations).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mtri/s + Gfrag/s at 1080p (frac of HBM peak); 64-light shade ms plain vs MLP"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(counts, scene, args, fused, cleared=True):
    """Per-pass algorithmic bytes of one frame: the input reads and output writes a pass
    must make at minimum.  With the fused pipeline the bin launch also clears the frame
    and depth, the raster pass also shades (no visibility buffer round trip) and the
    separate shade pass only covers overflowed tiles (none in this frame)."""
    W, H = args.image_width, args.image_height
    V = scene.vertices.shape[0]
    Vn = scene.normals.shape[0]
    T = scene.num_faces
    L = scene.lights.shape[0]
    tex = sum(int(m.texture.shape[0] * m.texture.shape[1]) for m in scene.materials if m.texture is not None)
    P16, Tf, Tvis = counts["bin_pairs16"], counts["front_triangles"], counts["visible_faces"]
    depth_in = 0 if cleared else 4 * W * H
    raster_in = 48 * Tf + 4 * P16 + depth_in
    shade_out = 7 * W * H + 108 * Tvis + 4 * tex + 40 * L   # frame + depth, winners' data, texture, lights
    # the bin pass transforms each face's corners itself (no vertex pass): it reads the
    # face indices, the positions (screen-space input: 16 B raster vertices, as given) and
    # the normals, and writes the bin entries
    raster_space = getattr(scene, "raster_space", False)
    out = {"bin": 12 * T + (16 if raster_space else 12) * V + 12 * Vn + 4 * P16}
    if fused:
        # cleared buffers: the bin launch writes the cleared frame and depth (7 B/px), the
        # raster pass shades the covered pixels only
        C = counts["covered_pixels"]
        out["bin"] += 7 * W * H
        out["raster"] = raster_in + 7 * C + 108 * Tvis + 4 * tex + 40 * L
        out["shade"] = 0
    else:
        out["raster"] = raster_in + 8 * W * H        # + visibility words written
        out["shade"] = 8 * W * H + shade_out         # + visibility words read
    return out


def per_pass_table(us, nbytes):
    table = {}
    for p, t in us.items():
        if p == "network" and t == 0:
            continue   # the learned shader's pass: not part of the plain frame
        if p in ("overflow", "vertex", "shade") and t == 0:
            continue   # single frames: the overflow loop runs inside the raster pass, the
                       # vertex transform inside the bin pass; no overflow-shade launch when
                       # no bin can overflow
        b = nbytes.get(p, 0)
        gbs = b / (t * 1e-6) / 1e9 if t > 0 else 0.0
        table[p] = {"avg_us": t, "alg_bytes": b, "GBps": gbs, "frac": gbs / HBM_PEAK_GBS}
    return table


def cpu_baseline(scene, args, seconds):
    import oracle
    oracle.draw_frame(scene, args)   # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.draw_frame(scene, args)
        n += 1
        el = time.perf_counter() - t0
        if (el >= seconds and n >= 3) or n >= 10000:
            break
    ms = el * 1000.0 / n
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": scene.num_faces / (ms / 1000.0) / 1e6,
        "unit": "Mtri/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} full frames (clear + draw_frame) of the same 81,920-face 1080p frame, "
                  f"{ms:.1f} ms/frame on 1 thread of {cpu} ({os.cpu_count()} logical CPUs)",
        "ms_per_frame": ms,
    }


MLP_FLOPS_PER_FRAGMENT = 2 * (12 * 64 + 64 * 64 + 64 * 3)   # 10,112, unpadded


def graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags):
    """ms per frame of a graph of `frames` whole frames (second launch)."""
    g = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, frames, flags=flags)
    try:
        g.launch()
        ctx.sync()
        t0 = time.perf_counter()
        g.launch()
        ctx.sync()
        return (time.perf_counter() - t0) * 1000.0 / frames
    finally:
        g.close()


def traffic_bytes(kernel):
    """HBM bytes per launch of `kernel` from the committed hardware-counter summary
    (profiles/hbm_traffic_icosphere_1080p.md: FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_icosphere_1080p.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch ") and line.split()[1] == kernel:
                    return float(line.split()[2])
    except OSError:
        pass
    return None


def learned_shading(ctx, rz, scenes, np, steps, batch, frames):
    import ctypes

    scene, args = scenes.learned_shading_scene()
    W, H = args.image_width, args.image_height
    mesh = ctx.upload(scene)
    frame = ctx.empty((3, H, W), np.uint8)
    depth = ctx.empty((H, W), np.float32)
    ids = ctx.empty((H, W), np.int32)
    n_target = 1 << 22
    samples = ctx.empty((n_target, rz.RZ_MLP_SAMPLE_FLOATS), np.float32)
    try:
        # training fragments: the plain shader at seeded rotations
        rng = np.random.default_rng(0x5EED2000)
        n, rotations = 0, 0
        t0 = time.perf_counter()
        while n < n_target:
            a_k = rz.make_args(W, H, angles=tuple(rng.uniform(-np.pi, np.pi, 3)))
            a_k.fovy_degrees, a_k.z_near, a_k.z_far = args.fovy_degrees, args.z_near, args.z_far
            a_k.scale, a_k.displacement[:] = args.scale, list(args.displacement)
            n = ctx.mlp_collect(mesh, scene.lights_copy(), a_k, samples, n)
            rotations += 1
        t_collect = time.perf_counter() - t0
        ctx.mlp_init(0x5EED3000)
        t0 = time.perf_counter()
        losses = ctx.mlp_train(samples, n, steps=steps, batch=batch, lr=3e-3, seed=0x5EED4000)
        t_train = time.perf_counter() - t0
        log(f"mlp: {n} samples from {rotations} rotations in {t_collect:.2f} s; {steps} steps x {batch} "
            f"in {t_train:.2f} s; loss {losses[0]:.5f} -> {losses[-100:].mean():.6f}")

        lights = scene.lights_copy()
        flags = rz.RZ_FLAG_CLEARED | rz.RZ_FLAG_ASYNC
        # the shade pass on its own (separate-pass pipeline), plain light loop vs network
        fused = ctx.fused_shade
        ctx.set_fused_shade(False)
        plain_ms = ctx.time_passes(mesh, lights, args, frame, depth, reps=frames)["shade"] / 1000.0
        t_mlp = ctx.time_passes(mesh, lights, args, frame, depth, reps=frames, flags=rz.RZ_FLAG_MLP_SHADE)
        mlp_records_ms = t_mlp["shade"] / 1000.0      # shade pass writing the fragment records
        mlp_network_ms = t_mlp["network"] / 1000.0    # the network on the matrix cores
        mlp_ms = mlp_records_ms + mlp_network_ms
        split_plain_frame_ms = graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags)
        split_mlp_frame_ms = graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags | rz.RZ_FLAG_MLP_SHADE)
        ctx.set_fused_shade(True)
        plain_frame_ms = graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags)
        mlp_frame_ms = graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags | rz.RZ_FLAG_MLP_SHADE)
        ctx.set_fused_shade(fused)
        ctx.draw_device(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, ids.ptr, flags)
        ctx.sync()
        fp, idp = frame.download(), ids.download()
        ctx.draw_device(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, ids.ptr,
                        flags | rz.RZ_FLAG_MLP_SHADE)
        ctx.sync()
        fm, idm = frame.download(), ids.download()
        cov = idp >= 0
        if not np.array_equal(idp, idm):
            raise RuntimeError("MLP frame changed coverage")
        err = np.abs(fm[:, cov].astype(np.int32) - fp[:, cov].astype(np.int32))
        frags = int(cov.sum())
        return {
            "workload": "icosphere-s6-1080p-64-lights-normal-map-1024",
            "training_samples": int(n), "training_rotations": rotations,
            "train_steps": steps, "batch": batch, "train_s": t_train, "collect_s": t_collect,
            "loss_first": float(losses[0]), "loss_last100": float(losses[-100:].mean()),
            "fragments": frags,
            "shade_plain_ms": plain_ms, "shade_mlp_ms": mlp_ms,
            "shade_mlp_records_ms": mlp_records_ms, "shade_mlp_network_ms": mlp_network_ms,
            "frame_plain_ms": plain_frame_ms, "frame_mlp_ms": mlp_frame_ms,
            "frame_plain_ms_separate_shade": split_plain_frame_ms,
            "frame_mlp_ms_separate_shade": split_mlp_frame_ms,
            "speedup": plain_ms / mlp_ms if mlp_ms > 0 else None,
            "mlp_network_tflops": (frags * MLP_FLOPS_PER_FRAGMENT / (mlp_network_ms * 1e-3) / 1e12
                                   if mlp_network_ms > 0 else None),
            "err_mean_per_channel_255": [float(e) for e in err.mean(axis=1)],
            "err_max_per_channel_255": [int(e) for e in err.max(axis=1)],
            "err_mean_255": float(err.mean()),
        }
    finally:
        for b in (samples, frame, depth, ids):
            b.close()
        mesh.close()


def batched(ctx, rz, scenes, np, frames):
    """The 256-mesh batched workload: autotuned tile size and bin capacity, then
    `frames` batch frames (every mesh, one launch per pass) in one HIP graph."""
    t0 = time.perf_counter()
    work = scenes.batch_scenes()
    t_gen = time.perf_counter() - t0
    meshes = [ctx.upload(sc) for sc, _ in work]
    outs = []
    for _, a in work:
        W, H = a.image_width, a.image_height
        outs.append((ctx.empty((3, H, W), np.uint8), ctx.empty((H, W), np.float32)))
    batch = [(m, sc.lights_copy(), a, f, d, None) for m, (sc, a), (f, d) in zip(meshes, work, outs)]
    try:
        t0 = time.perf_counter()
        ts, cap, tuned_ms = ctx.autotune(batch)
        t_tune = time.perf_counter() - t0
        st = ctx.draw_batch(batch, stats=True)
        g = ctx.batch_graph(batch, frames)
        g.launch()
        ctx.sync()
        t0 = time.perf_counter()
        g.launch()
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1000.0 / frames
        g.close()
        faces = sum(sc.num_faces for sc, _ in work)
        pixels = sum(a.image_width * a.image_height for _, a in work)
        return {"workload": "256 meshes, log-uniform 500-200k faces, viewports 256^2-1024^2",
                "meshes": len(work), "faces": faces, "pixels": pixels, "tile": ts, "bin_capacity": cap,
                "autotune_ms_per_frame": tuned_ms, "ms_per_batch_frame": ms,
                "mtri_per_s": faces / (ms * 1e-3) / 1e6, "gpix_per_s": pixels / (ms * 1e-3) / 1e9,
                "fragments": st["fragments"], "overflow_tiles": st["overflow_tiles"],
                "generate_s": t_gen, "autotune_s": t_tune, "frames_per_graph": frames}
    finally:
        for f, d in outs:
            f.close()
            d.close()
        for m in meshes:
            m.close()
        ctx.set_tuning(16, 0)


def stress(ctx, rz, scenes, np, frames):
    """The small-triangle depth stress: 4M screen-space triangles at 4096x4096 (about half
    back-facing), default tuning; frame time from a graph of `frames` frames, per-pass
    times from rz_time_passes."""
    scene, args = scenes.screen_soup_scene()
    W, H = args.image_width, args.image_height
    mesh = ctx.upload(scene)
    frame, depth = ctx.empty((3, H, W), np.uint8), ctx.empty((H, W), np.float32)
    ids = ctx.empty((H, W), np.int32)
    try:
        flags = rz.RZ_FLAG_CLEARED | rz.RZ_FLAG_ASYNC
        st = ctx.draw_device(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, ids.ptr, flags, stats=True)
        idh = ids.download()
        counts = {"bin_pairs16": st["bin_pairs"], "front_triangles": st["front_triangles"],
                  "visible_faces": int(np.unique(idh[idh >= 0]).size), "fragments": st["fragments"],
                  "covered_pixels": st["covered_pixels"]}
        ms = graph_ms(ctx, mesh, scene.lights_copy(), args, frame, depth, frames, flags)
        us = ctx.time_passes(mesh, scene.lights_copy(), args, frame, depth, reps=frames)
        fused = ctx.fused_shade
        table = per_pass_table(us, algorithmic_bytes(counts, scene, args, fused))
        ctx.set_fused_shade(not fused)
        us2 = ctx.time_passes(mesh, scene.lights_copy(), args, frame, depth, reps=frames)
        table2 = per_pass_table(us2, algorithmic_bytes(counts, scene, args, not fused))
        ctx.set_fused_shade(fused)
        raster_only = us2["raster"] if fused else us["raster"]
        return {"workload": "screen-soup-4M-4096x4096", "faces": scene.num_faces, "width": W, "height": H,
                "ms_per_frame": ms, "mtri_per_s": scene.num_faces / (ms * 1e-3) / 1e6,
                "gfrag_per_s": counts["fragments"] / (raster_only * 1e-6) / 1e9 if raster_only > 0 else None,
                "counts": counts, "overflow_tiles": st["overflow_tiles"], "passes": table,
                "passes_other_pipeline": table2, "frames_per_graph": frames}
    finally:
        for b in (frame, depth, ids):
            b.close()
        mesh.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--flat", action="store_true", help="flat instead of smooth normals")
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--cap", type=int, default=0, help="bin capacity (0: librz's automatic choice)")
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned-shading measurement")
    ap.add_argument("--stress", action="store_true",
                    help="also measure the 4M-triangle 4096x4096 depth stress frame")
    ap.add_argument("--batch", action="store_true",
                    help="also measure the 256-mesh batched workload (autotuned, one graph)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (k_copy), the "
                         "known-size reference for hardware byte counters")
    ap.add_argument("--mlp-steps", type=int, default=6000)
    ap.add_argument("--mlp-batch", type=int, default=16384)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the frame path is "
              f"built and measured on a single MI355X)", file=sys.stderr)
        return 2

    import numpy as np

    import rasteriser_amd as rz
    from rasteriser_amd import scenes

    scene, args = scenes.icosphere_scene(flat=a.flat)
    W, H = args.image_width, args.image_height
    ctx = rz.Context(0)
    if a.copy_mib > 0:
        n = a.copy_mib << 20
        src, dst = ctx.empty((n,), np.uint8), ctx.empty((n,), np.uint8)
        dst.copy_from(src)
        log(f"k_copy: {n} bytes read, {n} bytes written")
        src.close()
        dst.close()
    ctx.set_tuning(a.tile, a.cap)
    mesh = ctx.upload(scene)
    frame = ctx.empty((3, H, W), np.uint8)
    depth = ctx.empty((H, W), np.float32)
    ids = ctx.empty((H, W), np.int32)
    lights = scene.lights_copy()
    flags = rz.RZ_FLAG_CLEARED | rz.RZ_FLAG_ASYNC

    # counts of this frame (same counts every step: identical frames)
    st = ctx.draw_device(mesh, lights, args, frame.ptr, depth.ptr, ids.ptr,
                         flags, stats=True)
    idh = ids.download()
    counts = {"bin_pairs16": st["bin_pairs"] if a.tile == 16 else None,
              "front_triangles": st["front_triangles"],
              "visible_faces": int(np.unique(idh[idh >= 0]).size),
              "fragments": st["fragments"], "covered_pixels": st["covered_pixels"]}
    if counts["bin_pairs16"] is None:
        ctx.set_tuning(16, a.cap)
        counts["bin_pairs16"] = ctx.draw_device(mesh, lights, args, frame.ptr, depth.ptr,
                                                0, flags, stats=True)["bin_pairs"]
        ctx.set_tuning(a.tile, a.cap)
    log(f"frame counts: {counts}")

    # timed region: `steps` frames in one graph
    g_warm = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, max(1, a.warmup), flags=flags)
    g = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, a.steps, flags=flags)
    g_warm.launch()
    ctx.sync()
    t0 = time.perf_counter()
    g.launch()
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / a.steps
    value = scene.num_faces / (ms_per_step / 1000.0) / 1e6
    log(f"graph: {a.steps} frames in {(t1 - t0) * 1000:.2f} ms -> {ms_per_step * 1000:.2f} us/frame")

    # per-pass durations: for each pass a HIP graph of `reps` launches of that kernel on the
    # librz stream, timed with HIP events (rz_time_passes); first the default fused
    # pipeline, then with the deferred shade as its own pass over the visibility buffer
    reps = max(1, min(a.steps, 200))
    fused = ctx.fused_shade
    us = ctx.time_passes(mesh, lights, args, frame, depth, reps=reps)
    per_pass = per_pass_table(us, algorithmic_bytes(counts, scene, args, fused))
    ctx.set_fused_shade(not fused)
    us_other = ctx.time_passes(mesh, lights, args, frame, depth, reps=reps)
    per_pass_other = per_pass_table(us_other, algorithmic_bytes(counts, scene, args, not fused))
    ctx.set_fused_shade(fused)
    log("per pass:", json.dumps(per_pass))
    log("per pass, other pipeline:", json.dumps(per_pass_other))
    dom = max(per_pass, key=lambda p: per_pass[p]["avg_us"])
    kernel = {"bin": "k_bin",
              "raster": "k_raster_shade" if fused else "k_raster",
              "shade": "k_shade_tiles" if fused else "k_shade"}[dom]
    roofline = {"bound": "hbm", "achieved": per_pass[dom]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": per_pass[dom]["frac"], "traffic": traffic_bytes(kernel), "kernel": kernel,
                "alg_bytes_per_launch": per_pass[dom]["alg_bytes"], "avg_launch_us": per_pass[dom]["avg_us"]}
    # Gfrag/s: inside-test fragments over the raster/depth pass proper (the separate-shade
    # pipeline's raster kernel, which does nothing else)
    raster_only_us = us_other["raster"] if fused else us["raster"]
    gfrag = counts["fragments"] / (raster_only_us * 1e-6) / 1e9 if raster_only_us > 0 else None

    # the host-buffer boundary: draw_frame into host memory (frame and depth copied back
    # over PCIe every frame), for reference only
    host_frames = 20
    ctx.draw(mesh, scene.lights_copy(), args)
    t0 = time.perf_counter()
    for _ in range(host_frames):
        ctx.draw(mesh, scene.lights_copy(), args)
    host_ms = (time.perf_counter() - t0) * 1000.0 / host_frames
    log(f"host-buffer frames (PCIe copies included): {host_ms:.3f} ms")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(scene, args, a.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    g.close()
    g_warm.close()
    mesh.close()
    for buf in (frame, depth, ids):
        buf.close()

    stressed = None
    if a.stress:
        stressed = stress(ctx, rz, scenes, np, 20)
        log(f"stress: {stressed}")

    batch = None
    if a.batch:
        batch = batched(ctx, rz, scenes, np, 20)
        log(f"batched: {batch}")

    mlp = None
    if not a.no_mlp:
        ctx.set_tuning(16, 0)
        mlp = learned_shading(ctx, rz, scenes, np, a.mlp_steps, a.mlp_batch, max(10, min(a.steps, 100)))
        log(f"learned shading: {mlp}")
    ctx.close()
    out = {
        "metric": METRIC, "value": value, "unit": "Mtri/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "icosphere-s6-displaced-textured-1080p", "faces": scene.num_faces,
                   "width": W, "height": H, "shading": "flat" if a.flat else "smooth",
                   "lights": int(scene.lights.shape[0]), "texture": "1024x1024 RGBA8",
                   "tile": a.tile, "bin_capacity": a.cap, "frames_per_graph": a.steps},
        "gfrag_per_s": gfrag,
        "frame_ms_host_buffers": host_ms,
        "fragments_per_frame": counts["fragments"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu_frame_rate": (cpu["ms_per_frame"] / ms_per_step) if cpu else None,
        "pipeline": "fused raster+shade" if fused else "separate shade pass",
        "passes": per_pass,
        "passes_other_pipeline": per_pass_other,
        "shade64_plain_ms": mlp["shade_plain_ms"] if mlp else None,
        "shade64_mlp_ms": mlp["shade_mlp_ms"] if mlp else None,
        "learned_shading": mlp,
        "batched": batch,
        "stress": stressed,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
