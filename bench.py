#!/usr/bin/env python3
"""Frame-path benchmark on one MI355X.

Headline workload (`value`): the 81,920-face displaced, textured icosphere (6
subdivisions, smooth normals, 1024x1024 RGBA8 texture, three lights) at 1920x1080 with a
seeded YXZ rotation.  One step is one whole frame (bin with the vertex transform,
raster/depth, shade) with the mesh resident in HBM and the frame/depth buffers in device
memory.  Timed region: `steps` frames captured in one HIP graph, launched once between
two device synchronisations (after a graph of `warmup` frames); `value` = faces / frame
time in Mtri/s.

Every section reports its passes (`passes`): per pass the average launch time, its
algorithmic bytes (the inputs a pass must read and the outputs it must write, from the
per-unit figures below times the frame's own counts) and their rate against the 8 TB/s
HBM3E peak; `roofline` is the pass with the longest launch time, with `traffic`, the HBM
bytes per launch that hardware counters measured for that kernel on this workload
(profiles/hbm_traffic_*.md), when that file holds it.

Sections (each can be skipped):
  * the headline frame, fused pipeline (raster and shade in one pass) and the
    separate-shade pipeline (raster pass writing visibility words, then a shade pass);
  * `stress`: 4M screen-space triangles at 4096x4096 (edges 0.5-3 px, about half
    back-facing);
  * `batched`: 256 meshes (log-uniform 500-200k faces, icospheres and triangle soups,
    viewports 256^2-1024^2), tile size and bin capacity chosen by rz_autotune, batch
    frames in one HIP graph; per-pass times from HIP events around each launch;
  * `learned_shading`: the headline mesh with 64 lights and a 1024x1024 tangent-space
    normal map at 1920x1080; the BF16 MLP trained on 2^22 fragments of the plain shader,
    then plain and MLP shading timed three ways (whole frames, the shade pass, and the
    shade pass minus its per-pixel chain: light loop against network), with the MLP
    frame's error against the plain frame per channel.

`cpu_baseline` is the CPU oracle (a single-threaded restatement of the reference's
draw_frame) timed on this host for a bounded number of whole frames of the headline
workload.

Per-unit algorithmic bytes (T faces, T_front front faces, V vertices, Vn normals, P16
(front face, 16x16 tile) pairs, C covered pixels, T_vis faces winning a pixel, L lights,
tex texels, P_listed pixels of the tiles the raster pass visits):
  bin      12 T + 12 V + 12 Vn (16 V for screen-space vertices) + 4 P16
           + 7 W H when it clears the frame and depth (fused) or 8 W H the visibility words
  raster   fused: 48 T_front + 4 P16 + 7 C + 108 T_vis + 4 tex + 40 L
           separate: 48 T_front + 4 P16 + 8 P_listed (visibility words written)
  shade    separate: 8 W H + 7 W H + 108 T_vis + 4 tex + 40 L
  vertex   (batches) 28 V + 24 Vn
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
MLP_FLOPS_PER_FRAGMENT = 2 * (12 * 64 + 64 * 64 + 64 * 3)   # 10,112, unpadded
BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def texels(scene):
    return sum(int(m.texture.shape[0] * m.texture.shape[1]) for m in scene.materials if m.texture is not None)


def frame_bytes(scene, args, c, fused):
    """Algorithmic bytes per pass of one single frame (cleared buffers, no id buffer)."""
    W, H = args.image_width, args.image_height
    V, Vn, T, L = scene.vertices.shape[0], scene.normals.shape[0], scene.num_faces, scene.lights.shape[0]
    tex = texels(scene)
    shading = 108 * c["visible_faces"] + 4 * tex + 40 * L
    bin_ = 12 * T + (16 * V if scene.raster_space else 12 * V + 12 * Vn) + 4 * c["bin_pairs16"]
    raster_in = 48 * c["front_triangles"] + 4 * c["bin_pairs16"]
    if fused:
        return {"bin": bin_ + 7 * W * H, "raster": raster_in + 7 * c["covered_pixels"] + shading, "shade": 0}
    return {"bin": bin_ + 8 * W * H, "raster": raster_in + 8 * c["raster_pixels"],
            "shade": 8 * W * H + 7 * W * H + shading}


def pass_table(us, nbytes, names):
    """{pass: {avg_us, alg_bytes, GBps, frac, kernel}} for the passes with a launch."""
    out = {}
    for p, t in us.items():
        if t <= 0 or p not in names:
            continue
        b = nbytes.get(p, 0)
        gbs = b / (t * 1e-6) / 1e9
        out[p] = {"kernel": names[p], "avg_us": t, "alg_bytes": b, "GBps": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def traffic_bytes(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary
    profiles/hbm_traffic_<workload>.md (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def valu_time_us(workload, kernel):
    """Vector-unit time per launch of `kernel` on `workload` (its vector instructions x 4
    cycles over 1,024 SIMDs at 2.4 GHz) from the committed counter summary
    profiles/sq_counters_<workload>.md, or None."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "valu_time_us_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def roofline(table, workload):
    """The dominant pass against HBM, and beside it the bound it actually meets: the share
    of its launch the vector units would need to issue its instructions back to back."""
    dom = max(table, key=lambda p: table[p]["avg_us"])
    r = table[dom]
    vt = valu_time_us(workload, r["kernel"])
    issue = None if vt is None else {"valu_time_us": vt, "frac_of_launch": vt / r["avg_us"],
                                     "source": f"profiles/sq_counters_{workload}.md"}
    return {"bound": "hbm", "achieved": r["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac"],
            "traffic": traffic_bytes(workload, r["kernel"]), "kernel": r["kernel"], "pass": dom,
            "alg_bytes_per_launch": r["alg_bytes"], "avg_launch_us": r["avg_us"], "issue": issue}


def frame_counts(ctx, np, mesh, scene, args, frame, depth, ids, flags):
    """Counts of one frame from librz's statistics and its id buffer; P16 at 16-pixel tiles."""
    tuning = ctx.tuning()
    ctx.set_tuning(16, tuning[1])
    st = ctx.draw_device(mesh, scene.lights_copy(), args, frame.ptr, depth.ptr, ids.ptr, flags, stats=True)
    ctx.set_tuning(*tuning)
    idh = ids.download()
    return {"bin_pairs16": st["bin_pairs"], "front_triangles": st["front_triangles"],
            "visible_faces": int(np.unique(idh[idh >= 0]).size), "fragments": st["fragments"],
            "covered_pixels": st["covered_pixels"], "raster_tiles": st["raster_tiles"],
            "raster_pixels": st["raster_pixels"], "overflow_tiles": st["overflow_tiles"]}


def graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags, warm=None):
    """ms per frame of a graph of `frames` whole frames, launched after a warm-up graph."""
    gw = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, warm or frames, flags=flags)
    g = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, frames, flags=flags)
    try:
        gw.launch()
        ctx.sync()
        t0 = time.perf_counter()
        g.launch()
        ctx.sync()
        return (time.perf_counter() - t0) * 1000.0 / frames
    finally:
        g.close()
        gw.close()


def both_pipelines(ctx, np, mesh, scene, args, frame, depth, reps, counts, workload):
    """Per-pass tables of the fused and the separate-shade pipeline, and the roofline of
    the default (fused) one."""
    fused0 = ctx.fused_shade
    out = {}
    for fused in (True, False):
        ctx.set_fused_shade(fused)
        us = ctx.time_passes(mesh, scene.lights_copy(), args, frame, depth, reps=reps)
        names = {"bin": "k_bin", "raster": "k_raster_shade" if fused else "k_raster",
                 "shade": "k_shade_tiles" if fused else "k_shade"}
        out[fused] = (pass_table(us, frame_bytes(scene, args, counts, fused), names), us)
    ctx.set_fused_shade(fused0)
    return out[True][0], out[False][0], out[False][1]["raster"], roofline(out[True][0], workload)


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
        "sample": f"{n} whole frames (clear + draw_frame) of the same 81,920-face 1080p frame, "
                  f"{ms:.1f} ms/frame on 1 thread of {cpu} ({os.cpu_count()} logical CPUs)",
        "ms_per_frame": ms,
    }


def stress(ctx, rz, scenes, np, frames):
    """4M screen-space triangles at 4096x4096, default tuning."""
    scene, args = scenes.screen_soup_scene()
    W, H = args.image_width, args.image_height
    mesh = ctx.upload(scene)
    frame, depth = ctx.empty((3, H, W), np.uint8), ctx.empty((H, W), np.float32)
    ids = ctx.empty((H, W), np.int32)
    try:
        flags = rz.RZ_FLAG_CLEARED | rz.RZ_FLAG_ASYNC
        counts = frame_counts(ctx, np, mesh, scene, args, frame, depth, ids, flags)
        ms = graph_ms(ctx, mesh, scene.lights_copy(), args, frame, depth, frames, flags, warm=2)
        table, table_sep, raster_only, roof = both_pipelines(ctx, np, mesh, scene, args, frame, depth, frames,
                                                             counts, "stress_4m")
        return {"workload": "screen-soup-4M-4096x4096", "faces": scene.num_faces, "width": W, "height": H,
                "ms_per_frame": ms, "mtri_per_s": scene.num_faces / (ms * 1e-3) / 1e6,
                "gfrag_per_s": counts["fragments"] / (raster_only * 1e-6) / 1e9,
                "counts": counts, "passes": table, "passes_separate_shade": table_sep, "roofline": roof,
                "frames_per_graph": frames}
    finally:
        for b in (frame, depth, ids):
            b.close()
        mesh.close()


def batch_graph_ms(ctx, batch, frames):
    """ms per batch frame of a graph of `frames` batch frames, after one warm-up launch."""
    g = ctx.batch_graph(batch, frames)
    try:
        g.launch()
        ctx.sync()
        t0 = time.perf_counter()
        g.launch()
        ctx.sync()
        return (time.perf_counter() - t0) * 1000.0 / frames
    finally:
        g.close()


def batched(ctx, rz, scenes, np, frames):
    """The 256-mesh batch: autotuned, `frames` batch frames in one HIP graph."""
    t0 = time.perf_counter()
    work = scenes.batch_scenes()
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    meshes = [ctx.upload(sc) for sc, _ in work]
    t_up = time.perf_counter() - t0
    outs = []
    for _, a in work:
        W, H = a.image_width, a.image_height
        outs.append((ctx.empty((3, H, W), np.uint8), ctx.empty((H, W), np.float32), ctx.empty((H, W), np.int32)))
    batch = [(m, sc.lights_copy(), a, f, d, None) for m, (sc, a), (f, d, i) in zip(meshes, work, outs)]
    with_ids = [(m, sc.lights_copy(), a, f, d, i) for m, (sc, a), (f, d, i) in zip(meshes, work, outs)]
    try:
        t0 = time.perf_counter()
        ts, cap, tuned_ms = ctx.autotune(batch)
        t_tune = time.perf_counter() - t0
        # the context's batch launch shape (one persistent launch or five), and the other
        default_one = ctx.batch_persistent
        launch_ms = {}
        for one in (default_one, not default_one):
            ctx.set_batch_persistent(one)
            launch_ms["one_launch" if one else "five_launches"] = batch_graph_ms(ctx, batch, frames)
        ctx.set_batch_persistent(default_one)
        ms = launch_ms["one_launch" if default_one else "five_launches"]
        # counts: the tuned frame's statistics and ids; P16 at 16-pixel tiles
        st = ctx.draw_batch(with_ids, stats=True)
        ctx.set_tuning(16, 512)   # (pairs are counted from the faces' tile rectangles, overflow or not)
        p16 = ctx.draw_batch(batch, stats=True)["bin_pairs"]
        ctx.set_tuning(ts, cap)
        vis = 0
        for f, d, i in outs:
            idh = i.download()
            vis += int(np.unique(idh[idh >= 0]).size)
        # per-pass times: HIP events around each launch of one batch frame of five launches
        # (median of 5); the one-launch frame's phases as spans of the workgroups' clocks
        ctx.set_profiling(True)
        phases = {}
        for one in (False, True):
            ctx.set_batch_persistent(one)
            runs = [ctx.draw_batch(batch, stats=True) for _ in range(5)]
            phases[one] = {k: float(np.median([r["ms_" + k] for r in runs])) * 1000.0
                           for k in ("vertex", "bin", "raster", "overflow", "shade", "frame")}
        ctx.set_batch_persistent(default_one)
        ctx.set_profiling(False)
        med = phases[False]
        V = sum(sc.vertices.shape[0] for sc, _ in work)
        Vn = sum(sc.normals.shape[0] for sc, _ in work)
        T = sum(sc.num_faces for sc, _ in work)
        L = sum(sc.lights.shape[0] for sc, _ in work)
        pixels = sum(a.image_width * a.image_height for _, a in work)
        nbytes = {"vertex": 28 * V + 24 * Vn,
                  "bin": 12 * T + 16 * V + 4 * p16 + 7 * pixels,
                  "raster": 48 * st["front_triangles"] + 4 * p16 + 7 * st["covered_pixels"] + 108 * vis + 40 * L,
                  "overflow": 0, "shade": 0}
        names = {"vertex": "k_vertex_batch", "bin": "k_bin_batch", "raster": "k_raster_shade_batch",
                 "overflow": "k_overflow_batch", "shade": "k_shade_tiles_batch"}
        table = pass_table(med, nbytes, names)
        return {"workload": "256 meshes, log-uniform 500-200k faces, viewports 256^2-1024^2",
                "meshes": len(work), "faces": T, "pixels": pixels, "tile": ts, "bin_capacity": cap,
                "autotune_ms_per_frame": tuned_ms, "ms_per_batch_frame": ms,
                "launch": "one persistent launch" if default_one else "five launches",
                "ms_per_batch_frame_by_launch": launch_ms,
                "one_launch_phase_spans_us": {k: phases[True][k] for k in ("bin", "raster", "overflow", "shade")},
                "mtri_per_s": T / (ms * 1e-3) / 1e6, "gpix_per_s": pixels / (ms * 1e-3) / 1e9,
                "counts": {"front_triangles": st["front_triangles"], "bin_pairs16": p16, "visible_faces": vis,
                           "fragments": st["fragments"], "covered_pixels": st["covered_pixels"],
                           "overflow_tiles": st["overflow_tiles"]},
                "passes": table, "roofline": roofline(table, "batch_256"),
                "generate_s": t_gen, "upload_s": t_up, "autotune_s": t_tune, "frames_per_graph": frames}
    finally:
        for f, d, i in outs:
            f.close()
            d.close()
            i.close()
        for m in meshes:
            m.close()
        ctx.set_tuning(16, 0)


def learned_shading(ctx, rz, scenes, np, steps, batch, frames):
    """The learned shader on the 64-light normal-mapped 1080p icosphere, then the same
    frame with 256 and 1,024 lights (intensities scaled by 64 / lights, so the frame
    keeps its brightness): the light loop grows with the lights, the network does not."""
    out = learned_at(ctx, rz, scenes, np, 64, steps, batch, frames)
    sweep = []
    for nl in (64, 256, 1024):
        r = out if nl == 64 else learned_at(ctx, rz, scenes, np, nl, steps, batch, frames)
        sweep.append({"lights": nl, "shade_pass_us": r["shade_pass_us"], "light_loop_us": r["light_loop_us"],
                      "network_us": r["network_us"], "speedup_shade_pass": r["speedup_shade_pass"],
                      "speedup_frame": r["speedup_frame"], "frame_ms": r["frame_ms"],
                      "err_mean_255": r["err_mean_255"], "err_max_per_channel_255": r["err_max_per_channel_255"]})
        log(f"mlp at {nl} lights: shade pass plain {r['shade_pass_us']['plain']:.1f} us, "
            f"mlp {r['shade_pass_us']['mlp']:.1f} us, chain {r['shade_pass_us']['chain']:.1f} us; "
            f"mean error {r['err_mean_255']:.2f}/255")
    out["light_sweep"] = sweep
    return out


def learned_at(ctx, rz, scenes, np, nl, steps, batch, frames):
    scene, args = scenes.learned_shading_scene(lights=nl)
    scene.lights[:, 3] *= np.float32(64.0 / nl)
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
        fused0 = ctx.fused_shade
        timing = {}
        for fused in (False, True):
            ctx.set_fused_shade(fused)
            tp = ctx.time_passes(mesh, lights, args, frame, depth, reps=frames)
            tm = ctx.time_passes(mesh, lights, args, frame, depth, reps=frames, flags=rz.RZ_FLAG_MLP_SHADE)
            key = "raster" if fused else "shade"   # the pass that shades
            timing[fused] = {"plain_us": tp[key], "mlp_us": tm[key], "chain_us": tp["chain"],
                             "frame_plain_ms": graph_ms(ctx, mesh, lights, args, frame, depth, frames, flags),
                             "frame_mlp_ms": graph_ms(ctx, mesh, lights, args, frame, depth, frames,
                                                      flags | rz.RZ_FLAG_MLP_SHADE)}
        ctx.set_fused_shade(fused0)
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
        sep, fus = timing[False], timing[True]
        light_loop_us = sep["plain_us"] - sep["chain_us"]
        network_us = sep["mlp_us"] - sep["chain_us"]
        net_tflops = frags * MLP_FLOPS_PER_FRAGMENT / (network_us * 1e-6) / 1e12 if network_us > 0 else None
        return {
            "workload": f"icosphere-s6-1080p-{nl}-lights-normal-map-1024",
            "training_samples": int(n), "training_rotations": rotations,
            "train_steps": steps, "batch": batch, "train_s": t_train, "collect_s": t_collect,
            "loss_first": float(losses[0]), "loss_last100": float(losses[-100:].mean()),
            "fragments": frags,
            "shade_pass_us": {"plain": sep["plain_us"], "mlp": sep["mlp_us"], "chain": sep["chain_us"]},
            "raster_shade_pass_us": {"plain": fus["plain_us"], "mlp": fus["mlp_us"], "chain": fus["chain_us"]},
            "light_loop_us": light_loop_us, "network_us": network_us,
            "speedup_light_loop_vs_network": light_loop_us / network_us if network_us > 0 else None,
            "speedup_shade_pass": sep["plain_us"] / sep["mlp_us"],
            "speedup_frame": fus["frame_plain_ms"] / fus["frame_mlp_ms"],
            "frame_ms": {"plain": fus["frame_plain_ms"], "mlp": fus["frame_mlp_ms"],
                         "plain_separate_shade": sep["frame_plain_ms"], "mlp_separate_shade": sep["frame_mlp_ms"]},
            "network_tflops_in_pass": net_tflops,
            "network_frac_of_bf16_peak": net_tflops / BF16_PEAK_TFLOPS if net_tflops else None,
            "err_mean_per_channel_255": [float(e) for e in err.mean(axis=1)],
            "err_max_per_channel_255": [int(e) for e in err.max(axis=1)],
            "err_mean_255": float(err.mean()),
        }
    finally:
        for b in (samples, frame, depth, ids):
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
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned-shading section")
    ap.add_argument("--no-stress", action="store_true", help="skip the 4M-triangle 4096x4096 section")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-mesh batch section")
    ap.add_argument("--only", choices=["frame", "stress", "batch", "mlp"], default=None,
                    help="run only the headline frame (or it and one section): for profiling runs")
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
    if a.only is not None:
        a.no_stress, a.no_batch, a.no_mlp = a.only != "stress", a.only != "batch", a.only != "mlp"

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
    counts = frame_counts(ctx, np, mesh, scene, args, frame, depth, ids, flags)
    log(f"frame counts: {counts}")

    # timed region: `steps` frames in one graph (after a graph of `warmup` frames)
    g_warm = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, max(1, a.warmup), flags=flags)
    g = ctx.graph(mesh, lights, args, frame.ptr, depth.ptr, a.steps, flags=flags)
    g_warm.launch()
    ctx.sync()
    t0 = time.perf_counter()
    g.launch()
    ctx.sync()
    t1 = time.perf_counter()
    g.close()
    g_warm.close()
    ms_per_step = (t1 - t0) * 1000.0 / a.steps
    value = scene.num_faces / (ms_per_step / 1000.0) / 1e6
    log(f"graph: {a.steps} frames in {(t1 - t0) * 1000:.2f} ms -> {ms_per_step * 1000:.2f} us/frame")

    reps = max(1, min(a.steps, 200))
    per_pass, per_pass_sep, raster_only_us, roof = both_pipelines(ctx, np, mesh, scene, args, frame, depth, reps,
                                                                  counts, "icosphere_1080p")
    log("per pass:", json.dumps(per_pass))
    log("per pass, separate shade:", json.dumps(per_pass_sep))
    gfrag = counts["fragments"] / (raster_only_us * 1e-6) / 1e9 if raster_only_us > 0 else None

    # the host-buffer boundary: draw_frame into host memory (frame and depth copied back
    # over PCIe every frame), reported beside the device-buffer rate
    host_frames = 20
    ctx.draw(mesh, scene.lights_copy(), args)
    t0 = time.perf_counter()
    for _ in range(host_frames):
        ctx.draw(mesh, scene.lights_copy(), args)
    host_ms = (time.perf_counter() - t0) * 1000.0 / host_frames
    log(f"host-buffer frames (PCIe copies included): {host_ms:.3f} ms")

    cpu = None
    if not a.no_cpu and a.only is None:
        cpu = cpu_baseline(scene, args, a.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    mesh.close()
    for buf in (frame, depth, ids):
        buf.close()

    sections = {}
    if not a.no_stress:
        sections["stress"] = stress(ctx, rz, scenes, np, 20)
        log(f"stress: {json.dumps(sections['stress'])}")
    if not a.no_batch:
        sections["batched"] = batched(ctx, rz, scenes, np, 20)
        log(f"batched: {json.dumps(sections['batched'])}")
    if not a.no_mlp:
        ctx.set_tuning(16, 0)
        sections["learned_shading"] = learned_shading(ctx, rz, scenes, np, a.mlp_steps, a.mlp_batch,
                                                      max(10, min(a.steps, 100)))
        log(f"learned shading: {json.dumps(sections['learned_shading'])}")
    ctx.close()
    mlp = sections.get("learned_shading")
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
        "counts": counts,
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu_frame_rate": (cpu["ms_per_frame"] / ms_per_step) if cpu else None,
        "passes": per_pass,
        "passes_separate_shade": per_pass_sep,
        "shade64_plain_ms": mlp["shade_pass_us"]["plain"] / 1000.0 if mlp else None,
        "shade64_mlp_ms": mlp["shade_pass_us"]["mlp"] / 1000.0 if mlp else None,
        "learned_shading": mlp,
        "batched": sections.get("batched"),
        "stress": sections.get("stress"),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
