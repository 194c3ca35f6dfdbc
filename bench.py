#!/usr/bin/env python3
"""Benchmark of the rendering hot path on one MI355X.

One step = one frame of the instanced scene (1024 instances of 16 seeded
meshes with 2k-20k triangles, ~8.3M instanced triangles, two-level BVH,
3 quad lights + constant sky) at 1920x1080, 64 spp, max depth 8, rendered by
the wavefront path tracer into device-resident accumulation buffers (the
reference's running mean / variance with its early-out).  The scene is built
and uploaded before timing.

Printed: one JSON line with Mrays/s (rays of the accumulated samples / wall
time of the timed steps), the roofline of the traversal kernel (`extend`),
the reference CPU renderer's Mrays/s on the host cores, and the quality leg:
relMSE against a 4096-spp FP64-accumulated render of the same view for the
64-spp frame, 4 and 16 spp with the a-trous filter and 4 spp with the learned
denoiser, each with its ms/frame (render + G-buffer + filter).

Ray = one closest-hit scene query (camera, extension or shadow ray).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published
METRIC = "Mrays/s at 1080p; relMSE vs FP64 oracle at equal ms/frame (CNN vs à-trous)"

# algorithmic bytes per traversal unit: 48 B ray I/O (32 B ray in, 16 B hit out),
# 64 B per child-pair node, 48 B per triangle record, 64 B per instance record
B_RAY, B_NODE, B_TRI, B_INST = 48, 64, 48, 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traversal_bytes(st, shadow=False):
    if shadow:
        rays = st.shadow_rays
        return rays * B_RAY + st.shadow_node_visits * B_NODE + st.shadow_tri_tests * B_TRI + \
            st.shadow_instance_visits * B_INST
    rays = st.camera_rays + st.extension_rays
    return rays * B_RAY + st.node_visits * B_NODE + st.tri_tests * B_TRI + st.instance_visits * B_INST


def cpu_baseline(scene, cam, w, h, spp, depth, seed):
    """Pico's own CPU renderer (compiled reference subset) on the host cores,
    on a bounded sample of the same workload; rays counted by the oracle on the
    identical paths (same generators and seeds)."""
    import oracle
    import pico_amd as pa
    threads = min(16, os.cpu_count() or 1)
    p = pa.params(w, h, spp, depth, tile=64, frame_seed=seed)
    cs = oracle.CpuScene(scene)
    _, _, _, st = cs.render(cam, p, rng=oracle.RNG_REFERENCE, threads=threads)
    rays = st.camera_rays + st.extension_rays + st.shadow_rays
    if oracle.ref_available():
        rs = oracle.RefScene(scene)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            rs.render(cam, p, threads=threads)
            times.append(time.perf_counter() - t0)
        kind = "reference"
    else:
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            cs.render(cam, p, rng=oracle.RNG_REFERENCE, threads=threads)
            times.append(time.perf_counter() - t0)
        kind = "port"
    t = sorted(times)[1]
    return {"value": round(rays / t / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"same scene and camera at {w}x{h}, {spp} spp, depth {depth}, tiles 64x64; "
                      f"{rays} rays, median of 3 runs {t:.2f} s"}


def _median_ms(fn, reps, sync):
    times = []
    for _ in range(reps):
        sync()
        t0 = time.perf_counter()
        fn()
        sync()
        times.append((time.perf_counter() - t0) * 1e3)
    return sorted(times)[len(times) // 2]


def quality_leg(ctx, pa, scene, cam, W, H, depth, oracle_spp, weights, d_frame):
    """relMSE vs an FP64-accumulated oracle (frame seed 1000) and ms/frame of
    the a-trous and learned filters at 4 and 16 spp (frame seed 7)."""
    n = W * H
    d_sum = ctx.alloc(24 * n)
    d_sum.zero()
    t0 = time.time()
    chunk = 256
    for base in range(0, oracle_spp, chunk):
        p = pa.params(W, H, min(chunk, oracle_spp - base), depth, tile=16, frame_seed=1000)
        ctx.render_accum64_d(cam, p, base, d_sum)
        ctx.sync()
        log(f"[bench] oracle {base + chunk}/{oracle_spp} spp, {time.time() - t0:.1f} s")
    t_oracle = time.time() - t0
    out = {"oracle_spp": oracle_spp, "oracle_s": round(t_oracle, 1), "oracle_frame_seed": 1000,
           "frame_64spp_relmse": round(ctx.relmse_d(d_frame, d_sum, oracle_spp, n), 6)}
    d_pix, d_cnt, d_var = ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n)
    d_n, d_p, d_a = ctx.alloc(12 * n), ctx.alloc(12 * n), ctx.alloc(12 * n)
    d_out, d_scr = ctx.alloc(12 * n), ctx.alloc(12 * n)
    net = None
    if weights and os.path.exists(weights):
        net = pa.CNN(ctx, W, H, 1)
        net.load(weights)

    def frame(spp, filt):
        p = pa.params(W, H, spp, depth, tile=16, frame_seed=7)
        ctx.render_d(cam, p, d_pix, d_cnt, d_var)
        if filt is None:
            return d_pix
        ctx.gbuffer_d(cam, W, H, d_n, d_p, d_a)
        if filt == "atrous":
            ctx.atrous_fixed_d(d_pix, d_n, d_p, d_a, W, H, d_scr, d_out)
        elif filt == "atrous_ref":
            ctx.atrous_d(d_pix, d_n, d_p, d_a, W, H, d_scr, d_out)
        else:
            net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        return d_out

    ref = None
    rows = {}
    for name, spp, filt in (("raw_4spp", 4, None), ("raw_16spp", 16, None), ("atrous_4spp", 4, "atrous"),
                            ("atrous_16spp", 16, "atrous"), ("atrous_ref_4spp", 4, "atrous_ref"),
                            ("atrous_ref_16spp", 16, "atrous_ref"), ("cnn_4spp", 4, "cnn")):
        if filt == "cnn" and net is None:
            rows[name] = None
            continue
        frame(spp, filt)  # warm
        ms = _median_ms(lambda: frame(spp, filt), 5, ctx.sync)
        img = frame(spp, filt)
        if filt == "atrous_ref":
            # the reference's tile-local widths divide by zero on tiles with
            # no luminance or no depth range (NaN, as in the reference): the
            # error is taken over the finite pixels and the rest are counted
            if ref is None:
                ref = d_sum.download(np.float64, 3 * n).reshape(n, 3) / oracle_spp
            x = img.download(np.float32, 3 * n).reshape(n, 3).astype(np.float64)
            ok = np.isfinite(x).all(axis=1)
            e = (x[ok] - ref[ok]) ** 2 / (ref[ok] ** 2 + 1e-2)
            rows[name] = {"relmse_finite": round(float(e.mean()), 6), "nonfinite_pixels": int((~ok).sum()),
                          "ms_per_frame": round(ms, 2)}
        else:
            rows[name] = {"relmse": round(ctx.relmse_d(img, d_sum, oracle_spp, n), 6), "ms_per_frame": round(ms, 2)}
        log(f"[bench] {name}: {rows[name]}")
    out["atrous_fixed_sigmas"] = {"normal": pa.ATROUS_FIXED_SIGMAS[0], "depth_rel": pa.ATROUS_FIXED_SIGMAS[1],
                                  "albedo": pa.ATROUS_FIXED_SIGMAS[2], "colour": pa.ATROUS_FIXED_SIGMAS[3]}
    out.update(rows)
    if net is not None:
        # MFMA fraction of the network's forward at this size: logical FLOPs
        # (3x3 convolutions + 1x1 head) over the summed kernel time
        net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        ctx.sync()
        ctx.enable_kernel_timing(True)
        reps = 5
        for _ in range(reps):
            net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        ctx.sync()
        conv_n, conv_ms = ctx.kernel_time("cnn_conv_fwd")
        tot = 0.0
        for k in ("cnn_inputs", "cnn_pack", "cnn_conv_fwd", "cnn_pool", "cnn_up", "cnn_kpn_apply"):
            kn, kms = ctx.kernel_time(k)
            tot += kn * kms
        ctx.enable_kernel_timing(False)
        flops = pa.cnn_flops_per_pixel() * n
        conv_total_ms = conv_n * conv_ms / reps
        out["cnn_forward"] = {"flops_per_frame": flops, "conv_ms_per_frame": round(conv_total_ms, 3),
                              "kernels_ms_per_frame": round(tot / reps, 3),
                              "mfma": {"bound": "mfma", "achieved": round(flops / (conv_total_ms * 1e-3) / 1e12, 1),
                                       "peak": 2500.0, "unit": "TFLOP/s",
                                       "frac": round(flops / (conv_total_ms * 1e-3) / 1e12 / 2500.0, 4)}}
        out["cnn_weights"] = os.path.relpath(weights, ROOT)
    if out.get("cnn_4spp") and out.get("atrous_16spp"):
        c, a = out["cnn_4spp"], out["atrous_16spp"]
        out["cnn4_beats_atrous16"] = bool(c["relmse"] < a["relmse"] and c["ms_per_frame"] <= a["ms_per_frame"])
    for d in (d_sum, d_pix, d_cnt, d_var, d_n, d_p, d_a, d_out, d_scr):
        d.free()
    if net is not None:
        net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-count", action="store_true", help="skip the counting run (roofline bytes)")
    ap.add_argument("--cpu-spp", type=int, default=16, help="samples per pixel of the CPU baseline's sample")
    ap.add_argument("--oracle-spp", type=int, default=4096, help="samples of the FP64 quality oracle")
    ap.add_argument("--no-quality", action="store_true", help="skip the relMSE / ms-per-frame leg")
    ap.add_argument("--weights", default=None, help="trained denoiser parameters (.npy; default: the shipped ones)")
    ap.add_argument("--copy-mb", type=int, default=1024, help="size of the streaming-copy calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path would shard by "
              "image bands or by views, one scene replica per GPU)", file=sys.stderr)
        sys.exit(1)

    import pico_amd as pa
    ctx = pa.Context(0)
    log(f"[bench] device {ctx.device_name}")
    t0 = time.time()
    scene = pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)
    uniq, inst_tris = scene.n_triangles
    ctx.upload_scene(scene)
    info = ctx.scene_info()
    log(f"[bench] scene: {uniq} unique / {inst_tris} instanced triangles, upload+build {time.time() - t0:.1f}s, {info}")
    # streaming copy (16 B loads/stores) of a known size: the achievable HBM
    # bandwidth on this box, and the calibration point of the byte counters
    copy_bytes = args.copy_mb << 20
    src, dst = ctx.alloc(copy_bytes), ctx.alloc(copy_bytes)
    src.zero()
    ctx.enable_kernel_timing(True)
    for _ in range(5):
        ctx.copy_d(dst, src, copy_bytes)
    cn, cms = ctx.kernel_time("copy")
    ctx.enable_kernel_timing(False)
    copy_gbps = 2 * copy_bytes / (cms * 1e-3) / 1e9
    src.free()
    dst.free()
    log(f"[bench] copy {args.copy_mb} MiB: {cms:.3f} ms, {copy_gbps:.0f} GB/s (read + write)")
    W, H = args.width, args.height
    cam = scene.camera(W / H)
    n = W * H
    d_pix, d_cnt, d_var = ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n)
    p = pa.params(W, H, args.spp, args.depth, tile=16, frame_seed=1)

    count = None
    if not args.no_count:
        pc = pa.params(W, H, args.spp, args.depth, tile=16, frame_seed=1, flags=pa.RENDER_COUNT)
        count = ctx.render_d(cam, pc, d_pix, d_cnt, d_var)
        log(f"[bench] counting run: {count.camera_rays + count.extension_rays} extend rays, "
            f"{count.shadow_rays} shadow rays, {count.node_visits} node visits, {count.tri_tests} tri tests")
    for i in range(args.warmup):
        st = ctx.render_d(cam, p, d_pix, d_cnt, d_var)
        log(f"[bench] warmup {i}: {st.render_ms:.1f} ms, {st.accumulated_rays / st.render_ms / 1e3:.1f} Mrays/s")
    ctx.sync()
    ctx.enable_kernel_timing(True)
    rays = 0
    traced = 0
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.render_d(cam, p, d_pix, d_cnt, d_var)
        rays += st.accumulated_rays
        traced += st.rays
        log(f"[bench] step {i}: {st.render_ms:.1f} ms device")
    ctx.sync()
    dt = time.perf_counter() - t0
    ext_n, ext_ms = ctx.kernel_time("extend")
    sh_n, sh_ms = ctx.kernel_time("shadow")
    ctx.enable_kernel_timing(False)
    ms_step = dt * 1e3 / args.steps
    mrays = rays / dt / 1e6

    roof = None
    if count is not None and ext_n:
        launches_per_frame = ext_n / args.steps
        bytes_per_launch = traversal_bytes(count) / launches_per_frame
        achieved = bytes_per_launch / (ext_ms * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic_extend_instanced_1080p.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "measured_copy_GBps": round(copy_gbps, 1),
                "kernel": "extend", "launches": ext_n, "mean_launch_ms": round(ext_ms, 4),
                "algorithmic_bytes_per_launch": round(bytes_per_launch),
                "bytes_per_ray": round(traversal_bytes(count) / max(1, count.camera_rays + count.extension_rays), 1),
                "shadow_kernel_mean_launch_ms": round(sh_ms, 4),
                "shadow_achieved_GBps": round(traversal_bytes(count, True) / (sh_n / args.steps) / (sh_ms * 1e-3) / 1e9, 1)
                if sh_n else None}

    quality = None
    if not args.no_quality:
        quality = quality_leg(ctx, pa, scene, cam, W, H, args.depth, args.oracle_spp, args.weights or pa.weights_path(),
                              d_pix)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(scene, cam, W, H, args.cpu_spp, args.depth, 1)
        log(f"[bench] CPU baseline {cpu}")

    out = {"metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
           "config": {"workload": f"instanced scene, 1024 instances of 16 meshes (2k-20k tris), {inst_tris} instanced "
                                  f"triangles, {W}x{H}, {args.spp} spp, depth {args.depth}, GGX, wavefront path tracer",
                      "width": W, "height": H, "spp": args.spp, "max_depth": args.depth, "scene_seed": 2,
                      "unique_triangles": uniq, "instanced_triangles": inst_tris,
                      "rays_per_step": rays // args.steps, "traced_rays_per_step": traced // args.steps},
           "roofline": roof, "cpu_baseline": cpu, "quality": quality}
    if cpu:
        out["gpu_over_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
