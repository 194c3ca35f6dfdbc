#!/usr/bin/env python3
"""Benchmark of the rendering hot path on one MI355X.

One step = one frame of the instanced scene (1024 instances of 16 seeded
meshes with 2k-20k triangles, ~8.3M instanced triangles, two-level BVH,
3 quad lights + constant sky) at 1920x1080, 64 spp, max depth 8, rendered by
the wavefront path tracer into device-resident accumulation buffers (the
reference's running mean / variance with its early-out).  The scene is built
and uploaded before timing.

Printed: one JSON line with Mrays/s (rays of the accumulated samples / wall
time of the timed steps), the roofline of the traversal kernel (`trace`, the
extension rays of a bounce with the shadow rays of the bounce before it:
algorithmic bytes from a counting run over its HIP-event launch time, and its
HBM traffic from two rocprofv3 counter passes in this run), the reference CPU
renderer's Mrays/s on every logical CPU of the host, and the quality leg:
relMSE against a 4096-spp FP64-accumulated render of the same view (two
independent halves, whose mutual relMSE is the oracle's noise floor) for the
64-spp frame, 4 and 16 spp with the a-trous filters and 4 spp with the learned
denoiser, each with its ms/frame (render + G-buffer + filter), for GGX and for
the reference's Beckmann_All distribution.

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


def traversal_bytes(st):
    """Algorithmic bytes of every closest-hit query of a counting run: camera,
    extension and shadow rays."""
    rays = st.camera_rays + st.extension_rays + st.shadow_rays
    return rays * B_RAY + (st.node_visits + st.shadow_node_visits) * B_NODE + \
        (st.tri_tests + st.shadow_tri_tests) * B_TRI + (st.instance_visits + st.shadow_instance_visits) * B_INST


def finish_bytes(st):
    """The part of traversal_bytes traced by the kernel that ends a chunk's
    last bounces path by path (k_finish)."""
    return st.finish_rays * B_RAY + st.finish_node_visits * B_NODE + st.finish_tri_tests * B_TRI + \
        st.finish_instance_visits * B_INST


def cpu_baseline(scene, cam, w, h, spp, depth, seed):
    """Pico's own CPU renderer (compiled reference subset) on the host cores,
    on a bounded sample of the same workload; rays counted by the oracle on the
    identical paths (same generators and seeds)."""
    import oracle
    import pico_amd as pa
    # the reference's scene set-up prints progress on stdout; keep stdout for
    # the one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return _cpu_baseline(oracle, pa, scene, cam, w, h, spp, depth, seed)
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def host_cpu():
    """What the host offers this process: CPU model, sockets, physical cores,
    threads per core and logical CPUs from /proc/cpuinfo, the CPUs this
    process may run on, and the cgroup CPU quota."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            blocks = [b for b in f.read().split("\n\n") if b.strip()]
        procs = []
        for b in blocks:
            d = {}
            for line in b.splitlines():
                if ":" in line:
                    k, v = line.split(":", 1)
                    d[k.strip()] = v.strip()
            procs.append(d)
        info["model"] = procs[0].get("model name", "?") if procs else "?"
        cores = {(d.get("physical id", "0"), d.get("core id", str(i))) for i, d in enumerate(procs)}
        info["sockets"] = len({d.get("physical id", "0") for d in procs})
        info["physical_cores"] = len(cores)
        info["threads_per_core"] = round(len(procs) / max(1, len(cores)), 2)
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
        info["cgroup_cpu_quota"] = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    return info


def _cpu_baseline(oracle, pa, scene, cam, w, h, spp, depth, seed):
    # the reference's pool runs std::thread::hardware_concurrency() workers
    # (Source/Util/ThreadPool.hpp:31): every logical CPU of the host
    threads = os.cpu_count() or 1
    p = pa.params(w, h, spp, depth, tile=64, frame_seed=seed)
    cs = oracle.CpuScene(scene)
    _, _, _, st = cs.render(cam, p, rng=oracle.RNG_REFERENCE, threads=threads)
    rays = st.camera_rays + st.extension_rays + st.shadow_rays
    if oracle.ref_available():
        rs = oracle.RefScene(scene)
        run = lambda: rs.render(cam, p, threads=threads)  # noqa: E731
        kind = "reference"
    else:
        run = lambda: cs.render(cam, p, rng=oracle.RNG_REFERENCE, threads=threads)  # noqa: E731
        kind = "port"
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = sorted(times)[1]
    host = host_cpu()
    return {"value": round(rays / t / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"same scene and camera at {w}x{h}, {spp} spp, depth {depth}, tiles 64x64, {threads} threads "
                      f"(every logical CPU); {rays} rays, median of 3 runs {t:.2f} s "
                      f"(runs {min(times):.2f}-{max(times):.2f} s)",
            "host": host, "runs_s": [round(x, 3) for x in times],
            "mrays_per_s_per_thread": round(rays / t / 1e6 / threads, 3)}


def _median_ms(fn, reps, sync):
    times = []
    for _ in range(reps):
        sync()
        t0 = time.perf_counter()
        fn()
        sync()
        times.append((time.perf_counter() - t0) * 1e3)
    return sorted(times)[len(times) // 2]


def fp64_oracle(ctx, pa, cam, W, H, depth, spp, seed=1000, chunk=256):
    """Two independent FP64 accumulations of spp/2 samples each (samples
    [0, spp/2) and [spp/2, spp) of frame seed `seed`), and their sum: the
    spp-sample oracle.  Returns (d_sum, half_a, half_b) with the halves on the
    host as float64 [n, 3] means."""
    n = W * H
    half = spp // 2
    d_sum = ctx.alloc(24 * n)
    d_sum.zero()
    halves = []
    t0 = time.time()
    for h in range(2):
        for base in range(h * half, (h + 1) * half, chunk):
            p = pa.params(W, H, min(chunk, (h + 1) * half - base), depth, tile=16, frame_seed=seed)
            ctx.render_accum64_d(cam, p, base, d_sum)
        ctx.sync()
        halves.append(d_sum.download(np.float64, 3 * n).reshape(n, 3).copy())
        log(f"[bench] oracle half {h}: {half} spp, {time.time() - t0:.1f} s")
    a = halves[0] / half
    b = (halves[1] - halves[0]) / half
    return d_sum, a, b


def relmse_np(x, ref):
    return float(((x - ref) ** 2 / (ref ** 2 + 1e-2)).mean())


def quality_leg(ctx, pa, scene, cam, W, H, depth, oracle_spp, weights, d_frame=None):
    """relMSE vs an FP64-accumulated oracle (frame seed 1000) and ms/frame of
    the a-trous and learned filters at 4 and 16 spp (frame seed 7).  The
    oracle is two independent halves; their mutual relMSE is its noise floor
    (the full oracle's own error is about a quarter of it)."""
    n = W * H
    t0 = time.time()
    d_sum, half_a, half_b = fp64_oracle(ctx, pa, cam, W, H, depth, oracle_spp)
    t_oracle = time.time() - t0
    split = relmse_np(half_a, half_b)
    out = {"oracle_spp": oracle_spp, "oracle_s": round(t_oracle, 1), "oracle_frame_seed": 1000,
           "oracle_split_half_relmse": round(split, 6),
           "oracle_noise_floor_relmse": round(split / 4, 6),
           "oracle_split_half_worst_0.01pct_share": None}
    e = ((half_a - half_b) ** 2 / (half_b ** 2 + 1e-2)).sum(axis=1)
    k = max(1, n // 10000)
    out["oracle_split_half_worst_0.01pct_share"] = round(float(np.partition(e, n - k)[n - k:].sum() / max(e.sum(), 1e-300)), 4)
    del half_a, half_b
    if d_frame is not None:
        out["frame_relmse"] = round(ctx.relmse_d(d_frame, d_sum, oracle_spp, n), 6)
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
        if ref is None:
            ref = d_sum.download(np.float64, 3 * n).reshape(n, 3) / oracle_spp
        x = img.download(np.float32, 3 * n).reshape(n, 3).astype(np.float64)
        if filt == "atrous_ref":
            # the reference's tile-local widths divide by zero on tiles with
            # no luminance or no depth range (NaN, as in the reference): the
            # error is taken over the finite pixels and the rest are counted
            ok = np.isfinite(x).all(axis=1)
            e = (x[ok] - ref[ok]) ** 2 / (ref[ok] ** 2 + 1e-2)
            rows[name] = {"relmse_finite": round(float(e.mean()), 6), "nonfinite_pixels": int((~ok).sum()),
                          "ms_per_frame": round(ms, 2)}
        else:
            rows[name] = {"relmse": round(ctx.relmse_d(img, d_sum, oracle_spp, n), 6), "ms_per_frame": round(ms, 2)}
            # share of the error carried by the worst 0.01 % of pixels (fireflies)
            e = ((x - ref) ** 2 / (ref ** 2 + 1e-2)).sum(axis=1)
            k = max(1, n // 10000)
            rows[name]["worst_0.01pct_share"] = round(float(np.partition(e, n - k)[n - k:].sum() / max(e.sum(), 1e-300)), 4)
            rows[name]["max_value"] = round(float(x.max()), 2)
        log(f"[bench] {name}: {rows[name]}")
    out["atrous_fixed_sigmas"] = {"normal": pa.ATROUS_FIXED_SIGMAS[0], "depth_rel": pa.ATROUS_FIXED_SIGMAS[1],
                                  "albedo": pa.ATROUS_FIXED_SIGMAS[2], "colour": pa.ATROUS_FIXED_SIGMAS[3]}
    out.update(rows)
    if net is not None:
        # MFMA fraction of the network's forward at this size: logical FLOPs
        # (3x3 convolutions + 1x1 head) over the convolutions' time.  That time
        # is the GPU time of the forward (frames back to back, one stream,
        # wall clock over `wall_reps` frames) minus its input-packing kernel
        # (HIP events); per-launch HIP events on every convolution, which add
        # a dispatch gap to each, are reported beside it.
        net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        ctx.sync()
        ctx.enable_kernel_timing(True)
        reps = 5
        for _ in range(reps):
            net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        ctx.sync()
        conv_n, conv_ms = ctx.kernel_time("cnn_conv_fwd")
        pack_n, pack_ms = ctx.kernel_time("cnn_inputs_pack")
        tot = 0.0
        for k in ("cnn_inputs_pack", "cnn_conv_fwd"):
            kn, kms = ctx.kernel_time(k)
            tot += kn * kms
        ctx.enable_kernel_timing(False)
        wall_reps = 40
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(wall_reps):
            net.denoise_gbuffer_d(d_pix, d_n, d_p, d_a, W, H, d_out)
        ctx.sync()
        fwd_ms = (time.perf_counter() - t0) * 1e3 / wall_reps
        flops = pa.cnn_flops_per_pixel() * n
        conv_total_ms = fwd_ms - pack_n * pack_ms / reps
        out["cnn_forward"] = {"flops_per_frame": flops, "conv_ms_per_frame": round(conv_total_ms, 3),
                              "forward_ms_per_frame": round(fwd_ms, 3),
                              "conv_ms_per_frame_per_launch_events": round(conv_n * conv_ms / reps, 3),
                              "kernels_ms_per_frame_per_launch_events": round(tot / reps, 3),
                              "mfma": {"bound": "mfma", "achieved": round(flops / (conv_total_ms * 1e-3) / 1e12, 1),
                                       "peak": 2500.0, "unit": "TFLOP/s",
                                       "frac": round(flops / (conv_total_ms * 1e-3) / 1e12 / 2500.0, 4)}}
        out["cnn_weights"] = os.path.relpath(weights, ROOT)
    if out.get("cnn_4spp") and out.get("atrous_16spp"):
        c, a = out["cnn_4spp"], out["atrous_16spp"]
        out["cnn4_beats_atrous16"] = bool(c["relmse"] < a["relmse"] and c["ms_per_frame"] <= a["ms_per_frame"])
        out["cnn4_above_noise_floor"] = round(c["relmse"] / max(out["oracle_noise_floor_relmse"], 1e-300), 2)
    for d in (d_sum, d_pix, d_cnt, d_var, d_n, d_p, d_a, d_out, d_scr):
        d.free()
    if net is not None:
        net.close()
    return out


def look_at_box(rng, centre=50.0, r=(60.0, 110.0), jitter=15.0):
    """A camera outside the 100^3 box at a random direction, looking near its centre."""
    import pico_amd as pa
    d = rng.normal(size=3) * np.array([1.0, 0.5, 1.0])
    d /= np.linalg.norm(d)
    pos = centre + d * rng.uniform(*r)
    tgt = centre + rng.uniform(-jitter, jitter, 3)
    f = (tgt - pos) / np.linalg.norm(tgt - pos)
    return pa.Camera(tuple(pos), tuple(f), (0.0, 1.0, 0.0), 1.0, 10000.0)


def batched_leg(ctx, pa, depth, weights, n_views=64, seed=3, reps=3, oracle_spp=4096):
    """64 views of the scene with resolutions from {128^2, 256^2, 512^2,
    1024^2} and 4-64 spp in one wavefront launch (autotuned launch shape),
    each followed by the learned denoiser on a second stream."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([128, 256, 512, 1024], n_views)
    spps = rng.integers(4, 65, n_views)
    views, keep = [], []
    for i in range(n_views):
        w = h = int(sizes[i])
        b = (ctx.alloc(12 * w * h), ctx.alloc(4 * w * h), ctx.alloc(12 * w * h))
        v = pa.View()
        v.camera = look_at_box(rng)
        v.params = pa.params(w, h, int(spps[i]), depth, tile=16, frame_seed=11, view_id=i)
        v.d_pixels, v.d_sample_count, v.d_variance = b[0].ptr.value, b[1].ptr.value, b[2].ptr.value
        views.append(v)
        keep.append(b)
    t0 = time.time()
    saved = ctx.get_tuning()
    # calibrated on the batch itself: a scaled-down calibration cannot tell
    # queue depths apart (every depth gives one sample per pixel per chunk)
    tuning, cal_ms, def_ms = ctx.autotune_views(views, calib_spp=int(spps.max()))
    log(f"[bench] autotune {time.time() - t0:.1f} s: {tuning} ({cal_ms:.1f} ms vs default {def_ms:.1f} ms)")
    net = pa.CNN(ctx, 1024, 1024, 1)
    if weights and os.path.exists(weights):
        net.load(weights)
    outs = [ctx.alloc(12 * int(sizes[i]) ** 2) for i in range(n_views)]
    def timed(flags):
        ctx.set_tuning(tuning[0], tuning[1:3], flags)
        ctx.render_views_denoise_d(views, net, outs)  # warm
        runs = []
        for _ in range(reps):
            ctx.sync()
            t0 = time.perf_counter()
            st = ctx.render_views_denoise_d(views, net, outs)
            ctx.sync()
            runs.append((time.perf_counter() - t0, st))
        return sorted(runs, key=lambda r: r[0])[len(runs) // 2]

    # the denoiser between chunks on the launch's stream, or on a second
    # stream; and the views shortest first
    wall, st = timed(0)
    wall2, st2 = timed(pa.TUNE_SIDE_STREAM)
    wall3, st3 = timed(pa.TUNE_SHORTEST_FIRST)
    # the views one at a time, and the legs after this one, with the
    # context's own launch shape
    ctx.set_tuning(saved[0], saved[1:3], saved[3])
    lat3 = np.array([s.done_ms for s in st3])
    rays = sum(s.accumulated_rays for s in st)
    lat = np.array([s.done_ms for s in st])
    rlat = np.array([s.render_ms for s in st])
    lat2 = np.array([s.done_ms for s in st2])
    # the same views one at a time (render, G-buffer, network), for comparison
    d_n, d_p, d_a = (ctx.alloc(12 * 1024 * 1024) for _ in range(3))
    ctx.sync()
    t0 = time.perf_counter()
    for v, o in zip(views, outs):
        w, h = v.params.width, v.params.height
        ctx.render_d(v.camera, v.params, _ptr_array(ctx, v.d_pixels), _ptr_array(ctx, v.d_sample_count),
                     _ptr_array(ctx, v.d_variance))
        ctx.gbuffer_d(v.camera, w, h, d_n, d_p, d_a)
        net.denoise_gbuffer_d(_ptr_array(ctx, v.d_pixels), d_n, d_p, d_a, w, h, o)
    ctx.sync()
    seq = time.perf_counter() - t0
    # error of the smallest and the largest view against an FP64 oracle of
    # that view (oracle_spp samples, frame seed 1000): raw and denoised
    quality = {}
    for label, i in (("smallest", int(np.argmin(sizes))), ("largest", int(np.argmax(sizes)))):
        v = views[i]
        w, h = v.params.width, v.params.height
        d_sum, a, b = fp64_oracle(ctx, pa, v.camera, w, h, depth, oracle_spp, seed=1000)
        ref = (a + b) / 2
        raw = keep[i][0].download(np.float32, 3 * w * h).reshape(-1, 3).astype(np.float64)
        den = outs[i].download(np.float32, 3 * w * h).reshape(-1, 3).astype(np.float64)
        quality[label] = {"view": i, "size": w, "spp": v.params.spp, "oracle_spp": oracle_spp,
                          "relmse_raw": round(relmse_np(raw, ref), 6), "relmse_denoised": round(relmse_np(den, ref), 6),
                          "oracle_noise_floor_relmse": round(relmse_np(a, b) / 4, 6)}
        d_sum.free()
    net.close()
    for b in keep:
        for x in b:
            x.free()
    for x in outs + [d_n, d_p, d_a]:
        x.free()
    return {"workload": f"{n_views} views of the instanced scene, square resolutions from {{128, 256, 512, 1024}}, "
                        f"4-64 spp, depth {depth}, one launch + learned denoiser per view (view seed {seed})",
            "pixels": int((sizes.astype(np.int64) ** 2).sum()), "samples_requested": int((sizes.astype(np.int64) ** 2 * spps).sum()),
            "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s", "ms_per_batch": round(wall * 1e3, 2),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "render_latency_ms_p50": round(float(np.percentile(rlat, 50)), 2),
            "render_latency_ms_p99": round(float(np.percentile(rlat, 99)), 2),
            "shortest_first": {"value": round(sum(s.accumulated_rays for s in st3) / wall3 / 1e6, 2),
                               "ms_per_batch": round(wall3 * 1e3, 2),
                               "latency_ms_p50": round(float(np.percentile(lat3, 50)), 2),
                               "latency_ms_p99": round(float(np.percentile(lat3, 99)), 2)},
            "second_stream": {"ms_per_batch": round(wall2 * 1e3, 2),
                              "latency_ms_p50": round(float(np.percentile(lat2, 50)), 2),
                              "latency_ms_p99": round(float(np.percentile(lat2, 99)), 2)},
            "sequential_ms": round(seq * 1e3, 2), "view_relmse": quality,
            "tuning": {"queue_paths": tuning[0], "tile_w": tuning[1], "tile_h": tuning[2],
                       "calib_ms": round(cal_ms, 2), "default_calib_ms": round(def_ms, 2)}}


def _turned(xf, rng, angle, shift):
    """Column-major instance transforms, each turned about its own origin by a
    random rotation of up to `angle` radians and moved by up to `shift`."""
    n = xf.shape[0]
    m = xf.astype(np.float64).reshape(n, 4, 4).transpose(0, 2, 1)
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    a = rng.uniform(-angle, angle, n)[:, None, None]
    k = np.zeros((n, 3, 3))
    k[:, 0, 1], k[:, 0, 2], k[:, 1, 2] = -ax[:, 2], ax[:, 1], -ax[:, 0]
    k[:, 1, 0], k[:, 2, 0], k[:, 2, 1] = ax[:, 2], -ax[:, 1], ax[:, 0]
    r = np.eye(3) + np.sin(a) * k + (1 - np.cos(a)) * (k @ k)
    t = np.tile(np.eye(4), (n, 1, 1))
    t[:, :3, :3] = r
    c = m[:, :3, 3]
    t[:, :3, 3] = c - np.einsum("nij,nj->ni", r, c) + rng.uniform(-shift, shift, (n, 3))
    out = t @ m
    out[:, 3] = (0.0, 0.0, 0.0, 1.0)
    return np.ascontiguousarray(out.transpose(0, 2, 1).reshape(n, 16), dtype=np.float32)


def tlas_leg(ctx, pa, scene, cam, W, H, depth, d_pix, d_cnt, d_var, frames=8, spp=16):
    """Device TLAS rebuild: its time per frame of animation (every instance
    moved), the traversal cost of the linear-BVH TLAS against the host SAH
    TLAS on the same frame, and the rebuild against the host build at 65536
    instances."""
    import ctypes as C
    d = scene.desc
    n = d.n_instances
    base = np.ctypeslib.as_array(C.cast(d.transforms, C.POINTER(C.c_float)), (n * 16,)).copy().reshape(n, 16)
    p = pa.params(W, H, spp, depth, tile=16, frame_seed=5)
    host = ctx.render_d(cam, p, d_pix, d_cnt, d_var)
    img_host = d_pix.download(np.float32, 3 * W * H)
    st = ctx.set_instance_transforms(base)
    dev = ctx.render_d(cam, p, d_pix, d_cnt, d_var)
    img_dev = d_pix.download(np.float32, 3 * W * H)
    same = float(np.all(img_host.reshape(-1, 3) == img_dev.reshape(-1, 3), axis=1).mean())
    rng = np.random.default_rng(21)
    buf = ctx.alloc(64 * n)
    rebuild, render = [], []
    for f in range(frames + 2):
        buf.upload(_turned(base, rng, 0.05 * f, 0.5 * f))
        s = ctx.set_instance_transforms_d(buf)
        r = ctx.render_d(cam, p, d_pix, d_cnt, d_var)
        if f >= 2:
            rebuild.append(s.device_ms)
            render.append(r.render_ms)
    buf.free()
    ctx.upload_scene(scene)  # back to the host-built TLAS
    big = pa.Scene.instanced(3, 16, 65536, 20, 200, 400.0, pa.MICROFACET_GGX)
    t0 = time.perf_counter()
    ctx.upload_scene(big)
    host_ms = (time.perf_counter() - t0) * 1e3
    nb = big.desc.n_instances
    bx = np.ctypeslib.as_array(C.cast(big.desc.transforms, C.POINTER(C.c_float)), (nb * 16,)).copy().reshape(nb, 16)
    big_ms = []
    for f in range(6):
        s = ctx.set_instance_transforms(_turned(bx, rng, 0.5, 2.0))
        if f >= 2:
            big_ms.append(s.device_ms)
    ctx.upload_scene(scene)
    out = {"instances": n, "rebuild_ms": round(float(np.median(rebuild)), 3), "wide_nodes": st.wide_nodes,
           "binary_depth": st.binary_depth, "stack_bound": st.stack_bound,
           "render_ms_host_sah_tlas": round(host.render_ms, 2), "render_ms_device_tlas": round(dev.render_ms, 2),
           "identical_pixels": round(same, 5), "frame": f"{W}x{H}, {spp} spp",
           "animated_render_ms": round(float(np.median(render)), 2),
           "instances_65536": {"instances": nb, "rebuild_ms": round(float(np.median(big_ms)), 3),
                               "host_upload_ms": round(host_ms, 1),
                               "minstances_per_s": round(nb / (float(np.median(big_ms)) * 1e-3) / 1e6, 1)}}
    return out


def build_leg(ctx, pa, scene):
    """Scene build on the device against the host build: the bench scene
    (16 meshes, 2k-20k triangles each) and one mesh of 2M triangles; the
    device build's BVH kernels are event-timed, the host's is wall time.
    Both builders make the same trees (node counts checked here too)."""
    out = {}
    big = pa.Scene.instanced(5, 1, 4, 2000000, 2000000, 100.0, pa.MICROFACET_GGX)
    for name, sc in (("bench_scene", scene), ("mesh_2M", big)):
        h = ctx.upload_scene(sc)
        d = [ctx.upload_scene(sc, device_build=True) for _ in range(3)][-1]
        out[name] = {"triangles": sc.n_triangles[0], "host_bvh_ms": round(h.bvh_ms, 1),
                     "device_bvh_ms": round(d.bvh_ms, 2), "device_upload_total_ms": round(d.total_ms, 1),
                     "speedup": round(h.bvh_ms / d.bvh_ms, 1), "levels": d.levels,
                     "same_tree": (h.binary_nodes, h.blas_nodes4, h.blas_bound, h.levels) ==
                                  (d.binary_nodes, d.blas_nodes4, d.blas_bound, d.levels)}
    ctx.upload_scene(scene)
    return out


def ingest_leg(ctx, pa, tris=500000):
    """Scene ingest: a scene file in the reference's JSON format whose one mesh
    (about `tris` triangles, positions and normals) is a Wavefront OBJ file;
    the host load (JSON + OBJ parse, corner joining) and the device BVH build
    of the loaded scene are timed."""
    import ctypes as C
    import tempfile
    src = pa.Scene.instanced(11, 1, 1, tris, tris, 10.0, pa.MICROFACET_GGX)
    d = src.desc
    nv = int(np.ctypeslib.as_array(C.cast(d.mesh_vertex_offset, C.POINTER(C.c_uint32)), (2,))[1])
    nt = int(np.ctypeslib.as_array(C.cast(d.mesh_triangle_offset, C.POINTER(C.c_uint32)), (2,))[1])
    pos = np.ctypeslib.as_array(C.cast(d.positions, C.POINTER(C.c_float)), (3 * nv,)).reshape(-1, 3)
    nrm = np.ctypeslib.as_array(C.cast(d.normals, C.POINTER(C.c_float)), (3 * nv,)).reshape(-1, 3)
    idx = np.ctypeslib.as_array(C.cast(d.indices, C.POINTER(C.c_uint32)), (3 * nt,)).reshape(-1, 3) + 1
    with tempfile.TemporaryDirectory() as tmp:
        obj = os.path.join(tmp, "mesh.obj")
        with open(obj, "w") as f:
            f.write("".join(f"v {a:.9g} {b:.9g} {c:.9g}\n" for a, b, c in pos.tolist()))
            f.write("".join(f"vn {a:.9g} {b:.9g} {c:.9g}\n" for a, b, c in nrm.tolist()))
            f.write("".join(f"f {a}//{a} {b}//{b} {c}//{c}\n" for a, b, c in idx.tolist()))
        doc = {"MESH": {"Blob": {"Path": "mesh.obj"}, "Lamp": {"Path": "lamp.obj"}},
               "MATERIALS": {"Grey": {"Type": "Constant", "Albedo": [0.6, 0.6, 0.6], "Roughness": 0.4},
                             "Lamp": {"Type": "Constant", "Albedo": [1, 1, 1], "Emissive": [8, 8, 8]}},
               "INSTANCE": {"blob": {"Asset": "Blob", "BSRDF": "Dielectric", "Material": "Grey"},
                            "lamp": {"Asset": "Lamp", "BSRDF": "Light", "Material": "Lamp", "Position": [0, 30, 0],
                                     "Rotation": [1, 0, 0, 0], "Scale": [5, 1, 5]}},
               "CAMERA": {"main": {"Position": [5, 5, -25], "Direction": [0, 0, 1], "FarPlane": 100.0}}}
        with open(os.path.join(tmp, "lamp.obj"), "w") as f:
            f.write("v -1 0 -1\nv -1 0 1\nv 1 0 1\nv 1 0 -1\nf 1 2 3 4\n")
        path = os.path.join(tmp, "scene.json")
        with open(path, "w") as f:
            json.dump(doc, f)
        mb = os.path.getsize(obj) / 1e6
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            scene = pa.Scene.load_json(path)
            times.append((time.perf_counter() - t0) * 1e3)
    load_ms = sorted(times)[1]
    st = [ctx.upload_scene(scene, device_build=True) for _ in range(2)][-1]
    cam = scene.cameras()["main"]
    ld = _desc_counts(scene)
    out = {"obj_mb": round(mb, 1), "triangles": ld[1], "vertices": ld[0], "load_ms": round(load_ms, 1),
           "load_mb_per_s": round(mb / (load_ms * 1e-3), 1), "device_bvh_ms": round(st.bvh_ms, 2),
           "upload_total_ms": round(st.total_ms, 1)}
    p = pa.params(960, 540, 4, 8, frame_seed=1)
    out["first_frame_ms_960x540_4spp"] = round(ctx.render(cam, p)[3].render_ms, 2)
    return out


def _desc_counts(scene):
    import ctypes as C
    d = scene.desc
    vofs = np.ctypeslib.as_array(C.cast(d.mesh_vertex_offset, C.POINTER(C.c_uint32)), (d.n_meshes + 1,))
    tofs = np.ctypeslib.as_array(C.cast(d.mesh_triangle_offset, C.POINTER(C.c_uint32)), (d.n_meshes + 1,))
    return int(vofs[-1]), int(tofs[-1])


def traffic_child(args):
    """Child of the traffic leg, run under rocprofv3 --pmc: one streaming copy
    of --copy-mb MiB (the counters' calibration), then one frame of the bench
    workload.  Prints nothing on stdout."""
    import pico_amd as pa
    ctx = pa.Context(0)
    nbytes = args.copy_mb << 20
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    src.zero()
    ctx.copy_d(dst, src, nbytes)
    ctx.sync()
    scene = pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_GGX)
    ctx.upload_scene(scene)
    W, H = args.width, args.height
    n = W * H
    d_pix, d_cnt, d_var = ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n)
    ctx.render_d(scene.camera(W / H), pa.params(W, H, args.spp, args.depth, tile=16, frame_seed=1), d_pix, d_cnt, d_var)
    ctx.sync()
    ctx.close()


def _pmc_pass(counter, args):
    """One rocprofv3 counter pass over the traffic child: the counter's value
    per dispatch, averaged per kernel name."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        raise RuntimeError("rocprofv3 not found")
    tmp = tempfile.mkdtemp(prefix="pico_pmc_", dir="/tmp")
    try:
        cmd = [prof, "--pmc", counter, "--kernel-include-regex", "k_trace|k_copy", "--output-format", "csv",
               "-d", tmp, "-o", "run", "--", sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-child",
               "--width", str(args.width), "--height", str(args.height), "--spp", str(args.spp), "--depth",
               str(args.depth), "--copy-mb", str(args.copy_mb)]
        r = subprocess.run(cmd, cwd=tmp, env=dict(os.environ, TMPDIR="/tmp"), capture_output=True, text=True,
                           timeout=240)
        files = glob.glob(os.path.join(tmp, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise RuntimeError(f"rocprofv3 exit {r.returncode}, no counter file: {r.stderr[-400:]}")
        per = {}
        seen = set()
        for fn in files:
            with open(fn, newline="") as f:
                for row in csv.DictReader(f):
                    seen.add((row.get("Kernel_Name", "?")[:60], row.get("Counter_Name")))
                    if row.get("Counter_Name") != counter:
                        continue
                    key = (row["Kernel_Name"], row.get("Dispatch_Id", ""))
                    per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
        out = {}
        for (name, _), v in per.items():
            for short in ("k_trace", "k_copy"):
                if "::" + short + "<" in name or "::" + short + "(" in name:
                    out.setdefault(short, []).append(v)
        if "k_copy" not in out or "k_trace" not in out:
            raise RuntimeError(f"{counter}: dispatches of k_copy and k_trace expected, found {sorted(seen)[:8]}")
        return {k: (len(v), sum(v) / len(v)) for k, v in out.items()}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def traffic_leg(args):
    """HBM bytes per dispatch of the traversal kernel, measured in this run:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) from two rocprofv3 counter
    passes over one frame, each scaled by what the same counter reads for the
    streaming copy of known size."""
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        log(f"[bench] traffic: rocprofv3 --pmc {counter} over one frame ...")
        res[counter] = _pmc_pass(counter, args)
    copy_kib = float(args.copy_mb << 10)
    f_scale = copy_kib / res["FETCH_SIZE"]["k_copy"][1]
    w_scale = copy_kib / res["WRITE_SIZE"]["k_copy"][1]
    out = {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
           "calibration": f"k_copy of {args.copy_mb} MiB read + written: FETCH_SIZE "
                          f"{res['FETCH_SIZE']['k_copy'][1]:.0f} KiB, WRITE_SIZE {res['WRITE_SIZE']['k_copy'][1]:.0f} KiB"}
    for k in ("k_trace",):
        if k in res["FETCH_SIZE"] and k in res["WRITE_SIZE"]:
            fk, wk = res["FETCH_SIZE"][k][1], res["WRITE_SIZE"][k][1]
            nd = res["FETCH_SIZE"][k][0]
            per = fk * 1024 * f_scale + wk * 1024 * w_scale
            # the child's frame is its context's first render, which launches
            # every bounce (those after the hand-over return at once and move
            # nothing): the frame's total is what the timed frames' fewer
            # launches move
            out[k] = {"dispatches": nd, "fetch_kib_per_dispatch": round(fk, 1),
                      "write_kib_per_dispatch": round(wk, 1), "hbm_bytes_per_dispatch": round(per),
                      "hbm_bytes_per_frame": round(per * nd)}
    return out


def _ptr_array(ctx, ptr):
    """A non-owning device array at a raw pointer (the views' own buffers)."""
    import pico_amd as pa
    import ctypes as C
    a = pa.DeviceArray.__new__(pa.DeviceArray)
    a.ctx, a.nbytes, a._owner, a.ptr = ctx, 0, True, C.c_void_p(ptr)
    return a


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
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-views leg")
    ap.add_argument("--no-tlas", action="store_true", help="skip the device TLAS rebuild and scene build legs")
    ap.add_argument("--weights", default=None, help="trained denoiser parameters (.npy; default: the shipped ones)")
    ap.add_argument("--copy-mb", type=int, default=1024, help="size of the streaming-copy calibration")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes (roofline traffic)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(args)
        return
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
    tr_n, tr_ms = ctx.kernel_time("trace")
    fin_n, fin_ms = ctx.kernel_time("finish")
    ctx.enable_kernel_timing(False)
    ms_step = dt * 1e3 / args.steps
    mrays = rays / dt / 1e6

    roof = None
    if count is not None and tr_n:
        launches_per_frame = tr_n / args.steps
        # k_trace's rays: all but those of the last bounces' hand-over kernel
        trace_bytes = traversal_bytes(count) - finish_bytes(count)
        trace_rays = count.camera_rays + count.extension_rays + count.shadow_rays - count.finish_rays
        bytes_per_launch = trace_bytes / launches_per_frame
        achieved = bytes_per_launch / (tr_ms * 1e-3) / 1e9
        traffic, tdetail = None, None
        if not args.no_traffic:
            try:
                tdetail = traffic_leg(args)
                traffic = round(tdetail["k_trace"]["hbm_bytes_per_frame"] / launches_per_frame)
                log(f"[bench] traffic {tdetail}")
            except Exception as e:  # no profiler on this host: no traffic figure
                tdetail = {"error": str(e)[:300]}
                log(f"[bench] traffic leg failed: {e}")
        real = traffic / (tr_ms * 1e-3) / 1e9 if traffic else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic, "traffic_detail": tdetail,
                "traffic_GBps": round(real, 1) if real else None,
                "traffic_frac": round(real / HBM_PEAK_GBPS, 4) if real else None,
                "note": "achieved counts the logical bytes of every node, triangle and instance record a ray "
                        "visits; most are served from L2 and the Infinity Cache, so it can exceed the HBM peak. "
                        "traffic is the measured HBM bytes of a frame's k_trace launches per launch, "
                        "traffic_GBps their rate",
                "measured_copy_GBps": round(copy_gbps, 1),
                "kernel": "trace", "launches": tr_n, "mean_launch_ms": round(tr_ms, 4),
                "algorithmic_bytes_per_launch": round(bytes_per_launch),
                "bytes_per_ray": round(trace_bytes / max(1, trace_rays), 1),
                "rays_per_frame": trace_rays,
                "stack_spill_pushes_per_frame": count.spill_pushes, "stack_spill_pops_per_frame": count.spill_pops,
                # the last bounces of each chunk, path by path in one launch
                # (traversal and shading): rays, algorithmic traversal bytes
                # and launch time per frame
                "finish_kernel": {"rays_per_frame": count.finish_rays,
                                  "traversal_bytes_per_frame": round(finish_bytes(count)),
                                  "launches_per_frame": round(fin_n / args.steps, 2),
                                  "ms_per_frame": round(fin_ms * fin_n / args.steps, 3)}}

    quality = None
    if not args.no_quality:
        quality = quality_leg(ctx, pa, scene, cam, W, H, args.depth, args.oracle_spp, args.weights or pa.weights_path(),
                              d_pix)
        # the same rows with the reference's own distribution, Beckmann_All,
        # whose clamped D bounds the next-event weight
        beck = pa.Scene.instanced(2, 16, 1024, 2000, 20000, 100.0, pa.MICROFACET_BECKMANN_ALL)
        ctx.upload_scene(beck)
        quality["beckmann_all"] = quality_leg(ctx, pa, beck, cam, W, H, args.depth, args.oracle_spp,
                                              args.weights or pa.weights_path())
        ctx.upload_scene(scene)

    batched = None
    if not args.no_batched:
        batched = batched_leg(ctx, pa, args.depth, args.weights or pa.weights_path(), oracle_spp=args.oracle_spp)
        log(f"[bench] batched views: {batched}")

    tlas = build = None
    if not args.no_tlas:
        tlas = tlas_leg(ctx, pa, scene, cam, W, H, args.depth, d_pix, d_cnt, d_var)
        log(f"[bench] device TLAS rebuild: {tlas}")
        build = build_leg(ctx, pa, scene)
        log(f"[bench] device scene build: {build}")
        build["json_ingest"] = ingest_leg(ctx, pa)
        log(f"[bench] JSON ingest: {build['json_ingest']}")
        ctx.upload_scene(scene)

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
           "roofline": roof, "cpu_baseline": cpu, "quality": quality,
           "batched_views": batched, "tlas_rebuild": tlas, "scene_build": build}
    if cpu:
        out["gpu_over_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
