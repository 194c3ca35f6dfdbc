"""Training of the learned denoiser on images the path tracer renders itself.

A training pair is one seeded procedural scene (``Scene.training(seed)``: a
jittered Cornell box, a small instanced cluster or a 100^3 instanced field) at
256x256:

* input: the 4-spp render with the reference's accumulation (frame seed
  2*seed + 1) and its G-buffer (albedo, normal, distance), 10 channels;
* target: a 1024-spp render accumulated in FP64 (frame seed 2*seed + 2).

Everything stays in device memory: the pairs are rendered straight into the
dataset arrays, batches are gathered with device copies, and each step is one
forward, backward and Adam update of the network's own HIP kernels.

Run: ``python -m pico_amd.kpn_train --out weights.npy`` (options below).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

import pico_amd as pa


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class PairRenderer:
    """Renders (input, target) pairs of one size into device memory."""

    def __init__(self, ctx, w, h, spp_in=4, spp_target=1024, depth=8):
        self.ctx, self.w, self.h = ctx, w, h
        self.spp_in, self.spp_target, self.depth = spp_in, spp_target, depth
        n = w * h
        self.d_pix, self.d_cnt, self.d_var = ctx.alloc(12 * n), ctx.alloc(4 * n), ctx.alloc(12 * n)
        self.d_n, self.d_p, self.d_a = ctx.alloc(12 * n), ctx.alloc(12 * n), ctx.alloc(12 * n)
        self.d_sum = ctx.alloc(24 * n)
        self.rays = 0

    def render(self, scene, seed, d_x, d_t):
        """Input planes into d_x (10 x H x W f32), target RGB into d_t (H*W x 3 f32)."""
        ctx, w, h = self.ctx, self.w, self.h
        ctx.upload_scene(scene)
        cam = scene.camera(w / h)
        p = pa.params(w, h, self.spp_in, self.depth, tile=16, frame_seed=2 * seed + 1)
        st = ctx.render_d(cam, p, self.d_pix, self.d_cnt, self.d_var)
        self.rays += st.rays
        ctx.gbuffer_d(cam, w, h, self.d_n, self.d_p, self.d_a)
        ctx.cnn_inputs_d(self.d_pix, self.d_n, self.d_p, self.d_a, w, h, d_x)
        self.d_sum.zero()
        pt = pa.params(w, h, self.spp_target, self.depth, tile=16, frame_seed=2 * seed + 2)
        st = ctx.render_accum64_d(cam, pt, 0, self.d_sum)
        self.rays += st.rays
        s = self.d_sum.download(np.float64, 3 * w * h)
        d_t.upload((s / self.spp_target).astype(np.float32))


def relmse(x, r):
    x, r = np.asarray(x, np.float64), np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def lr_at(step, steps, lr, warm=100):
    if step < warm:
        return lr * (step + 1) / warm
    return 1e-2 * lr + 0.5 * (lr - 1e-2 * lr) * (1 + math.cos(math.pi * (step - warm) / max(1, steps - warm)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--seed0", type=int, default=10000)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--spp-in", type=int, default=4)
    ap.add_argument("--spp-target", type=int, default=1024)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--clip", type=float, default=1.0, help="global gradient-norm clip (0: off)")
    ap.add_argument("--val", type=int, default=48, help="held-out pairs (the last seeds)")
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--out", required=True, help=".npy file for the trained parameters")
    ap.add_argument("--log", default=None, help="JSON training log")
    args = ap.parse_args(argv)

    ctx = pa.Context(0)
    S, N = args.size, args.scenes
    px = S * S
    xs, ts = 40 * px, 12 * px
    X, T = ctx.alloc(N * xs), ctx.alloc(N * ts)
    pr = PairRenderer(ctx, S, S, args.spp_in, args.spp_target, args.depth)
    t0 = time.time()
    for k in range(N):
        scene = pa.Scene.training(args.seed0 + k, pa.MICROFACET_GGX)
        pr.render(scene, args.seed0 + k, X.view(k * xs, xs), T.view(k * ts, ts))
        del scene
        if (k + 1) % 100 == 0 or k + 1 == N:
            ctx.sync()
            log(f"[data] {k + 1}/{N} pairs, {time.time() - t0:.1f} s, {pr.rays / 1e9:.2f} G rays")
    t_data = time.time() - t0

    net = pa.CNN(ctx, S, S, args.batch)
    net.init(args.init_seed)
    net.set_clip(args.clip)
    bx, bt = ctx.alloc(args.batch * xs), ctx.alloc(args.batch * ts)
    ntrain = N - args.val
    rng = np.random.default_rng(args.init_seed)
    history = []
    t1 = time.time()
    for step in range(args.steps):
        idx = rng.choice(ntrain, args.batch, replace=False)
        for j, i in enumerate(idx):
            ctx.copy_d(bx.view(j * xs, xs), X.view(int(i) * xs, xs), xs)
            ctx.copy_d(bt.view(j * ts, ts), T.view(int(i) * ts, ts), ts)
        want = step % 50 == 0 or step == args.steps - 1
        loss = net.train_step_d(bx, bt, args.batch, S, S, lr_at(step, args.steps, args.lr), want_loss=want)
        if want:
            gn = net.last_grad_norm()
            history.append((step, loss, gn))
            if step % 250 == 0 or step == args.steps - 1:
                log(f"[train] step {step}: loss {loss:.5f}, grad norm {gn:.4g}, {time.time() - t1:.1f} s")
    ctx.sync()
    t_train = time.time() - t1
    net.save(args.out)

    # held-out pairs: raw 4 spp, the two a-trous modes and the network
    # against the target (the reference-mode filter's NaN pixels, from its
    # tile-local widths, are left out of its error and counted)
    res = {"raw": [], "atrous_fixed": [], "atrous_ref_finite": [], "atrous_ref_nonfinite_frac": [], "cnn": []}
    d_out, d_scr = ctx.alloc(ts), ctx.alloc(ts)
    for k in range(ntrain, N):
        x = X.view(k * xs, xs)
        tgt = T.view(k * ts, ts).download(np.float32, 3 * px).reshape(px, 3)
        xin = x.download(np.float32, 10 * px).reshape(10, px)
        rgb = np.ascontiguousarray(xin[0:3].T)
        res["raw"].append(relmse(rgb, tgt))
        net.denoise_d(x, S, S, d_out)
        res["cnn"].append(relmse(d_out.download(np.float32, 3 * px).reshape(px, 3), tgt))
        scene = pa.Scene.training(args.seed0 + k, pa.MICROFACET_GGX)
        ctx.upload_scene(scene)
        cam = scene.camera(1.0)
        ctx.gbuffer_d(cam, S, S, pr.d_n, pr.d_p, pr.d_a)
        pr.d_pix.upload(rgb)
        ctx.atrous_fixed_d(pr.d_pix, pr.d_n, pr.d_p, pr.d_a, S, S, d_scr, d_out)
        res["atrous_fixed"].append(relmse(d_out.download(np.float32, 3 * px).reshape(px, 3), tgt))
        ctx.atrous_d(pr.d_pix, pr.d_n, pr.d_p, pr.d_a, S, S, d_scr, d_out)
        ar = d_out.download(np.float32, 3 * px).reshape(px, 3)
        ok = np.isfinite(ar).all(axis=1)
        res["atrous_ref_finite"].append(relmse(ar[ok], tgt[ok]) if ok.any() else float("nan"))
        res["atrous_ref_nonfinite_frac"].append(float((~ok).mean()))
        del scene
    summary = {k: float(np.nanmean(v)) for k, v in res.items()}
    out = {"scenes": N, "val": args.val, "size": S, "spp_in": args.spp_in, "spp_target": args.spp_target,
           "steps": args.steps, "batch": args.batch, "lr": args.lr, "data_s": round(t_data, 1),
           "train_s": round(t_train, 1), "ms_per_step": round(1e3 * t_train / max(1, args.steps), 2),
           "val_relmse": summary, "history": history}
    log(f"[val] relMSE over {args.val} held-out pairs: {summary}")
    if args.log:
        with open(args.log, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "history"}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
