#!/usr/bin/env python3
"""Throughput of the gfx950 path tracer on the 1080p / 1M-triangle forward workload.

One step = one full frame: 1920x1080 pixels x 64 samples, 8 bounces, on a 1,048,576-triangle
fBm-displaced sphere plus ground under a 2048x1024 log-normal environment map (CDF + MIS). The
scene is uploaded once; each timed step renders into HBM-resident AOVs (no host copies inside the
timed region).

Prints one JSON line: Msamples/s, the dominant kernel's roofline (the trace kernel: closest hit of
a segment's rays plus the previous segment's shadow rays; algorithmic bytes per launch from the
reference's traversal counters over its average launch time, measured with HIP events on the render
stream), the CPU oracle's throughput on a bounded row band of the same frame, and, after the timed
frames, the metric's second half: "denoise" (relMSE against a 4096-spp render of the learned
denoiser at 1 spp and the bilateral filter at 4 spp, and the plain tracer's time to the same
relMSE), "fog" (the 128^3 grid-fog frame at 4 spp: throughput and the 7x7 bilateral filter's
relMSE) and "adjoint" (forward and backward time at 512 x 512, 16 spp, 100k triangles).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes of the trace kernel per unit of work (per-unit figures of the reference's
# counters, bvh.cs:188-203): 32 B per child-box test (a 64-byte sibling pair per interior visit),
# 40 B per triangle test (TriangleLean 36 B + item index 4 B), 48 B per closest-hit ray (queue
# index 4 B, origin + direction 24 B, hit record 20 B), 104 B per shadow ray (80 B deferred record,
# read-modify-write of the path's 12 B radiance).
BYTES_PER_BOX, BYTES_PER_TRI, BYTES_PER_RAY, BYTES_PER_SHADOW = 32, 40, 48, 104


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def processor_count():
    """The CPUs this process may run on, as .NET's Environment.ProcessorCount counts them (the
    reference's worker count, renderer.cs:88): the affinity mask, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(period)
    except (OSError, ValueError):
        try:
            with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
                q = int(f.read())
            with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                period = int(f.read())
            if q > 0:
                quota = q / period
        except (OSError, ValueError):
            pass
    if quota is not None:
        n = max(1, min(n, int(-(-quota // 1))))
    return n, quota


def host_info():
    """nproc, os.cpu_count(), the cgroup quota and the CPU model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    n, quota = processor_count()
    return {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(), "cgroup_cpu_quota": quota,
            "processor_count": n, "cpu_model": model}


def cpu_baseline(spec, seconds_target=15.0):
    """Times the CPU oracle (the C++ restatement with the reference's 32x32-block scheduler, built
    here with -O3 -march=native -ffp-contract=off) on a centred row band of the same frame, sized
    to about `seconds_target` seconds, on as many threads as .NET's ProcessorCount would give the
    reference's renderer on this host."""
    import tempfile
    from oracle import oracle as O
    host = host_info()
    threads = host["processor_count"]
    with tempfile.TemporaryDirectory() as tmp:
        O.build_native(tmp)
        t0 = time.perf_counter()
        osc = O.OracleScene(spec, precision="f32_native")
        setup = time.perf_counter() - t0
        mid = spec.height // 2
        probe_rows = 2
        t0 = time.perf_counter()
        osc.render(mid, mid + probe_rows, threads=threads)
        dt = max(time.perf_counter() - t0, 1e-3)
        rows = int(max(probe_rows, min(spec.height, probe_rows * seconds_target / dt)))
        y0 = max(0, mid - rows // 2)
        y1 = min(spec.height, y0 + rows)
        t0 = time.perf_counter()
        out = osc.render(y0, y1, threads=threads)
        dt = time.perf_counter() - t0
        del osc
    samples = float(out["samples"].astype("int64").sum() - (y1 - y0) * spec.width)  # count paths traced
    msps = samples / dt / 1e6
    return {"value": round(msps, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"rows {y0}-{y1} of the {spec.width}x{spec.height} frame at {spec.max_samples} spp "
                      f"({int(samples)} samples, {dt:.1f} s, scene build {setup:.1f} s; oracle built -O3 "
                      f"-march=native -ffp-contract=off; {threads} threads)",
            "host": host}


def bench_adjoint(args, spec, scene, view):
    """Inverse-rendering step: forward render at perturbed parameters (materials +-30 %, sky texels
    +-30 %) and the backward pass of 0.5 * |image - target|^2, the target being rendered at the
    scene's own parameters. value = backward paths per second."""
    import numpy as np
    from sharpray_amd import Adjoint
    adj = Adjoint(scene)
    fw = (view, spec.width, spec.height, spec.max_samples, spec.bounces, spec.indirect_clamp)
    p_true = adj.params()
    target = adj.forward(*fw, seed=1)
    rng = np.random.default_rng(0)
    p0 = (p_true * rng.uniform(0.7, 1.3, size=p_true.shape)).astype(np.float32)
    n_obj = len(spec.objects)
    mat = p0[:n_obj * 4].reshape(n_obj, 4)
    mat[:, :3] = np.clip(mat[:, :3], 0.0, 1.0)
    mat[:, 3] = np.clip(mat[:, 3], 0.05, 1.0)
    adj.set_params(p0)
    for i in range(args.warmup):
        img = adj.forward(*fw, seed=2)
        adj.backward(img - target)
    tf, tb = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        img = adj.forward(*fw, seed=2)
        t1 = time.perf_counter()
        g = adj.backward(img - target)
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
        log(f"[bench] adjoint step {i}: forward {1e3 * tf[-1]:.1f} ms, backward {1e3 * tb[-1]:.1f} ms, "
            f"loss {0.5 * float(((img - target) ** 2).sum()):.4e}, |grad| {float(np.linalg.norm(g)):.4e}")
    paths = spec.width * spec.height * spec.max_samples
    mb = sum(tb) / len(tb)
    mf = sum(tf) / len(tf)
    line = {
        "metric": "adjoint backward Mpaths/s @512px/100k tris",
        "value": round(paths / mb / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (mb + mf), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": spec.name, "width": spec.width, "height": spec.height, "spp": spec.max_samples,
                   "bounces": spec.bounces, "triangles": spec.triangle_count, "params": adj.param_count},
        "forward_ms": round(1e3 * mf, 3),
        "backward_ms": round(1e3 * mb, 3),
        "backward_over_forward": round(mb / mf, 3),
    }
    print(json.dumps(line), flush=True)


def bench_cnn(args, spec, scene, view):
    """The kernel-predicting denoiser (shipped trained weights) on device-resident AOVs of one
    render. value = frames per second."""
    from sharpray_amd import KpnDenoiser, Renderer
    r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                 spec.variance_threshold, spec.bounces, spec.indirect_clamp)
    r.wait()
    den = KpnDenoiser(scene)
    den.load(KpnDenoiser.pretrained_weights())
    den.benchmark(r, max(1, args.warmup))
    ms = den.benchmark(r, max(1, args.steps))
    flops = KpnDenoiser.flops(spec.width, spec.height)
    tflops = flops / (ms * 1e-3) / 1e12
    line = {
        "metric": "kpn denoiser frames/s @1080p",
        "value": round(1e3 / ms, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded procedural scene, shipped trained weights)",
        "config": {"workload": "kpn_unet_1080p", "width": spec.width, "height": spec.height,
                   "input_spp": spec.max_samples, "gflop_per_frame": round(flops / 1e9, 2)},
        "roofline": {"kernel": "k_conv3x3 (bf16 MFMA implicit GEMM)", "bound": "mfma", "achieved": round(tflops, 2),
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4), "traffic": None},
    }
    print(json.dumps(line), flush=True)


def _relmse_fn(ref, H, W):
    import numpy as np

    def rel(x):
        x = np.asarray(x, np.float64).reshape(H, W, 3)
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))
    return rel


def denoise_summary(spec, scene, view, ref_spp, bilateral_radius=0.000694):
    """Equal-quality / equal-time comparison at the spec's size: the plain tracer's relMSE =
    mean (x - r)^2 / (r^2 + 0.01) against a ref_spp render (counter RNG, other seed) as a function
    of its samples, the learned denoiser (shipped weights) and the bilateral filter (7x7 window at
    1080p, strength chosen by grid search against the reference, which favours it; on the per-pixel
    sample mean and on the Radiance AOV) on 1 and 4 spp renders, and the time the plain tracer needs
    to reach each denoiser's relMSE (log-log interpolation of its relMSE-time curve)."""
    import math
    import numpy as np
    from sharpray_amd import RNG_COUNTER, Compositor, KpnDenoiser, Renderer
    W, H = spec.width, spec.height

    def render(spp, seed):
        t0 = time.perf_counter()
        r = Renderer(scene, view, W, H, 32, spp, spp, 0.0, spec.bounces, spec.indirect_clamp,
                     rng_policy=RNG_COUNTER, rng_seed=seed)
        r.wait()
        return r, time.perf_counter() - t0

    render(1, 3)  # warm-up
    t0 = time.perf_counter()
    ref = render(ref_spp, 999)[0].sample_mean.reshape(H, W, 3).astype(np.float64)
    log(f"[bench] reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    rel = _relmse_fn(ref, H, W)
    # The plain tracer's per-pixel sample mean (its Radiance AOV divides pixels at the sample cap by
    # M + 1, like the reference; relmse_radiance_aov is that image's error).
    plain = []
    for spp in (1, 2, 4, 8, 16, 32, 64):
        r, t = render(spp, 7)
        plain.append({"spp": spp, "ms": round(1e3 * t, 2), "relmse": round(rel(r.sample_mean), 6),
                      "relmse_radiance_aov": round(rel(r.radiance), 6)})
        log(f"[bench] plain {spp} spp: {1e3 * t:.1f} ms, relMSE {plain[-1]['relmse']:.5f}")

    def matched_ms(e):
        """Plain-tracer time at relMSE e (log-log interpolation; 1/spp extrapolation past the ends)."""
        pts = [(math.log(p["ms"]), math.log(p["relmse"])) for p in plain]
        for (t0_, e0), (t1_, e1) in zip(pts, pts[1:]):
            if e1 <= math.log(e) <= e0:
                f = (math.log(e) - e0) / (e1 - e0) if e1 != e0 else 0.0
                return math.exp(t0_ + f * (t1_ - t0_))
        tl, el = pts[-1] if math.log(e) < pts[-1][1] else pts[0]
        return math.exp(tl + (el - math.log(e)))

    den = KpnDenoiser(scene)
    den.load(KpnDenoiser.pretrained_weights())
    rows = {}
    for spp in (1, 4):
        r, t = render(spp, 7)
        kpn = den.denoise(r)
        ms_den = den.benchmark(r, 10)
        rows[f"kpn_{spp}spp"] = {"relmse": round(rel(kpn), 6), "ms": round(1e3 * t + ms_den, 2),
                                 "denoise_ms": round(ms_den, 3)}
        for tag, mean in (("", True), ("_radiance_aov", False)):
            best = None
            for strength in (0.05, 0.2, 1.0, 5.0, 25.0):
                comp = Compositor(bilateral_radius, strength, 1.75, 0.25, 2.0, 0.125, 0.2)
                tf = time.perf_counter()
                out = comp.filter(r, sample_mean=mean)
                tf = time.perf_counter() - tf
                e = rel(out)
                if best is None or e < best[0]:
                    best = (e, tf, strength)
            rows[f"bilateral_{spp}spp{tag}"] = {"relmse": round(best[0], 6), "ms": round(1e3 * (t + best[1]), 2),
                                                "strength": best[2]}
    for k, v in rows.items():
        v["plain_ms_at_same_relmse"] = round(matched_ms(v["relmse"]), 2)
        v["speedup_vs_plain"] = round(v["plain_ms_at_same_relmse"] / v["ms"], 3)
        log(f"[bench] {k}: {v}")
    return plain, rows


def bench_denoise(args, spec, scene, view):
    """The learned denoiser's relMSE-matched speedup over the plain tracer (denoise_summary).
    value = the plain tracer's time to the relMSE of the learned denoiser at 1 spp divided by the
    denoiser's own time (render + denoise)."""
    plain, rows = denoise_summary(spec, scene, view, args.ref_spp)
    line = {
        "metric": "relMSE-matched time: learned denoiser vs plain @1080p/1M tris",
        "value": rows["kpn_1spp"]["speedup_vs_plain"],
        "unit": "x (plain-tracer time to the same relMSE / render + denoise time)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": rows["kpn_1spp"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded procedural scene, shipped trained weights)",
        "config": {"workload": spec.name, "width": spec.width, "height": spec.height, "bounces": spec.bounces,
                   "triangles": spec.triangle_count, "reference_spp": args.ref_spp, "relmse_eps": 0.01},
        "plain": plain,
        "denoisers": rows,
    }
    print(json.dumps(line), flush=True)


def fog_summary(ref_spp, frames=3):
    """The fog configuration at 1080p: the 1M-triangle scene in a 128^3 fp16 density grid, 4 spp,
    depth 8. Throughput over `frames` frames, then relMSE against a ref_spp render of the noisy
    4 spp sample mean, the 7x7 bilateral filter (radius 3, 29 taps) with and without the albedo
    guide, and the learned denoiser (trained on fog-free scenes)."""
    import numpy as np
    from sharpray_amd import RNG_COUNTER, Compositor, KpnDenoiser, Renderer, scenes
    spec = scenes.config_fog(scale=1.0, spp=4)
    scene, view = scenes.instantiate(spec)
    scene.build()
    W, H = spec.width, spec.height

    def render(spp, seed):
        r = Renderer(scene, view, W, H, 32, spp, spp, 0.0, spec.bounces, spec.indirect_clamp,
                     rng_policy=RNG_COUNTER, rng_seed=seed)
        r.wait()
        return r

    # The reference first: a render's AOVs live on the scene's context until its next render.
    t0 = time.perf_counter()
    ref = render(ref_spp, 999).sample_mean.reshape(H, W, 3).astype(np.float64)
    log(f"[bench] fog reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    render(4, 1)
    t0 = time.perf_counter()
    for i in range(frames):
        r = render(4, 7)
    ms = 1e3 * (time.perf_counter() - t0) / frames
    rel = _relmse_fn(ref, H, W)
    out = {"workload": spec.name, "ms_per_frame": round(ms, 3), "msamples_per_s": round(W * H * 4 / ms / 1e3, 3),
           "reference_spp": ref_spp, "relmse_noisy_4spp": round(rel(r.sample_mean), 6)}
    radius = 0.000694  # ceil(3 * 0.000694 * sqrt(1920 * 1080)) = 3: a 7x7 window
    for tag, kw in (("bilateral_4spp", {"sample_mean": True}), ("bilateral_4spp_albedo_guided", {"albedo_limit": 0.15})):
        best = min((rel(Compositor(radius, st, 1.75, 0.25, 2.0, 0.125, 0.2).filter(r, **kw)), st)
                   for st in (0.05, 0.2, 1.0, 5.0, 25.0))
        out[tag] = {"relmse": round(best[0], 6), "strength": best[1]}
    den = KpnDenoiser(scene)
    den.load(KpnDenoiser.pretrained_weights())
    out["kpn_4spp_relmse"] = round(rel(den.denoise(r)), 6)
    log(f"[bench] fog: {out}")
    return out


def adjoint_summary(steps=5):
    """Forward and backward time of the adjoint configuration (512 x 512, 16 spp, 100k triangles,
    16 materials + the 256 x 128 environment) at parameters perturbed +-30 % from the target's."""
    import numpy as np
    from sharpray_amd import Adjoint, scenes
    spec = scenes.config_adjoint(scale=1.0, spp=16)
    scene, view = scenes.instantiate(spec)
    adj = Adjoint(scene)
    fw = (view, spec.width, spec.height, spec.max_samples, spec.bounces, spec.indirect_clamp)
    target = adj.forward(*fw, seed=1)
    p = adj.params()
    p0 = (p * np.random.default_rng(0).uniform(0.7, 1.3, size=p.shape)).astype(np.float32)
    m = p0[:len(spec.objects) * 4].reshape(-1, 4)
    m[:, :3] = np.clip(m[:, :3], 0.0, 1.0)
    m[:, 3] = np.clip(m[:, 3], 0.05, 1.0)
    adj.set_params(p0)
    img = adj.forward(*fw, seed=2)
    adj.backward(img - target)
    tf, tb = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        img = adj.forward(*fw, seed=2)
        t1 = time.perf_counter()
        adj.backward(img - target)
        tf.append(t1 - t0)
        tb.append(time.perf_counter() - t1)
    mf, mb = 1e3 * sum(tf) / steps, 1e3 * sum(tb) / steps
    paths = spec.width * spec.height * spec.max_samples
    out = {"workload": spec.name, "forward_ms": round(mf, 3), "backward_ms": round(mb, 3),
           "backward_over_forward": round(mb / mf, 3), "backward_mpaths_per_s": round(paths / mb / 1e3, 3),
           "params": adj.param_count}
    log(f"[bench] adjoint: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="image scale (1.0 = 1920x1080)")
    ap.add_argument("--spp", type=int, default=None, help="samples per pixel (default 64; 4 for --workload fog)")
    ap.add_argument("--ref-spp", type=int, default=2048, help="reference samples for --workload denoise")
    ap.add_argument("--workload", choices=["plain", "fog", "adjoint", "cnn", "denoise"], default="plain",
                    help="plain: the 1080p / 1M-triangle forward frame; fog: the same scene inside a 128^3 fp16 "
                         "density grid (delta tracking); adjoint: forward + backward at 512x512, 16 spp, 100k triangles; "
                         "cnn: the U-Net denoiser on a 1080p 4 spp render; denoise: relMSE-matched time of the learned "
                         "denoiser and the bilateral filter against the plain tracer at 1080p")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="plain workload: skip the denoiser, fog and adjoint summaries after the timed frames")
    ap.add_argument("--extras-ref-spp", type=int, default=4096,
                    help="reference samples of the denoiser and fog summaries' relMSE")
    ap.add_argument("--light-sampler", choices=["cdf", "alias"], default="cdf",
                    help="environment light sampling: the reference's row/column CDFs or the alias table")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--in-flight", type=int, default=32,
                    help="samples of every pixel in flight at once with the counter policy (1 to 32)")
    ap.add_argument("--rng", choices=["counter", "xorwow"], default="counter",
                    help="random numbers of the plain workload: the counter-based policy (Philox keyed by "
                         "pixel and sample; samples of a pixel independent) or the reference's per-pixel xorwow "
                         "stream (samples in order)")
    ap.add_argument("--tail-segment", type=int, default=None,
                    help="segment from which paths finish in one fused launch (default: the library's)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (torch clone), to calibrate HBM counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        for _ in range(3):
            y = x.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "checksum": float(y[:16].sum())}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (pixels shard trivially across GPUs; see the "
              "sharding notes); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from sharpray_amd import scenes

    t0 = time.perf_counter()
    if args.workload == "fog":
        spec = scenes.config_fog(scale=args.scale, spp=args.spp or 4)
    elif args.workload == "adjoint":
        spec = scenes.config_adjoint(scale=args.scale, spp=args.spp or 16)
    elif args.workload in ("cnn", "denoise"):
        spec = scenes.config_plain_gpu(scale=args.scale, spp=args.spp or 4)
    else:
        spec = scenes.config_plain_gpu(scale=args.scale, spp=args.spp or 64)
    log(f"[bench] scene spec: {spec.triangle_count} triangles, {spec.width}x{spec.height}, "
        f"{spec.max_samples} spp ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    scene, view = scenes.instantiate(spec)
    if args.tail_segment is not None:
        scene.set_tail_segment(args.tail_segment)
    scene.set_samples_in_flight(args.in_flight)
    scene.build()
    log(f"[bench] BVH build + upload {time.perf_counter() - t0:.1f} s")
    if args.workload == "adjoint":
        bench_adjoint(args, spec, scene, view)
        return
    if args.workload == "cnn":
        bench_cnn(args, spec, scene, view)
        return
    if args.workload == "denoise":
        bench_denoise(args, spec, scene, view)
        return

    from sharpray_amd import LIGHT_ALIAS, LIGHT_CDF, RNG_COUNTER, RNG_REFERENCE, Renderer

    light = LIGHT_ALIAS if args.light_sampler == "alias" else LIGHT_CDF
    policy = RNG_COUNTER if args.rng == "counter" else RNG_REFERENCE

    def frame(**kw):
        kw.setdefault("light_sampler", light)
        kw.setdefault("rng_policy", policy)
        return Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                        spec.variance_threshold, spec.bounces, spec.indirect_clamp, **kw)

    for i in range(args.warmup):
        r = frame()
        r.wait()
        log(f"[bench] warmup {i} done")
    # Timed region: K full frames back to back, host-synchronised on both sides.
    walls, gpus, ext_ms, ext_launch, tail_ms, busy_ms = [], [], 0.0, 0, 0.0, 0.0
    for i in range(args.steps):
        t0 = time.perf_counter()
        r = frame(kernel_timing=True)
        r.wait()
        walls.append(time.perf_counter() - t0)
        gpus.append(r.gpu_seconds())
        kt = r.kernel_times()
        kb = r.kernel_busy()
        ext_ms += kt["trace"][0]
        ext_launch += kt["trace"][1]
        tail_ms += kt["trace_tail"][0]
        busy_ms += kb["trace"]
        log(f"[bench] step {i}: {walls[-1] * 1e3:.1f} ms wall, {gpus[-1] * 1e3:.1f} ms GPU, "
            f"trace {kt['trace'][0]:.1f} ms / {kt['trace'][1]} launches, shade {kt['shade'][0]:.1f} ms, "
            f"fused tail {kt['trace_tail'][0]:.1f} ms / {kt['trace_tail'][1]} launches, "
            f"other {kt['other'][0]:.1f} ms (launch spans summed over the concurrent slices); "
            f"trace busy {kb['trace']:.1f} ms")
    samples = spec.width * spec.height * spec.max_samples
    ms_per_step = 1e3 * sum(walls) / len(walls)
    msps = samples / (sum(walls) / len(walls)) / 1e6
    # The same frame with the reference's per-pixel xorwow stream, whose samples run one at a time.
    xorwow = None
    if args.rng == "counter":
        xw = []
        for i in range(2):
            t0 = time.perf_counter()
            frame(rng_policy=RNG_REFERENCE).wait()
            xw.append(time.perf_counter() - t0)
        xorwow = {"ms_per_step": round(1e3 * sum(xw) / len(xw), 3),
                  "msamples_per_s": round(samples / (sum(xw) / len(xw)) / 1e6, 3),
                  "rng": "reference xorwow per pixel (a pixel's samples in order, one in flight)"}
        log(f"[bench] xorwow policy: {xorwow}")

    # Traversal counters of one (untimed) frame give the algorithmic bytes of the trace kernel
    # (its node and triangle tests include the shadow rays'): all traversal minus the fused tail
    # launch's share.
    r = frame(counters=True)
    r.wait()
    c = r.counters()
    fz = r.counters_fused()
    # The reference's way of tracing shadow rays (rays toward a light behind the surface, closest-hit
    # walks) is followed only while counting; the timed frames skip that share.
    sc = r.counters_reference_only()
    ext_bytes = (BYTES_PER_BOX * (c["BvhIntersectNode"] - fz["BvhIntersectNode"] - sc["BvhIntersectNode"]) +
                 BYTES_PER_TRI * (c["BvhIntersectShape"] - fz["BvhIntersectShape"] - sc["BvhIntersectShape"]) +
                 BYTES_PER_RAY * (c["SceneTrace"] - fz["SceneTrace"]) +
                 BYTES_PER_SHADOW * (c["SceneOcclude"] - fz["SceneOcclude"] - sc["SceneOcclude"]))
    # With concurrent row slices (sr_set_slices; one by default) trace launches overlap each other
    # and other kernels: achieved = the frame's algorithmic trace bytes / the time at least one trace
    # launch ran (union of the launch spans); with one slice that is the launches' summed duration.
    launches_per_frame = ext_launch / args.steps
    avg_launch_s = ext_ms / ext_launch / 1e3
    busy_launch_s = busy_ms / ext_launch / 1e3
    bytes_per_launch = ext_bytes / launches_per_frame
    achieved = bytes_per_launch / busy_launch_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "trace_hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    log(f"[bench] counters: {c}; fused tail share: {fz}; reference-only share: {sc}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU oracle baseline ...")
        cpu = cpu_baseline(spec, args.cpu_seconds)
        log(f"[bench] CPU: {cpu}")

    # The metric's second half (relMSE-matched time of the denoisers and the adjoint against the
    # plain tracer), measured after the timed frames on the same GPU.
    extras = {}
    if args.workload == "plain" and not args.no_extras:
        _, rows = denoise_summary(spec, scene, view, args.extras_ref_spp)
        k1, b4 = rows["kpn_1spp"], rows["bilateral_4spp"]
        extras["denoise"] = {
            "reference_spp": args.extras_ref_spp, "relmse_eps": 0.01,
            "kpn_1spp": {k: k1[k] for k in ("relmse", "ms", "denoise_ms", "plain_ms_at_same_relmse",
                                            "speedup_vs_plain")},
            "bilateral_4spp": {k: b4[k] for k in ("relmse", "ms", "strength", "plain_ms_at_same_relmse",
                                                  "speedup_vs_plain")},
            "bilateral_4spp_radiance_aov_relmse": rows["bilateral_4spp_radiance_aov"]["relmse"],
            "kpn_4spp_relmse": rows["kpn_4spp"]["relmse"],
            "kpn_1spp_below_bilateral_4spp": k1["relmse"] < b4["relmse"],
        }
        del scene
        extras["fog"] = fog_summary(args.extras_ref_spp)
        extras["adjoint"] = adjoint_summary()

    line = {
        "metric": "Msamples/s @1080p/1M tris" + (" + 128^3 grid fog" if args.workload == "fog" else ""),
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": spec.name, "width": spec.width, "height": spec.height, "spp": spec.max_samples,
                   "bounces": spec.bounces, "triangles": spec.triangle_count, "env_map": "2048x1024 f32",
                   "rng": ("counter-based (Philox4x32-10 keyed by pixel, counter by sample)" if args.rng == "counter"
                           else "reference xorwow per pixel"),
                   "samples_in_flight": args.in_flight if args.rng == "counter" else 1,
                   "light_sampler": args.light_sampler},
        "gpu_ms_per_step": round(1e3 * sum(gpus) / len(gpus), 3),
        "roofline": {"kernel": "k_trace (closest hit + deferred shadow rays)", "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic, "busy_us_per_launch": round(busy_launch_s * 1e6, 2),
                     "avg_launch_us": round(avg_launch_s * 1e6, 2), "launches_per_frame": round(launches_per_frame),
                     "trace_busy_ms_per_frame": round(busy_ms / args.steps, 3),
                     "algorithmic_bytes_per_launch": round(bytes_per_launch)},
        "fused_tail_ms_per_step": round(tail_ms / args.steps, 3),
        "xorwow": xorwow,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(msps / cpu["value"], 2)
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
