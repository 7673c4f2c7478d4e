"""Trains the kernel-predicting denoiser on pairs the tracer renders itself, and compares it with
the bilateral filter against high-spp references.

    python -m sharpray_amd.train_denoiser --scenes 96 --views 4 --size 128 --steps 4000 --out weights.npy

Training scenes are seeded procedural scenes (fBm-displaced spheres, `--tris` triangles in all, at
random positions over a ground quad, random GGX materials, log-normal or sun-and-gradient environment
maps). Each scene is seen from `--views` random cameras, and each view contributes one example:
the AOVs of a low-spp render (radiance, albedo, normals, depth; the sample count cycles through
`--noisy-spp`) and the radiance of a high-spp render of the same view, both as per-pixel sample
means (the renderer, like the reference, divides pixels that reach the sample cap by M + 1). Held-out scenes are
then rendered at the evaluation size; the learned denoiser at `--eval-spp` samples is compared
with the bilateral filter at 1x and 4x those samples (its strength chosen by grid search on the
references, which favours it), in relMSE = mean (x - r)^2 / (r^2 + 0.01).
"""
import argparse
import json
import math
import sys
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import (RNG_COUNTER, Compositor, KpnDenoiser, Quat, Renderer, Scene, Transform, View, scenes)


def random_scene(seed, size, spp, tris=(10000, 100000)):
    """A seeded training / evaluation scene; `tris` bounds its noise meshes' total triangle count
    (drawn log-uniformly, split over 1-3 meshes)."""
    rng = np.random.default_rng(seed)
    s = scenes.SceneSpec(name=f"train_{seed}")
    s.width = s.height = size
    s.min_samples = s.max_samples = spp
    s.bounces = int(rng.integers(3, 7))
    s.indirect_clamp = 10.0
    if rng.random() < 0.6:
        s.textures.append(scenes.lognormal_env(128, 64, float(rng.uniform(0.5, 1.5)), seed + 11))
    else:
        sun = rng.normal(size=3)
        sun[1] = abs(sun[1]) + 0.3
        sun /= np.linalg.norm(sun)
        s.textures.append(scenes.procedural_sky(128, 64, peak=float(rng.uniform(10, 60)), sun_dir=tuple(sun),
                                                sun_width=float(rng.uniform(0.03, 0.1)), seed=seed + 12))
    s.sky_texture = 0
    s.sky_yaw = float(rng.uniform(0, 2 * math.pi))
    gp, guv = scenes.quad(-1.0, 8.0)
    s.meshes.append(scenes.triangles_build(gp, None, guv))
    s.mesh_bins.append(8)
    s.objects.append(scenes.ObjectSpec("ground", scenes._identity(),
                                       scenes._mat(rng.uniform(0.2, 0.8, 3), float(rng.uniform(0.3, 1.0))), 0, 0))
    blobs = int(rng.integers(1, 4))
    # The default range keeps the original draw (40-159 rings per mesh: 10k-100k triangles in all);
    # another range draws the total from a stream of its own, so the rest of the scene is unchanged.
    default = tuple(tris) == (10000, 100000)
    total = math.exp(np.random.default_rng(seed ^ 0x3A7).uniform(math.log(tris[0]), math.log(tris[1])))
    for k in range(blobs):
        rings = int(rng.integers(40, 160))
        if not default:
            rings = max(8, int(round(math.sqrt(total / blobs / 4.0))))  # 4 rings^2 triangles a mesh
        pos, uv, _ = scenes.fbm_sphere(rings, 2 * rings, float(rng.uniform(0.02, 0.2)), 4, seed * 7 + k,
                                       radius=float(rng.uniform(0.3, 0.8)))
        pos = pos.reshape(-1, 3, 3) + np.array([rng.uniform(-1.5, 1.5), 0.0, rng.uniform(1.5, 4.0)])
        pos[..., 1] += 0.0 - pos[..., 1].min() - 1.0
        s.meshes.append(scenes.triangles_build(pos.reshape(-1, 9).astype(np.float32), None, uv, smooth=True))
        s.mesh_bins.append(8)
        s.objects.append(scenes.ObjectSpec(f"blob{k}", scenes._identity(),
                                           scenes._mat(rng.uniform(0.05, 0.95, 3), float(rng.uniform(0.05, 1.0)),
                                                       float(rng.integers(0, 2))), 0, len(s.meshes) - 1))
    s.view = scenes._view((float(rng.uniform(-0.8, 0.8)), float(rng.uniform(0.0, 0.8)), -1.5),
                          float(rng.uniform(0, 15)), float(rng.uniform(45, 70)))
    return s


def random_view(rng):
    """A camera behind the scene's objects, turned by up to 25 degrees and pitched down by up to 20."""
    pos = (float(rng.uniform(-1.2, 1.2)), float(rng.uniform(-0.3, 1.0)), float(rng.uniform(-2.5, -0.8)))
    yaw = Quat.angle_axis(math.radians(rng.uniform(-25, 25)), (0.0, 1.0, 0.0))
    pitch = Quat.angle_axis(math.radians(rng.uniform(-5, 20)), (1.0, 0.0, 0.0))
    return View(Transform(pos, yaw * pitch), float(math.radians(rng.uniform(40, 75))), 0.1)


def render_pair(spec, noisy_spp, target_spp, seed, scene=None, view=None):
    """Noisy AOVs and the target radiance of one view (counter RNG: independent seeds)."""
    if scene is None:
        scene, view = scenes.instantiate(spec)
    out = {}
    for key, spp, sd in (("noisy", noisy_spp, seed), ("target", target_spp, seed + 7777)):
        r = Renderer(scene, view, spec.width, spec.height, 32, spp, spp, 0.0, spec.bounces, spec.indirect_clamp,
                     rng_policy=RNG_COUNTER, rng_seed=sd)
        r.wait()
        out[key] = r
        if key == "noisy":
            H, W = spec.height, spec.width
            aov = dict(rad=r.sample_mean.reshape(H, W, 3).copy(), raw=r.radiance.reshape(H, W, 3).copy(),
                       alb=r.albedo.reshape(H, W, 3).copy(),
                       nrm=r.normals.reshape(H, W, 3).copy(), dep=r.depth.reshape(H, W).copy(),
                       var=r.variance.copy())
    target = out["target"].sample_mean.reshape(spec.height, spec.width, 3).copy()
    return scene, view, aov, target


def rel_mse(x, ref):
    ref = np.asarray(ref, np.float64)
    x = np.asarray(x, np.float64).reshape(ref.shape)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scenes", type=int, default=96)
    ap.add_argument("--size", type=int, default=128, help="training render size")
    ap.add_argument("--crop", type=int, default=0, help="split each training render into crops of this size")
    ap.add_argument("--views", type=int, default=1, help="cameras per training scene")
    ap.add_argument("--noisy-spp", default="1", help="comma-separated sample counts of the noisy inputs")
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--workers", type=int, default=1,
                    help="training scenes rendered at once, each on its own device context (results unchanged)")
    ap.add_argument("--augment", type=int, default=1, choices=[1, 2, 3, 4],
                    help="examples per crop: the crop, then its left-right, top-bottom and double mirror images")
    ap.add_argument("--slices", type=int, default=1,
                    help="render training images as this many concurrent row slices (results unchanged)")
    ap.add_argument("--tris", default="10000,100000",
                    help="min,max total triangles of a training scene's noise meshes (log-uniform)")
    ap.add_argument("--data-seconds", type=float, default=0.0,
                    help="stop rendering training scenes after this many seconds (0: no limit)")
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--eval-scenes", type=int, default=8)
    ap.add_argument("--eval-size", type=int, default=256)
    ap.add_argument("--eval-spp", default="1,4", help="comma-separated input sample counts to evaluate")
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0x5EED0100)
    ap.add_argument("--init", default=None, help="initial weights (.npy); default He initialisation")
    ap.add_argument("--out", default=None, help="write the trained weights (.npy float32, or .npz as shipped)")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    trainer = Scene(0)  # a device context that only holds the denoiser and its examples
    den = KpnDenoiser(trainer)
    den.load(np.load(args.init) if args.init else KpnDenoiser.random_weights(args.seed))
    den.train_begin()
    t0 = time.perf_counter()
    noisy_spp = [int(v) for v in str(args.noisy_spp).split(",")]
    tris = tuple(int(v) for v in str(args.tris).split(","))
    vrng = np.random.default_rng(args.seed ^ 0x7E3)
    # Extra cameras in the order a one-by-one pass draws them (scene-major), so that the examples do
    # not depend on --workers.
    extra_views = [[random_view(vrng) for _ in range(args.views - 1)] for _ in range(args.scenes)]

    def scene_examples(i):
        """Builds training scene i and renders its views: (triangles, build s, render s, examples)."""
        ts = time.perf_counter()
        spec = random_scene(args.seed + i, args.size, noisy_spp[0], tris)
        scene, view = scenes.instantiate(spec)
        if args.slices > 1:
            scene.set_slices(args.slices)  # concurrent row slices: same images, launches overlap
        tb = time.perf_counter() - ts
        ex = []
        for v, cam in enumerate([view] + extra_views[i]):
            spp = noisy_spp[(i * args.views + v) % len(noisy_spp)]
            _, _, a, target = render_pair(spec, spp, args.target_spp, args.seed + 100000 + 97 * i + v, scene, cam)
            ex.append((a, target))
        return spec.triangle_count, tb, time.perf_counter() - ts - tb, ex

    n_ex = 0
    checksum = 0.0  # sum of every example's noisy and target radiance: equal for any --workers
    t_build = t_render = 0.0
    n_scenes = 0
    # Each worker renders a scene of its own on its own device context and stream: a 256 x 256
    # render fills a fraction of the GPU, so several at once raise the sample rate.
    pool = ThreadPoolExecutor(max_workers=max(1, args.workers))
    pending = deque()
    submitted = 0
    while submitted < args.scenes or pending:
        while submitted < args.scenes and len(pending) < 2 * max(1, args.workers):
            if args.data_seconds and time.perf_counter() - t0 > args.data_seconds:
                log(f"[train] data time limit reached: no scenes after {submitted}")
                args.scenes = submitted
                break
            pending.append(pool.submit(scene_examples, submitted))
            submitted += 1
        if not pending:
            break
        ntri, tb, tr, ex = pending.popleft().result()
        t_build += tb
        t_render += tr
        for a, target in ex:
            checksum += float(np.sum(target, dtype=np.float64)) + float(np.sum(a["rad"], dtype=np.float64))
            c = args.crop or args.size
            for y0 in range(0, args.size - c + 1, c):
                for x0 in range(0, args.size - c + 1, c):
                    sl = (slice(y0, y0 + c), slice(x0, x0 + c))
                    # The crop, then its mirror images (left-right, top-bottom, both) with --augment.
                    for fl in ((), (1,), (0,), (0, 1))[:max(1, args.augment)]:
                        def crop(x):
                            return np.ascontiguousarray(np.flip(x[sl], fl) if fl else x[sl])
                        den.train_add(crop(a["rad"]), crop(a["alb"]), crop(a["nrm"]), crop(a["dep"]), crop(target),
                                      c, c)
                        n_ex += 1
        n_scenes += 1
        log(f"[train] {n_scenes}/{args.scenes} scenes ({ntri} triangles), {n_ex} examples "
            f"({time.perf_counter() - t0:.1f} s: scene build {t_build:.1f} s, rendering {t_render:.1f} s, "
            f"summed over workers)")
    pool.shutdown()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    chunk = 250
    losses = []
    for k in range(0, args.steps, chunk):
        n = min(chunk, args.steps - k)
        lr = args.lr * (0.1 ** (k / max(args.steps, 1)))  # exponential decay to lr / 10
        losses.append(den.train(n, lr, args.seed + k))
        log(f"[train] step {k + n}/{args.steps}: mean loss {losses[-1]:.5f} (lr {lr:.2e}, "
            f"{time.perf_counter() - t0:.1f} s)")
    t_train = time.perf_counter() - t0
    weights = den.export()
    if args.out:
        if args.out.endswith(".npz"):
            KpnDenoiser.save_weights(args.out, weights)
        else:
            np.save(args.out, weights)

    # Held-out evaluation.
    report = {"train_scenes": n_scenes, "train_examples": n_ex, "train_size": args.size, "noisy_spp": args.noisy_spp,
              "target_spp": args.target_spp, "tris": args.tris, "augment": args.augment, "data_checksum": checksum,
              "steps": args.steps,
              "data_seconds": round(t_data, 1),
              "train_seconds": round(t_train, 1), "final_loss": losses[-1] if losses else None,
              "eval_size": args.eval_size, "ref_spp": args.ref_spp, "eval": {}}
    for espp in [int(v) for v in str(args.eval_spp).split(",")]:
        res = []
        for j in range(args.eval_scenes):
            seed = args.seed + 50000 + j
            spec = random_scene(seed, args.eval_size, espp)
            scene, view, a, ref = render_pair(spec, espp, args.ref_spp, seed)
            H = W = args.eval_size
            ev = KpnDenoiser(scene)
            ev.load(weights)
            r1 = Renderer(scene, view, W, H, 32, espp, espp, 0.0, spec.bounces, spec.indirect_clamp,
                          rng_policy=RNG_COUNTER, rng_seed=seed)
            r1.wait()
            row = {"noisy": rel_mse(a["rad"], ref), "noisy_raw": rel_mse(a["raw"], ref),
                   "kpn": rel_mse(ev.denoise(r1), ref)}
            for mult in (1, 4):
                spp = espp * mult
                rb = Renderer(scene, view, W, H, 32, spp, spp, 0.0, spec.bounces, spec.indirect_clamp,
                              rng_policy=RNG_COUNTER, rng_seed=seed + 31 * mult)
                rb.wait()
                # 7x7 circular window (radius 3 px at this size), strength by grid search; on the
                # reference's Radiance AOV and on the per-pixel sample mean.
                for tag, mean in (("", True), ("_raw", False)):
                    best = rel_mse(rb.sample_mean if mean else rb.radiance, ref)
                    for strength in (0.05, 0.2, 1.0, 5.0, 25.0):
                        comp = Compositor(1.0 / W, strength, 1.75, 0.25, 2.0, 0.125, 0.2)
                        best = min(best, rel_mse(comp.filter(rb, sample_mean=mean), ref))
                    row[f"bilateral_{mult}x{tag}"] = best
            res.append(row)
            log(f"[eval {espp} spp] scene {j}: " + ", ".join(f"{k} {v:.4f}" for k, v in row.items()))
        if res:
            report["eval"][f"{espp}spp"] = {k: round(float(np.mean([r[k] for r in res])), 5) for k in res[0]}
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
