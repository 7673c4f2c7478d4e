"""GPU parity: the gfx950 path (through the C ABI) against the CPU oracle on the same seeded scenes.

Bars (per the tracer's parity contract):
  - BVH node order and item permutation: bit-exact.
  - Sample-0 primary hits (object, triangle, distance): bit-exact; only +, -, *, /, sqrt, min and
    max are involved once tan(fov/2) is fixed.
  - Per-pixel radiance: within 1e-3 relative (absolute below 1e-3) except on a reported fraction of
    divergent pixels. Transcendentals (sin, cos, log, exp, pow, atan2, asin) come from the GPU's
    math library on one side and the host libm on the other; a last-bit difference that flips an
    RNG-consuming branch shifts the pixel's xorwow stream for all later samples.
  - Sample counts, depth and normals of non-divergent pixels follow the same bar.
"""
import numpy as np
import pytest

from tests.helpers import bits, rel_err, rel_mse, record_divergence

pytestmark = pytest.mark.gpu

RAD_TOL = 1e-3
MAX_DIVERGENT = 0.02
# Per-case bars on the fraction of divergent pixels (radiance off by more than RAD_TOL relative):
# about 5-10x the fraction measured on MI355X, and never below 8 pixels of the image.
# Measured (MI355X, this suite): plain_cpu xorwow 1.5e-5, counter 4.6e-5, analytic 1.6e-3, height fog
# 6.1e-5 (fog radiance 0), transparent 1.7e-3, alias 1.5e-5, homogeneous medium 6.1e-5, grid 0.
BARS = {"plain_cpu_xorwow": 2e-4, "plain_cpu_counter": 4e-4, "analytic_xorwow": 1e-2, "height_fog": 5e-4,
        "height_fog_fog_radiance": 0.0, "transparent": 1e-2, "alias_sampler": 2e-4, "medium_1": 5e-4,
        "medium_1_fog_radiance": 0.0, "medium_2": 0.0, "medium_2_fog_radiance": 0.0, "adjoint_small_forward": 0.0}


def check_divergence(name, frac, pixels):
    """Records the measured fraction for the session summary and applies the case's bar."""
    bar = max(BARS.get(name, MAX_DIVERGENT), 8.0 / pixels)
    record_divergence(name, frac, bar, pixels)
    assert frac <= bar, (name, frac, bar)


@pytest.fixture(scope="module")
def plain_cpu():
    from sharpray_amd import scenes
    return scenes.config_plain_cpu()


def _render_both(spec, O, **kw):
    from sharpray_amd import scenes
    scene, view, r = scenes.render(spec, **kw)
    r.wait()
    osc = O.OracleScene(spec)
    ref = osc.render()
    return scene, r, osc, ref


def test_primary_hits_bit_exact_plain_cpu(plain_cpu, oracle_lib):
    from sharpray_amd import scenes
    spec = plain_cpu
    scene, view = scenes.instantiate(spec)
    obj, tri, dist = scene.trace_primary(view, spec.width, spec.height)
    osc = oracle_lib.OracleScene(spec)
    oobj, otri, odist, _ = osc.trace_primary()
    print(f"plain_cpu {spec.width}x{spec.height}: hits={np.count_nonzero(oobj >= 0)}", flush=True)
    assert np.array_equal(obj, oobj)
    assert np.array_equal(tri, otri)
    assert np.array_equal(bits(dist), bits(odist))


def test_primary_hits_bit_exact_1m_triangles(oracle_lib):
    from sharpray_amd import scenes
    spec = scenes.config_plain_gpu(scale=0.25, spp=1)
    scene, view = scenes.instantiate(spec)
    obj, tri, dist = scene.trace_primary(view, spec.width, spec.height)
    osc = oracle_lib.OracleScene(spec)
    for m in range(len(spec.meshes)):
        gn, gi = scene.bvh(m)
        on, oi = osc.bvh(m)
        assert np.array_equal(bits(gn), bits(on)), f"mesh {m} node order"
        assert np.array_equal(gi, oi), f"mesh {m} items"
    oobj, otri, odist, _ = osc.trace_primary()
    print(f"plain_gpu 1/4 scale {spec.width}x{spec.height}, {spec.triangle_count} tris", flush=True)
    assert np.array_equal(obj, oobj)
    assert np.array_equal(tri, otri)
    assert np.array_equal(bits(dist), bits(odist))


def test_radiance_parity_plain_cpu(plain_cpu, oracle_lib):
    spec = plain_cpu
    scene, r, osc, ref = _render_both(spec, oracle_lib)
    err = rel_err(r.radiance, ref["radiance"])
    div = err > RAD_TOL
    frac = float(div.mean())
    print(f"plain_cpu radiance: divergent fraction {frac:.5f}, median rel err {np.median(err):.2e}", flush=True)
    check_divergence("plain_cpu_xorwow", frac, err.size)
    ok = ~div
    assert np.array_equal(r.samples, ref["samples"])
    assert np.all(r.samples == spec.max_samples + 1)  # the reference's M/(M+1) quirk (renderer.cs:173,215)
    assert np.array_equal(np.isinf(r.depth), np.isinf(ref["depth"]))
    fin = ok & np.isfinite(ref["depth"])
    assert np.allclose(r.depth[fin], ref["depth"][fin], rtol=1e-5, atol=0)
    assert np.allclose(r.normals[ok], ref["normals"][ok], atol=1e-5)
    assert np.allclose(r.uv[ok], ref["uv"][ok], atol=1e-5)
    assert np.all(rel_err(r.variance[ok, None], ref["variance"][ok, None]) <= 1e-2)


def _counter_case(case):
    from sharpray_amd import RNG_COUNTER, scenes
    if case == "plain":
        return scenes.config_plain_cpu(), {}, {}
    ctr = (dict(rng_policy=RNG_COUNTER, rng_seed=77), dict(rng_policy=1, rng_seed=77))
    if case == "height_fog":
        spec = scenes.config_plain_cpu(scale=0.5, spp=8)
        spec.fog = (0.35, (0.9, 0.85, 0.8), 0.3, 0.6)
        return (spec,) + ctr
    spec = scenes.config_analytic(scale=1.0, spp=8)
    for o in spec.objects[1:6]:
        o.material = o.material.copy()
        o.material[4] = 0.0   # metallic
        o.material[5] = 0.85  # transparency
        o.material[6] = 1.45  # ior
    return (spec,) + ctr


@pytest.mark.parametrize("case", ["plain", "height_fog", "transparent"])
def test_render_counters_match_oracle(case, oracle_lib):
    """Counters.Type render-path counters (counters.cs:24-79) of the GPU render against the oracle's
    on the same scene and random streams. Sample and pixel counts are exact; the per-ray counters
    (bounces, hits, fog events, traces, node and shape tests) differ only through the few divergent
    paths (see the radiance bar) and must agree to 0.2 % (measured on the plain scene: 1e-6). Shadow
    rays count as the reference traces them: also toward a light behind the surface
    (scene.cs:774-783), and with the closest-hit walk (scene.cs:576-600); skipping either
    undercounts SceneOcclude by ~40 % or the node tests by ~1.4 % on the plain scene."""
    from sharpray_amd import scenes
    spec, kw, okw = _counter_case(case)
    scene, view, r = scenes.render(spec, counters=True, **kw)
    r.wait()
    g = r.counters()
    o = oracle_lib.OracleScene(spec).render(counters=True, **okw)["counters"]
    print(case, "gpu", g, "\noracle", o, flush=True)
    for k in ("Pixel", "Sample", "SamplePixelConverged", "SamplePixelMaxed"):
        assert g[k] == o[k], k
    assert g["Pixel"] == spec.width * spec.height
    for k in ("SampleBounce", "SampleHit", "SampleMiss", "SampleTerminate", "SampleFogScatter", "SampleFogEscape",
              "SceneTrace", "SceneOcclude", "MeshIntersect", "BvhIntersectNode", "BvhIntersectShape"):
        assert abs(g[k] - o[k]) <= 0.002 * max(o[k], 1000), (k, g[k], o[k])
    for k in ("SampleBounce", "SampleHit", "SceneOcclude", "BvhIntersectNode"):
        assert o[k] > 0, k
    if case == "height_fog":
        assert o["SampleFogScatter"] > 0
    # Every traced segment is one SceneTrace.
    assert g["SceneTrace"] == g["SampleBounce"]
    assert g["BvhIntersectNode"] % 2 == 0  # two child boxes per interior visit (bvh.cs:188-189)


def test_radiance_parity_counter_rng(plain_cpu, oracle_lib):
    """Counter policy: each sample draws from Philox(seed, pixel, sample), so a divergent path cannot
    shift the draws of the pixel's other samples."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = plain_cpu
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=1234)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=1234)
    err = rel_err(r.radiance, ref["radiance"])
    frac = float((err > RAD_TOL).mean())
    print(f"plain_cpu counter-rng radiance: divergent fraction {frac:.5f}", flush=True)
    check_divergence("plain_cpu_counter", frac, err.size)


def test_radiance_parity_analytic_shapes(oracle_lib):
    from sharpray_amd import scenes
    spec = scenes.config_analytic(scale=1.0, spp=4)
    scene, r, osc, ref = _render_both(spec, oracle_lib)
    err = rel_err(r.radiance, ref["radiance"])
    frac = float((err > RAD_TOL).mean())
    print(f"analytic radiance: divergent fraction {frac:.5f}", flush=True)
    check_divergence("analytic_xorwow", frac, err.size)


def test_bilateral_filter_parity(plain_cpu, oracle_lib):
    from sharpray_amd import Compositor
    spec = plain_cpu
    scene, r, osc, ref = _render_both(spec, oracle_lib)
    comp = Compositor(0.01, 0.1, 1.75, 0.25, 2.0, 0.125, 0.2)
    out = comp.filter(r)
    oref, _ = oracle_lib.bilateral(comp.as_array(), r.radiance, r.normals, r.depth, r.variance, spec.width,
                                   spec.height)
    err = rel_err(out, oref, floor=1e-3)
    print(f"bilateral: max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4


def test_radiance_parity_height_fog(oracle_lib):
    """Exponential height fog (scene.cs:438-505): free flight, fog NEE with HG MIS, the fog
    radiance AOV of primary segments, fog transmittance of shadow rays. Counter RNG policy, so a
    path whose decision flips cannot shift the rest of the pixel."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=0.5, spp=8)
    spec.fog = (0.35, (0.9, 0.85, 0.8), 0.3, 0.6)
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=77)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=77)
    assert float(np.abs(ref["radiance_fog"]).max()) > 0.0
    err = rel_err(r.radiance, ref["radiance"])
    errf = rel_err(r.radiance_fog, ref["radiance_fog"])
    frac = float((err > RAD_TOL).mean())
    fracf = float((errf > RAD_TOL).mean())
    print(f"height fog: divergent fraction radiance {frac:.5f}, fog radiance {fracf:.5f}", flush=True)
    check_divergence("height_fog", frac, err.size)
    check_divergence("height_fog_fog_radiance", fracf, err.size)


def test_radiance_parity_transparent_objects(oracle_lib):
    """Transparent spheres: TransparentScatter (scene.cs:121-143) and the shadow walk through
    transparent surfaces (scene.cs:576-600), i.e. the closest-hit shadow path of the trace kernel."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = scenes.config_analytic(scale=1.0, spp=8)
    for o in spec.objects[1:6]:
        o.material = o.material.copy()
        o.material[4] = 0.0   # metallic
        o.material[5] = 0.85  # transparency
        o.material[6] = 1.45  # ior
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=5)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=5)
    err = rel_err(r.radiance, ref["radiance"])
    frac = float((err > RAD_TOL).mean())
    print(f"transparent: divergent fraction {frac:.5f}", flush=True)
    check_divergence("transparent", frac, err.size)


def test_converged_image_matches_float64_oracle(oracle_lib):
    """The GPU's float32 estimate converges to the float64 oracle's: its relMSE against a float64
    render (independent random numbers) is no larger than the noise between two GPU renders with
    different seeds, up to a 25 % margin. relMSE = mean (x - r)^2 / (r^2 + 1e-2)."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=0.25, spp=512)
    _, _, a = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=1)
    a.wait()
    ra = a.radiance.copy()
    _, _, b = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=2)
    b.wait()
    rb = b.radiance.copy()
    ref = oracle_lib.OracleScene(spec, precision="f64").render(rng_policy=1, rng_seed=3)["radiance"]
    noise = rel_mse(ra, rb)
    vs64 = rel_mse(ra, ref)
    print(f"converged 64x64 @512spp: relMSE gpu-vs-fp64 {vs64:.3e}, gpu-vs-gpu {noise:.3e}", flush=True)
    assert vs64 <= 1.25 * noise


def test_alias_light_sampler_parity(plain_cpu, oracle_lib):
    """The alias-table light sampler (same two uniforms, same pdf) against the oracle's alias
    sampler: the host-built tables must agree entry for entry for the per-sample parity to hold."""
    from sharpray_amd import LIGHT_ALIAS, RNG_COUNTER, scenes
    spec = plain_cpu
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=9, light_sampler=LIGHT_ALIAS)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=9, light_sampler=1)
    err = rel_err(r.radiance, ref["radiance"])
    frac = float((err > RAD_TOL).mean())
    print(f"alias sampler radiance: divergent fraction {frac:.5f}", flush=True)
    check_divergence("alias_sampler", frac, err.size)


def test_albedo_aov(plain_cpu, oracle_lib):
    """Mean primary-hit albedo over the samples that hit a surface (counter policy: the same camera
    jitter per sample on both sides, so the hits and hence the averages agree)."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = plain_cpu
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=21)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=21)
    assert np.any(ref["albedo"] > 0) and np.any(ref["albedo"] == 0)
    np.testing.assert_allclose(r.albedo, ref["albedo"], rtol=1e-6, atol=1e-7)


def test_bilateral_filter_albedo_guided_parity(plain_cpu, oracle_lib):
    from sharpray_amd import Compositor, RNG_COUNTER, scenes
    spec = plain_cpu
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=4)
    r.wait()
    comp = Compositor(0.01, 0.1, 1.75, 0.25, 2.0, 0.125, 0.2)
    out = comp.filter(r, albedo_limit=0.15)
    oref, _ = oracle_lib.bilateral(comp.as_array(), r.radiance, r.normals, r.depth, r.variance, spec.width,
                                   spec.height, albedo=r.albedo, albedo_limit=0.15)
    err = rel_err(out, oref, floor=1e-3)
    print(f"albedo-guided bilateral: max rel err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-4


@pytest.mark.parametrize("kind", [1, 2])
def test_radiance_parity_bounded_media(plain_cpu, oracle_lib, kind):
    """Homogeneous box (kind 1) and fp16 density grid with delta tracking (kind 2): free flight,
    fog NEE with exact voxel-walk transmittance, HG scattering. Counter RNG policy."""
    from sharpray_amd import FogVolume, RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=0.5, spp=8)
    box = ((-3.0, -1.0, -2.0), (3.0, 2.0, 4.0))
    if kind == 1:
        spec.fog_volume = FogVolume(1, 0.25, (0.9, 0.9, 0.95), 0.3, box[0], box[1])
    else:
        grid, _ = scenes.fbm_density_grid(24, 99, box=box)
        spec.fog_volume = FogVolume(2, 0.6, (0.9, 0.9, 0.95), 0.3, box[0], box[1], grid)
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=31)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=31)
    assert float(np.abs(ref["radiance_fog"]).max()) > 0.0
    err = rel_err(r.radiance, ref["radiance"])
    errf = rel_err(r.radiance_fog, ref["radiance_fog"])
    frac = float((err > RAD_TOL).mean())
    fracf = float((errf > RAD_TOL).mean())
    print(f"medium kind {kind}: divergent fraction radiance {frac:.5f}, fog radiance {fracf:.5f}", flush=True)
    check_divergence(f"medium_{kind}", frac, err.size)
    check_divergence(f"medium_{kind}_fog_radiance", fracf, err.size)


def test_adjoint_matches_oracle(oracle_lib):
    """The GPU backward pass against the oracle's float32 adjoint on the same counter-RNG paths:
    material gradients per parameter, sky-texel gradients as a vector."""
    from sharpray_amd import Adjoint, scenes
    spec = scenes.config_adjoint_small(48, 36, 16, 2)
    scene, view = scenes.instantiate(spec)
    adj = Adjoint(scene)
    img = adj.forward(view, spec.width, spec.height, spec.max_samples, spec.bounces, spec.indirect_clamp, seed=7)
    osc = oracle_lib.OracleScene(spec)
    ref = osc.render(rng_policy=1, rng_seed=7)["radiance"]
    check_divergence("adjoint_small_forward", float((rel_err(img, ref) > RAD_TOL).mean()), img.shape[0])
    rng = np.random.default_rng(3)
    delta = rng.normal(size=(spec.width * spec.height, 3)).astype(np.float32)
    g = adj.backward(delta)
    go = osc.adjoint(delta, seed=7)
    nm = len(spec.objects) * 4
    scale = float(np.abs(go[:nm]).max())
    err_m = float(np.abs(g[:nm] - go[:nm]).max()) / scale
    err_e = float(np.linalg.norm(g[nm:] - go[nm:]) / np.linalg.norm(go[nm:]))
    print(f"adjoint vs oracle: material max err {err_m:.2e} (of max |g|), sky rel L2 err {err_e:.2e}", flush=True)
    assert err_m <= 5e-3 and err_e <= 1e-2


def test_adjoint_albedo_gradients_match_gpu_finite_differences():
    """Albedo gradients of the GPU backward pass against central differences of GPU forward renders
    with common random numbers (sampling does not move with albedo for metallic 0 or 1)."""
    from sharpray_amd import Adjoint, scenes
    spec = scenes.config_adjoint_small(48, 36, 16, 2)
    scene, view = scenes.instantiate(spec)
    adj = Adjoint(scene)
    args = (view, spec.width, spec.height, spec.max_samples, spec.bounces, spec.indirect_clamp)
    rng = np.random.default_rng(4)
    delta = rng.normal(size=(spec.width * spec.height, 3))
    p0 = adj.params()
    adj.forward(*args, seed=11)
    g = adj.backward(delta)
    idx = [o * 4 + c for o in range(len(spec.objects)) for c in range(3)]
    gmax = float(np.abs(g[idx]).max())
    worst = 0.0
    for i in idx:
        eps = 1e-2
        hi, lo = p0.copy(), p0.copy()
        hi[i] += eps
        lo[i] -= eps
        adj.set_params(hi)
        fh = float((adj.forward(*args, seed=11).astype(np.float64) * delta).sum())
        adj.set_params(lo)
        fl = float((adj.forward(*args, seed=11).astype(np.float64) * delta).sum())
        fd = (fh - fl) / (2 * eps)
        worst = max(worst, abs(g[i] - fd) / max(abs(fd), 0.05 * gmax))
    adj.set_params(p0)
    print(f"adjoint albedo vs GPU finite differences: worst rel err {worst:.2e}", flush=True)
    assert worst <= 1e-2


def test_kpn_denoiser_layers_match_cpu_restatement():
    """The bf16-MFMA U-Net against the float32 restatement (oracle/cnn_ref.py): every layer
    recomputed on the CPU from the GPU's own input activations agrees to one bf16 rounding step;
    the whole network end to end stays within 2 % of the radiance scale."""
    from oracle import cnn_ref as C
    from sharpray_amd import KpnDenoiser, RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=0.16, spp=2)  # 41 x 41: odd sizes at every level
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=3)
    r.wait()
    den = KpnDenoiser(scene)
    blob = KpnDenoiser.random_weights(17)
    den.load(blob)
    out = den.denoise(r)
    H, W = spec.height, spec.width
    rad = r.sample_mean.reshape(H, W, 3)  # the network's input: Radiance * Samples / min(Samples, spp)
    alb, nrm, dep = r.albedo.reshape(H, W, 3), r.normals.reshape(H, W, 3), r.depth.reshape(H, W)
    feat = den.layer(-1)
    ref_feat = C.features(rad, alb, nrm, dep)
    assert np.all(np.abs(feat - ref_feat) <= np.abs(ref_feat) * 2.0 ** -7 + 1e-7)
    wb = C.split_weights(blob)
    gpu = {-1: feat}
    worst = 0.0
    for i in range(len(C.LAYERS)):
        gpu[i] = den.layer(i)
        ref = C.layer_forward(i, wb, gpu)
        assert ref.shape == gpu[i].shape, (i, ref.shape, gpu[i].shape)
        scale = max(float(np.abs(ref).max()), 1e-6)
        err = np.abs(gpu[i] - ref)
        tol = np.abs(ref) * 2.0 ** -7 + 1e-4 * scale
        worst = max(worst, float((err / (np.abs(ref) + 1e-4 * scale)).max()))
        assert np.all(err <= tol), (i, float(err.max()), scale)
    acts, ref_out = C.run(blob, rad, alb, nrm, dep)
    rel = np.abs(out.reshape(H, W, 3) - ref_out) / (np.abs(ref_out) + 1e-2)
    print(f"kpn denoiser: worst per-layer rel diff {worst:.2e}, end-to-end max rel diff {rel.max():.2e}", flush=True)
    assert rel.max() <= 2e-2


def test_kpn_full_resolution_layers_match_cpu_restatement_on_many_tiles():
    """The full-resolution layers (0, 1, 16 with its upsampled 48-channel source and 32-channel skip,
    17) and the logit head (18) at 648 x 372, where each resident workgroup of their persistent
    kernel walks several tiles (984 tiles of 16 x 16 pixels, 3813 of 16 x 4 for the head, ragged at
    the right and bottom edges), against the float32 restatement from the GPU's own input
    activations."""
    from oracle import cnn_ref as C
    from sharpray_amd import KpnDenoiser, RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=0.25, spp=1)
    spec.width, spec.height = 648, 372
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=5)
    r.wait()
    den = KpnDenoiser(scene)
    blob = KpnDenoiser.random_weights(23)
    den.load(blob)
    den.denoise(r)
    wb = C.split_weights(blob)
    gpu = {i: den.layer(i) for i in (-1, 0, 1, 15, 16, 17, 18)}
    for i in (0, 1, 16, 17, 18):
        ref = C.layer_forward(i, wb, gpu)
        assert ref.shape == gpu[i].shape == (372, 648, C.LAYERS[i][2]), (i, ref.shape, gpu[i].shape)
        scale = max(float(np.abs(ref).max()), 1e-6)
        err = np.abs(gpu[i] - ref)
        bad = err > np.abs(ref) * 2.0 ** -7 + 1e-4 * scale
        print(f"layer {i}: max err {float(err.max()):.3e} of scale {scale:.3e}, {int(bad.sum())} over", flush=True)
        assert not bad.any(), (i, np.argwhere(bad)[:5])


def _kpn_pair(spec_fn, seed):
    from sharpray_amd import RNG_COUNTER, scenes
    spec = spec_fn()
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=seed)
    r.wait()
    H, W = spec.height, spec.width
    noisy = dict(rad=r.sample_mean.reshape(H, W, 3).copy(), alb=r.albedo.reshape(H, W, 3).copy(),
                 nrm=r.normals.reshape(H, W, 3).copy(), dep=r.depth.reshape(H, W).copy())
    spec.min_samples = spec.max_samples = 64
    _, _, rt = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=seed + 1000)
    rt.wait()
    return scene, noisy, rt.sample_mean.reshape(H, W, 3).copy(), W, H


def test_kpn_training_gradient_matches_cpu_restatement():
    """One backward pass of the hand-written training kernels (loss, kernel cascade, ReLU, MFMA
    data and weight gradients, pooling / upsampling / concatenation) against the reverse-mode
    restatement, step by step: each layer's output gradient recomputed on the CPU from the GPU's
    own activations and its consumers' GPU gradients, and each layer's weight and bias gradient
    from the GPU's output gradient and staged input. (A whole-network comparison would compound
    bf16 rounding flips of a random 19-layer network, ReLU masks included, into noise.)"""
    from oracle import cnn_ref as C
    from sharpray_amd import KpnDenoiser, scenes
    scene, ex, target, W, H = _kpn_pair(lambda: scenes.config_plain_cpu(scale=0.16, spp=2), 8)
    den = KpnDenoiser(scene)
    blob = KpnDenoiser.random_weights(23)
    den.load(blob)
    den.train_begin()
    den.train_add(ex["rad"], ex["alb"], ex["nrm"], ex["dep"], target, W, H)
    loss, g = den.gradient(0)
    lref, _ = C.backward(blob, ex["rad"], ex["alb"], ex["nrm"], ex["dep"], target)
    assert abs(loss - lref) <= 1e-4 * lref, (loss, lref)
    acts = {i: den.layer(i).astype(np.float64) for i in range(-1, 19)}
    dys = {i: den.layer(1000 + i).astype(np.float64) for i in range(19)}
    wb = C.split_weights(blob)
    # Output gradient of the last layer: the kernel cascade's logit gradient of the loss.
    out_logits = acts[18]
    r = ex["rad"].astype(np.float64)
    out = C.kpn_apply(out_logits, r)
    t = target.astype(np.float64)
    gr = 2.0 * (out - t) / (t ** 2 + 0.01) / out.size
    dl = C.bf16(C.kpn_cascade_back(out_logits, r, gr).astype(np.float32))
    errs = {"dY18": float(np.linalg.norm(dys[18] - dl) / np.linalg.norm(dl))}
    dact = {i: None for i in range(-1, 19)}

    def add(k, v):
        dact[k] = v if dact[k] is None else dact[k] + v

    k = 0
    offsets = []
    for ca, cb, cout, *_ in C.LAYERS:
        offsets.append(k)
        k += cout * (ca + cb) * 9 + cout
    for i in range(18, -1, -1):
        ca, cb, cout, mode, sa, sb, lvl, relu = C.LAYERS[i]
        if i < 18:
            ref = C.bf16((dact[i] * (acts[i] > 0)).astype(np.float32))
            errs[f"dY{i}"] = float(np.linalg.norm(dys[i] - ref) / max(np.linalg.norm(ref), 1e-30))
        x = C.staged_input(i, acts, C.bf16).astype(np.float64)
        hh, ww, cin = x.shape
        dwm = C._cols(x).reshape(hh * ww, 9 * cin).T @ dys[i].reshape(hh * ww, cout)
        gw_ref = dwm.reshape(3, 3, cin, cout).transpose(3, 2, 0, 1).reshape(-1)
        gb_ref = dys[i].sum((0, 1))
        nw = gw_ref.size
        gw, gb = g[offsets[i]:offsets[i] + nw], g[offsets[i] + nw:offsets[i] + nw + cout]
        errs[f"W{i}"] = float(np.linalg.norm(gw - gw_ref) / max(np.linalg.norm(gw_ref), 1e-30))
        errs[f"b{i}"] = float(np.linalg.norm(gb - gb_ref) / max(np.linalg.norm(gb_ref), 1e-30))
        if i == 0:
            break
        w = wb[i][0].astype(np.float64)
        wt = w[:, :, ::-1, ::-1].transpose(1, 0, 2, 3)
        dx = C.bf16(C.conv3x3(dys[i], wt, np.zeros(cin), False, C.exact).astype(np.float32)).astype(np.float64)
        if mode == 0:
            add(sa, dx)
        elif mode == 1:
            fh, fw = acts[sa].shape[:2]
            up_ = dx[np.arange(fh) // 2][:, np.arange(fw) // 2] * 0.25
            if fw % 2:
                up_[:, -1] *= 2.0
            if fh % 2:
                up_[-1] *= 2.0
            add(sa, up_)
        else:
            lh, lw = acts[sa].shape[:2]
            da = np.zeros((lh, lw, ca))
            np.add.at(da, (np.arange(hh)[:, None] // 2, np.arange(ww)[None, :] // 2), dx[..., :ca])
            add(sa, da)
            add(sb, dx[..., ca:])
    worst = max(errs.values())
    print(f"kpn training gradient: loss {loss:.5f} (cpu {lref:.5f}), worst step-wise rel L2 err {worst:.2e} "
          f"({max(errs, key=errs.get)})", flush=True)
    assert worst <= 2e-3, errs


def test_kpn_training_reduces_loss():
    from sharpray_amd import KpnDenoiser, scenes
    scene, ex, target, W, H = _kpn_pair(lambda: scenes.config_plain_cpu(scale=0.25, spp=2), 9)
    den = KpnDenoiser(scene)
    den.load(KpnDenoiser.random_weights(1))
    den.train_begin()
    den.train_add(ex["rad"], ex["alb"], ex["nrm"], ex["dep"], target, W, H)
    first = den.train(10, 1e-3, 0)
    later = den.train(150, 1e-3, 1)
    last = den.train(10, 1e-3, 2)
    print(f"kpn training: loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.7 * first


@pytest.mark.parametrize("scene_index", [0, 1])
def test_kpn_shipped_weights_denoise_held_out_scene(scene_index):
    """The shipped weights on held-out procedural scenes (index 0: sun-and-gradient sky, 1:
    log-normal sky) at 128^2 and 1 spp, against a 1024 spp reference: the learned denoiser must
    cut relMSE = mean (x - r)^2 / (r^2 + 0.01) well below the noisy input's, below the best
    bilateral filter (strength grid) at the same sample count, and below the best bilateral filter
    at 4 spp (the learned denoiser reaching the bilateral filter's error at 1/4 of the samples)."""
    from sharpray_amd import RNG_COUNTER, Compositor, KpnDenoiser, train_denoiser as T
    seed = 0x5EED0100 + 50000 + scene_index
    spec = T.random_scene(seed, 128, 1)
    scene, view, a, ref = T.render_pair(spec, 1, 1024, seed)
    den = KpnDenoiser(scene)
    den.load(KpnDenoiser.pretrained_weights())
    from sharpray_amd import Renderer
    r = Renderer(scene, view, 128, 128, 32, 1, 1, 0.0, spec.bounces, spec.indirect_clamp,
                 rng_policy=RNG_COUNTER, rng_seed=seed)
    r.wait()
    noisy = T.rel_mse(a["rad"], ref)
    kpn = T.rel_mse(den.denoise(r), ref)
    bil = min(T.rel_mse(Compositor(1.0 / 128, st, 1.75, 0.25, 2.0, 0.125, 0.2).filter(r), ref)
              for st in (0.05, 0.2, 1.0, 5.0, 25.0))
    # The claim: the learned denoiser at 1 spp reaches below the bilateral filter's error at 4x the
    # samples (its best strength on the 4-spp render, on Radiance and on the sample mean).
    r4 = Renderer(scene, view, 128, 128, 32, 4, 4, 0.0, spec.bounces, spec.indirect_clamp,
                  rng_policy=RNG_COUNTER, rng_seed=seed + 7)
    r4.wait()
    bil4 = min(T.rel_mse(Compositor(1.0 / 128, st, 1.75, 0.25, 2.0, 0.125, 0.2).filter(r4, sample_mean=sm), ref)
               for st in (0.05, 0.2, 1.0, 5.0, 25.0) for sm in (False, True))
    print(f"kpn shipped weights, scene {scene_index}: noisy {noisy:.4f}, kpn@1spp {kpn:.4f}, bilateral@1spp "
          f"{bil:.4f}, bilateral@4spp {bil4:.4f}", flush=True)
    assert kpn < 0.5 * noisy and kpn < bil
    assert kpn < bil4


def test_compose_tail_parity(oracle_lib):
    """Compose (compose.cs:74-99): Filter + FilterFog + Tonemap + ToPixel, and Preview, against the
    oracle on the GPU's own AOVs (a height-fog render, so the fog radiance is non-zero). FilterFog
    within 1e-5 relative; bytes equal except where a transcendental's last bit moves a value
    across a rounding boundary (at most 1 step, on under 1 % of the bytes)."""
    from sharpray_amd import RNG_COUNTER, TONEMAP_LINEAR_SMOOTH, TONEMAP_REINHARD, Compositor, scenes
    spec = scenes.config_plain_cpu(scale=0.5, spp=4)
    spec.fog = (0.35, (0.9, 0.85, 0.8), 0.3, 0.6)
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=11)
    r.wait()
    W, H = spec.width, spec.height
    assert float(np.abs(r.radiance_fog).max()) > 0.0
    for tm, exposure in [(TONEMAP_REINHARD, 1.0), (TONEMAP_LINEAR_SMOOTH, 0.7)]:
        comp = Compositor(0.01, 0.1, 1.75, 0.25, 2.0, 0.125, 0.2, denoise_fog_radius=0.01,
                          denoise_fog_strength=0.5, tonemapper=tm, exposure=exposure)
        fog = comp.filter_fog(r)
        fog_ref = oracle_lib.filter_fog(0.01, 0.5, r.radiance_fog, W, H)
        ferr = rel_err(fog, fog_ref, floor=1e-3)
        main_ref, _ = oracle_lib.bilateral(comp.as_array(), r.radiance, r.normals, r.depth, r.variance, W, H)
        img = comp.compose(r).pixels.reshape(-1, 3).astype(int)
        ref = oracle_lib.tonemap(main_ref, fog_ref, tm, exposure).astype(int)
        d = np.abs(img - ref)
        prev = comp.preview(r).pixels.reshape(-1, 3).astype(int)
        pref = oracle_lib.tonemap(r.radiance, r.radiance_fog, tm, exposure).astype(int)
        dp = np.abs(prev - pref)
        print(f"compose tm={tm}: fog filter max rel err {ferr.max():.2e}; compose bytes off {(d > 0).mean():.5f} "
              f"(max {d.max()}), preview bytes off {(dp > 0).mean():.5f} (max {dp.max()})", flush=True)
        assert ferr.max() <= 1e-5
        assert d.max() <= 1 and (d > 0).mean() <= 0.01
        assert dp.max() <= 1 and (dp > 0).mean() <= 0.01


def _fbm_sphere_tris(rings, seed):
    from sharpray_amd import scenes
    pos, uv, _ = scenes.fbm_sphere(rings, 2 * rings, 0.15, 6, seed)
    return pos.reshape(-1, 9).astype(np.float32)


@pytest.mark.parametrize("case", ["ico3_8", "ico5_16", "clusters_32", "fbm_64", "fbm1m_8", "tiny_2"])
def test_device_bvh_builder_bit_exact(case, oracle_lib):
    """The device builder (level-wise large nodes, per-wave subtrees, depth-first renumbering)
    against the oracle's restatement of bvh.cs and against the host builder: node arrays (bounds,
    child, count) and item permutations bit-identical."""
    import time
    from sharpray_amd import Mesh, Scene, scenes, triangles_build
    rng = np.random.default_rng(7)
    if case == "ico3_8":
        tris, bins = scenes.icosphere(3, seed=3), 8
    elif case == "ico5_16":
        tris, bins = scenes.icosphere(5, seed=5), 16
    elif case == "clusters_32":  # two far clusters plus duplicates: exact ties in centroids
        a = rng.normal(0, 0.1, (3000, 9)).astype(np.float32)
        b = (rng.normal(0, 0.1, (3000, 9)) + 5).astype(np.float32)
        tris, bins = np.concatenate([a, b, a[:500]]), 32
    elif case == "fbm_64":
        tris, bins = _fbm_sphere_tris(96, 11), 64
    elif case == "fbm1m_8":
        tris, bins = _fbm_sphere_tris(512, 1), 8
    else:
        tris, bins = rng.normal(0, 1, (3, 9)).astype(np.float32), 2
    recs = triangles_build(tris, None, None)
    out = {}
    for dev in (False, True):
        sc = Scene(0)
        sc.set_bvh_builder(dev)
        m = Mesh(recs, split_bins=bins)
        t0 = time.perf_counter()
        mid = sc.mesh_id(m)
        dt = time.perf_counter() - t0
        out[dev] = sc.bvh(mid) + (dt,)
    (hn, hi, ht), (dn, di, dtm) = out[False], out[True]
    print(f"device BVH {case}: {len(tris)} triangles, {len(hn)} nodes, host {ht * 1e3:.1f} ms, "
          f"device {dtm * 1e3:.1f} ms", flush=True)
    assert hn.shape == dn.shape
    assert np.array_equal(hn.view(np.uint32), dn.view(np.uint32))
    assert np.array_equal(hi, di)
    spec = scenes.SceneSpec(name=f"bvh_{case}")
    spec.textures.append(np.ones((4, 8, 3), np.float32))
    spec.sky_texture = 0
    spec.meshes.append(recs.reshape(-1, 36))
    spec.mesh_bins.append(bins)
    on, oi = oracle_lib.OracleScene(spec).bvh(0)
    assert on.shape == dn.shape, (on.shape, dn.shape)
    assert np.array_equal(bits(on), bits(dn)) and np.array_equal(oi, di)


@pytest.mark.parametrize("tail", [0, 1, 2, 5])
def test_fused_tail_launch_matches_wavefront(tail):
    """Finishing paths in one fused launch from segment `tail` on gives bit-identical AOVs and
    counters to the all-wavefront render (tail disabled), for odd and even switch segments and two
    samples in a row (the second sample starts from the queues the first left)."""
    from sharpray_amd import RNG_COUNTER, Renderer, scenes
    spec = scenes.config_plain_cpu(scale=0.25, spp=3)
    out = {}
    for t in (99, tail):
        scene, view = scenes.instantiate(spec)
        scene.set_tail_segment(t)
        r = Renderer(scene, view, spec.width, spec.height, 32, 3, 3, 0.0, spec.bounces, spec.indirect_clamp,
                     rng_policy=RNG_COUNTER, rng_seed=5, counters=True)
        r.wait()
        out[t] = (r.radiance.copy(), r.radiance_fog.copy(), r.depth.copy(), r.counters())
    a, b = out[99], out[tail]
    assert np.array_equal(bits(a[0]), bits(b[0])) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3], (a[3], b[3])


@pytest.mark.parametrize("case", ["plain_gpu", "textured_transformed"])
def test_speculative_traversal_matches_counting_render(case):
    """Renders without counters traverse mesh BVHs speculatively (postponed leaves) and bound each
    object's search by the closest world distance found so far; counting renders use the plain
    order over whole meshes. Both must give bit-identical AOVs, on the 1M-triangle scene (nine
    objects sharing one transform) and on a scene of rotated, translated and scaled objects."""
    from sharpray_amd import Renderer, scenes
    if case == "plain_gpu":
        spec = scenes.config_plain_gpu(scale=0.25, spp=4)
    else:
        spec = scenes.config_textured_transformed(scale=1.0, spp=16)
    scene, view = scenes.instantiate(spec)
    scene.build()
    out = {}
    for counting in (False, True):
        r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                     spec.variance_threshold, spec.bounces, spec.indirect_clamp, counters=counting)
        r.wait()
        out[counting] = (r.radiance.copy(), r.radiance_fog.copy(), r.depth.copy(), r.normals.copy())
    print(f"{case} {spec.width}x{spec.height}, {spec.max_samples} spp", flush=True)
    assert np.abs(out[True][0]).sum() > 0
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(bits(a), bits(b))


@pytest.mark.parametrize("slices", [2, 3, 5])
def test_concurrent_slices_match_single_wavefront(slices):
    """A band run as concurrent row slices (own queues and streams) gives bit-identical AOVs and
    counters to the band run as one wavefront, including slice counts that do not divide the tile
    rows and a height that is not a multiple of 8."""
    from sharpray_amd import RNG_COUNTER, Renderer, scenes
    spec = scenes.config_plain_gpu(scale=0.25, spp=4)
    spec.height = 268
    scene, view = scenes.instantiate(spec)
    scene.build()
    out = {}
    for k in (1, slices):
        scene.set_slices(k)
        r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                     spec.variance_threshold, spec.bounces, spec.indirect_clamp, counters=True,
                     rng_policy=RNG_COUNTER, rng_seed=3)
        r.wait()
        out[k] = (r.radiance.copy(), r.radiance_fog.copy(), r.depth.copy(), r.normals.copy(), r.samples.copy(),
                  r.counters())
    print(f"plain_gpu {spec.width}x{spec.height}, {spec.max_samples} spp, {slices} slices", flush=True)
    for a, b in zip(out[1][:5], out[slices][:5]):
        assert np.array_equal(bits(a), bits(b)) if a.dtype == np.float32 else np.array_equal(a, b)
    assert out[1][5] == out[slices][5]


def test_cdf_guide_tables_bit_identical(monkeypatch):
    """Environment-light sampling confined to the CDFs' bucket bounds picks the texels of the full
    binary searches: a textured-sky render with and without the tables (SHARPRAY_CDF_GUIDE=0 when
    the scene is built) gives bit-identical AOVs and counters."""
    from sharpray_amd import Renderer, scenes
    spec = scenes.config_plain_gpu(scale=0.25, spp=4)
    assert spec.sky_texture >= 0
    out = []
    for guide in ("1", "0"):
        monkeypatch.setenv("SHARPRAY_CDF_GUIDE", guide)
        scene, view = scenes.instantiate(spec)
        scene.build()
        r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                     spec.variance_threshold, spec.bounces, spec.indirect_clamp, counters=True)
        r.wait()
        out.append((r.radiance.copy(), r.radiance_fog.copy(), r.depth.copy(), r.normals.copy(), r.samples.copy(),
                    r.counters()))
    print(f"plain_gpu {spec.width}x{spec.height}, {spec.max_samples} spp, guide tables on / off", flush=True)
    for a, b in zip(out[0][:5], out[1][:5]):
        assert np.array_equal(bits(a), bits(b)) if a.dtype == np.float32 else np.array_equal(a, b)
    assert out[0][5] == out[1][5]


@pytest.mark.parametrize("case", ["cdf_counting", "alias_counting", "cdf_in_flight", "fog_grid"])
def test_sky_light_table_bit_identical(case, monkeypatch):
    """Light samples of a texture sky read their direction, pdf and radiance from a per-texel table
    (tabulated by the functions that evaluate them per sample): renders with and without it
    (SHARPRAY_SKY_TABLE=0 when the scene is built) give bit-identical AOVs and counters, for the CDF
    and alias samplers, the counter policy with samples in flight, and fog NEE."""
    from sharpray_amd import LIGHT_ALIAS, LIGHT_CDF, RNG_COUNTER, RNG_REFERENCE, Renderer, scenes
    spec = scenes.config_fog(scale=0.1, spp=4) if case == "fog_grid" else scenes.config_plain_gpu(scale=0.2, spp=8)
    assert spec.sky_texture >= 0
    counting = case.endswith("counting")
    sampler = LIGHT_ALIAS if case.startswith("alias") else LIGHT_CDF
    policy = RNG_REFERENCE if counting else RNG_COUNTER
    out = []
    for table in ("1", "0"):
        monkeypatch.setenv("SHARPRAY_SKY_TABLE", table)
        scene, view = scenes.instantiate(spec)
        scene.build()
        r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                     spec.variance_threshold, spec.bounces, spec.indirect_clamp, rng_policy=policy, rng_seed=5,
                     counters=counting, light_sampler=sampler)
        r.wait()
        out.append((r.radiance.copy(), r.radiance_fog.copy(), r.depth.copy(), r.normals.copy(), r.albedo.copy(),
                    r.samples.copy(), r.counters() if counting else None))
    print(f"{case}: {spec.width}x{spec.height}, {spec.max_samples} spp, sky table on / off", flush=True)
    assert np.abs(out[0][0]).sum() > 0
    for a, b in zip(out[0][:6], out[1][:6]):
        assert np.array_equal(bits(a), bits(b)) if a.dtype == np.float32 else np.array_equal(a, b)
    assert out[0][6] == out[1][6]


@pytest.mark.parametrize("size", [(1, 1), (7, 3), (13, 9), (9, 17)])
def test_odd_image_sizes_match_oracle(size, oracle_lib):
    """Images smaller than one 8x8 tile or with partial tiles on both edges: every pixel is rendered
    once (sample counts exact) and matches the oracle up to the divergent-path bar."""
    from sharpray_amd import RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=1.0, spp=4)
    spec.width, spec.height = size
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=9)
    r.wait()
    ref = oracle_lib.OracleScene(spec).render(rng_policy=1, rng_seed=9)
    assert r.radiance.shape == ref["radiance"].shape == (size[0] * size[1], 3)
    assert np.array_equal(r.samples, ref["samples"])
    div = int((rel_err(r.radiance, ref["radiance"]) > RAD_TOL).sum())
    print(f"{size}: divergent pixels {div}", flush=True)
    assert div <= max(1, int(MAX_DIVERGENT * size[0] * size[1]))
