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

from tests.helpers import bits, rel_err

pytestmark = pytest.mark.gpu

RAD_TOL = 1e-3
MAX_DIVERGENT = 0.02


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
    assert frac <= MAX_DIVERGENT
    ok = ~div
    assert np.array_equal(r.samples, ref["samples"])
    assert np.all(r.samples == spec.max_samples + 1)  # the reference's M/(M+1) quirk (renderer.cs:173,215)
    assert np.array_equal(np.isinf(r.depth), np.isinf(ref["depth"]))
    fin = ok & np.isfinite(ref["depth"])
    assert np.allclose(r.depth[fin], ref["depth"][fin], rtol=1e-5, atol=0)
    assert np.allclose(r.normals[ok], ref["normals"][ok], atol=1e-5)
    assert np.allclose(r.uv[ok], ref["uv"][ok], atol=1e-5)
    assert np.all(rel_err(r.variance[ok, None], ref["variance"][ok, None]) <= 1e-2)


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
    assert frac <= MAX_DIVERGENT


def test_radiance_parity_analytic_shapes(oracle_lib):
    from sharpray_amd import scenes
    spec = scenes.config_analytic(scale=1.0, spp=4)
    scene, r, osc, ref = _render_both(spec, oracle_lib)
    err = rel_err(r.radiance, ref["radiance"])
    frac = float((err > RAD_TOL).mean())
    print(f"analytic radiance: divergent fraction {frac:.5f}", flush=True)
    assert frac <= MAX_DIVERGENT


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
