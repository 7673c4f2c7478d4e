"""sr_compose_device_noalloc (no counterpart in the reference): the allocation-free compose on caller
device arrays gives the bytes sr_compose gives on the same render, on two calls in a row, the
second on a smaller image; images above its scratch size are refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_aovs(scene):
    from sharpray_amd import _native as N
    ptrs = [ctypes.c_void_p() for _ in range(7)]
    N.check(N.lib().sr_render_device_aovs(scene.ctx, *[ctypes.byref(p) for p in ptrs]), "device aovs")
    return [p.value for p in ptrs]  # radiance, fog radiance, normals, uv, depth, samples, variance


def _render(scale, seed):
    from sharpray_amd import RNG_COUNTER, scenes
    spec = scenes.config_plain_cpu(scale=scale, spp=4)
    spec.fog = (0.35, (0.9, 0.85, 0.8), 0.3, 0.6)
    scene, view, r = scenes.render(spec, rng_policy=RNG_COUNTER, rng_seed=seed)
    r.wait()
    return scene, r, spec.width, spec.height


def _hip():
    """The HIP runtime the library itself loaded (the caller's allocations for the output)."""
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                return ctypes.CDLL(line.split()[-1])
    raise RuntimeError("the HIP runtime is not loaded")


def _noalloc(scene, comp, W, H):
    from sharpray_amd import _native as N
    hip = _hip()
    n = W * H * 3
    out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), ctypes.c_size_t(n)) == 0
    try:
        assert hip.hipMemset(out, 7, ctypes.c_size_t(n)) == 0
        rad, fog, nrm, _, dep, _, var = _device_aovs(scene)
        N.check(N.lib().sr_compose_device_noalloc(scene.ctx, ctypes.byref(comp.compose_params), rad, fog, nrm, dep,
                                                  var, W, H, out), "compose noalloc")
        assert hip.hipDeviceSynchronize() == 0
        host = np.empty(n, np.uint8)
        assert hip.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), out, ctypes.c_size_t(n), 2) == 0  # device to host
    finally:
        hip.hipFree(out)
    return host.reshape(-1, 3)


def test_noalloc_compose_matches_compose_on_two_calls():
    from sharpray_amd import TONEMAP_LINEAR_SMOOTH, Compositor
    comp = Compositor(0.01, 0.1, 1.75, 0.25, 2.0, 0.125, 0.2, denoise_fog_radius=0.01, denoise_fog_strength=0.5,
                      tonemapper=TONEMAP_LINEAR_SMOOTH, exposure=0.8)
    for scale, seed in ((0.5, 3), (0.25, 4)):  # 128 x 128, then a smaller 64 x 64 image
        scene, r, W, H = _render(scale, seed)
        assert float(np.abs(r.radiance_fog).max()) > 0
        want = comp.compose(r).pixels.reshape(-1, 3)
        got = _noalloc(scene, comp, W, H)
        print(f"noalloc compose {W}x{H}: bytes off {(got != want).mean():.6f}", flush=True)
        assert np.array_equal(got, want)


def test_noalloc_compose_rejects_images_above_its_scratch():
    from sharpray_amd import Compositor, _native as N
    scene, r, W, H = _render(0.125, 5)
    rad, fog, nrm, _, dep, _, var = _device_aovs(scene)
    comp = Compositor(0.01, 0.1, 1.75, 0.25, 2.0, 0.125, 0.2)
    # 2048 x 1025 pixels exceed the 2,097,152-pixel scratch: refused before any launch (the
    # pointers are never read).
    rc = N.lib().sr_compose_device_noalloc(scene.ctx, ctypes.byref(comp.compose_params), rad, fog, nrm, dep, var,
                                           2048, 1025, ctypes.c_void_p(rad))
    assert rc == -4 and b"2,097,152" in N.lib().sr_last_error()
