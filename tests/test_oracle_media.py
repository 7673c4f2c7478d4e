"""Pins of the oracle's bounded media (homogeneous box, half-precision density grid) against what
the mathematics fixes: closed-form transmittance of a homogeneous slab, brute-force quadrature of a
grid's optical depth, and the free-flight distribution of delta tracking (a Kolmogorov-Smirnov test
against 1 - exp(-tau(t)))."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

BMIN = np.array([-1.0, -0.5, 0.0])
BMAX = np.array([2.0, 1.5, 2.5])


def _scene(L, kind, sigma, grid=None):
    h = L.oc_scene_new()
    sky = np.ones((4, 8, 3), np.float32)
    L.oc_scene_add_texture(h, O.f(sky), 8, 4, 0, 1.0, 1.0)
    L.oc_scene_set_sky_texture(h, 0, 1.0, 0.0)
    res = np.array([0, 0, 0] if grid is None else grid.shape[::-1], np.uint32)
    bits = None if grid is None else np.ascontiguousarray(grid, np.float16).view(np.uint16)
    L.oc_scene_set_fog_volume(h, kind, sigma, O.f(np.array([0.9, 0.9, 0.9], np.float32)), 0.3,
                              O.f(BMIN.astype(np.float32)), O.f(BMAX.astype(np.float32)), res.ctypes.data_as(O._U),
                              None if bits is None else bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    assert L.oc_scene_build(h) == 0
    return h, bits


def _rays(rng, n):
    o = rng.uniform(-3, 4, (n, 3))
    target = rng.uniform(BMIN, BMAX, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def _span(o, d, start, end):
    """Box overlap [a, b) of each segment, float64."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / np.where(np.abs(d) >= 1e-8, d, np.copysign(1e-8, d))
    t1 = (BMIN - o) * inv
    t2 = (BMAX - o) * inv
    lo = np.minimum(t1, t2).max(1)
    hi = np.maximum(t1, t2).min(1)
    return np.maximum(lo, start), np.minimum(hi, end)


def test_homogeneous_optical_depth_closed_form(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(1)
    h, _ = _scene(L, 1, 0.37)
    try:
        o, d = _rays(rng, 3000)
        se = np.stack([rng.uniform(0, 2, 3000), rng.uniform(2, 9, 3000)], 1).astype(np.float32)
        se[::7, 1] = np.inf
        out = np.empty(3000, np.float32)
        L.oc_medium_optical_depth(h, O.f(np.concatenate([o, d], 1)), O.f(se), 3000, O.f(out))
    finally:
        L.oc_scene_free(h)
    a, b = _span(o.astype(np.float64), d.astype(np.float64), se[:, 0].astype(np.float64), se[:, 1].astype(np.float64))
    expect = 0.37 * np.maximum(0.0, b - a)
    np.testing.assert_allclose(out, expect, rtol=1e-4, atol=2e-5)


def _grid(rng, shape=(5, 6, 8)):
    g = rng.uniform(0.0, 1.0, shape).astype(np.float16)
    g[1, 2, :] = 0.0
    return g


def _grid_tau(grid, sigma, o, d, a, ts):
    """tau(a, t) for each t in ts by midpoint quadrature with nearest-voxel lookups (float64)."""
    nz, ny, nx = grid.shape
    cell = (BMAX - BMIN) / np.array([nx, ny, nz])
    g = grid.astype(np.float64)
    m = 20000
    t = a + (np.arange(m) + 0.5) / m * (ts.max() - a)
    p = o[None, :] + t[:, None] * d[None, :]
    idx = np.clip(np.floor((p - BMIN) / cell).astype(np.int64), 0, [nx - 1, ny - 1, nz - 1])
    s = sigma * g[idx[:, 2], idx[:, 1], idx[:, 0]]
    cum = np.concatenate([[0.0], np.cumsum(s) * (ts.max() - a) / m])
    grid_t = np.concatenate([[a], t + 0.5 * (ts.max() - a) / m])
    return np.interp(ts, grid_t, cum)


def test_grid_optical_depth_matches_quadrature(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(2)
    grid = _grid(rng)
    sigma = 0.8
    h, _ = _scene(L, 2, sigma, grid)
    try:
        n = 60
        o, d = _rays(rng, n)
        se = np.stack([np.zeros(n), np.full(n, np.inf)], 1).astype(np.float32)
        out = np.empty(n, np.float32)
        L.oc_medium_optical_depth(h, O.f(np.concatenate([o, d], 1)), O.f(se), n, O.f(out))
    finally:
        L.oc_scene_free(h)
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    a, b = _span(o64, d64, 0.0, np.inf)
    for i in range(n):
        if not a[i] < b[i]:
            assert out[i] == 0.0
            continue
        ref = _grid_tau(grid, sigma, o64[i], d64[i], a[i], np.array([b[i]]))[0]
        assert abs(out[i] - ref) <= 2e-3 * max(ref, 1e-3), (i, out[i], ref)


@pytest.mark.parametrize("kind", [1, 2])
def test_free_flight_distribution_ks(oracle_lib, kind):
    """Scatter distances (and the escape mass) follow P(T <= t) = 1 - exp(-tau(a, t)); for the grid
    this checks delta tracking against the voxel quadrature."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(3)
    grid = _grid(rng) if kind == 2 else None
    sigma = 0.9
    h, _ = _scene(L, kind, sigma, grid)
    o = np.array([-2.0, 0.3, 0.9])
    d = np.array([1.0, 0.15, 0.4])
    d /= np.linalg.norm(d)
    n = 40000
    out = np.empty(n, np.float32)
    try:
        L.oc_medium_scatter(h, O.f(np.concatenate([o, d]).astype(np.float32)), 0.0, np.inf, 5, 6, n, O.f(out))
    finally:
        L.oc_scene_free(h)
    a, b = _span(o[None], d[None], 0.0, np.inf)
    a, b = float(a[0]), float(b[0])
    hits = np.sort(out[out >= 0].astype(np.float64))
    assert np.all((hits >= a - 1e-5) & (hits < b + 1e-5))
    ts = np.linspace(a, b, 400)
    if kind == 1:
        cdf = 1.0 - np.exp(-sigma * (ts - a))
    else:
        cdf = 1.0 - np.exp(-_grid_tau(grid, sigma, o, d, a, ts))
    emp = np.searchsorted(hits, ts, side="right") / n
    ks = float(np.abs(emp - cdf).max())
    assert ks < 1.63 / np.sqrt(n) * 1.5, ks
    escape = float((out < 0).mean())
    assert abs(escape - (1 - cdf[-1])) < 4 * np.sqrt(cdf[-1] * (1 - cdf[-1]) / n) + 1e-3
