"""Pins of the oracle's environment light (SkyTexture + Cdf2) and joint bilateral filter against
independent float64 restatements written from the reference's definitions, so that a dropped
term, a flipped axis or a wrong index in the oracle fails here before it can hide a GPU mismatch.
"""

import numpy as np
import pytest

from oracle import oracle as O


def _sky_scene(L, tex, mult, yaw):
    h = L.oc_scene_new()
    rgb = np.ascontiguousarray(tex, np.float32)
    L.oc_scene_add_texture(h, O.f(rgb), rgb.shape[1], rgb.shape[0], 0, 1.0, 1.0)
    L.oc_scene_set_sky_texture(h, 0, mult, yaw)
    assert L.oc_scene_build(h) == 0
    return h, rgb


def _lum(c):
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]


def _yaw_rotate(d, yaw):
    """Rotation by `yaw` about +Y (Quat.AngleAxis(yaw, Up) applied to d)."""
    c, s = np.cos(yaw), np.sin(yaw)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    return np.stack([c * x + s * z, y, -s * x + c * z], -1)


def _equirect_uv(d):
    u = np.arctan2(d[..., 2], d[..., 0]) / (2 * np.pi) + 0.5
    v = 0.5 - np.arcsin(np.clip(d[..., 1], -1, 1)) / np.pi
    return u, v


def _bilinear(tex, u, v):
    """Texture.Sample with repeat wrap and the (size - 1) scaling of texture.cs:32-60."""
    H, W = tex.shape[:2]
    u = u - np.floor(u)
    v = v - np.floor(v)
    fx, fy = u * (W - 1), v * (H - 1)
    x0, y0 = fx.astype(np.int64), fy.astype(np.int64)
    x1, y1 = (x0 + 1) % W, (y0 + 1) % H
    tx, ty = (fx - x0)[..., None], (fy - y0)[..., None]
    top = tex[y0, x0] + (tex[y0, x1] - tex[y0, x0]) * tx
    bot = tex[y1, x0] + (tex[y1, x1] - tex[y1, x0]) * tx
    return top + (bot - top) * ty


@pytest.mark.parametrize("yaw", [0.0, 0.7])
def test_sky_radiance_and_mis_pdf_closed_form(oracle_lib, yaw):
    """SkyTexture.Radiance = bilinear(EquirectUv(R^-1 d)) * mult and LightDir's pdf = point-looked-up
    luminance * W*H / (total * 2 pi^2), total = sum lum * sin(pi (y + .5) / H) (scene.cs:389-425)."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(11)
    H, W, mult = 8, 16, 1.7
    tex = rng.uniform(0.05, 3.0, (H, W, 3)).astype(np.float32)
    h, _ = _sky_scene(L, tex, mult, yaw)
    try:
        d = rng.normal(size=(4000, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d = d.astype(np.float32)
        rad = np.empty((len(d), 3), np.float32)
        pdf = np.empty(len(d), np.float32)
        L.oc_sky_eval(h, O.f(d), len(d), O.f(rad), O.f(pdf))
        t64 = tex.astype(np.float64)
        local = _yaw_rotate(d.astype(np.float64), -yaw)
        u, v = _equirect_uv(local)
        expect = _bilinear(t64, u, v) * mult
        np.testing.assert_allclose(rad, expect, rtol=2e-4, atol=1e-5)
        total = float((_lum(t64) * np.sin(np.pi * (np.arange(H)[:, None] + 0.5) / H)).sum())
        scale = W * H / (total * 2 * np.pi ** 2)
        cx = (u * W).astype(np.int64) % W
        cy = (v * H).astype(np.int64) % H
        # Directions within float rounding of a texel edge may land in either texel.
        inner = (np.abs(u * W - np.round(u * W)) > 1e-3) & (np.abs(v * H - np.round(v * H)) > 1e-3)
        np.testing.assert_allclose(pdf[inner], (_lum(t64[cy, cx]) * scale)[inner], rtol=2e-5)
    finally:
        L.oc_scene_free(h)


@pytest.mark.parametrize("sampler", [0, 1])
def test_sky_light_sampling_pmf_chi_square(oracle_lib, sampler):
    """LightDirRand draws texel centres with probability lum * sin(theta) / total (Cdf2, cdf.cs:62-98;
    sampler 1: the alias table over the same weights) and returns pdf = lum * pdfScale: a chi-square
    test of 200k draws against the closed-form pmf."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(5)
    H, W = 6, 12
    tex = rng.uniform(0.0, 2.0, (H, W, 3)).astype(np.float32)
    tex[2, 3] = 0.0  # a black texel is never drawn
    yaw = 0.3
    h, _ = _sky_scene(L, tex, 1.0, yaw)
    try:
        n = 200_000
        out = np.empty((n, 4), np.float32)
        L.oc_sky_sample(h, 3, 9, n, sampler, O.f(out))
    finally:
        L.oc_scene_free(h)
    t64 = tex.astype(np.float64)
    w = _lum(t64) * np.sin(np.pi * (np.arange(H)[:, None] + 0.5) / H)
    pmf = w / w.sum()
    local = _yaw_rotate(out[:, :3].astype(np.float64), -yaw)
    u, v = _equirect_uv(local)
    x = np.floor(u * W).astype(np.int64) % W
    y = np.floor(v * H).astype(np.int64) % H
    # Every draw is a texel centre.
    np.testing.assert_allclose(u * W - x, 0.5, atol=1e-3)
    np.testing.assert_allclose(v * H - y, 0.5, atol=1e-3)
    counts = np.bincount(y * W + x, minlength=H * W).reshape(H, W)
    assert counts[2, 3] == 0
    mask = pmf > 0
    expect = pmf[mask] * n
    chi2 = float(((counts[mask] - expect) ** 2 / expect).sum())
    dof = int(mask.sum()) - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)
    scale = W * H / (w.sum() * 2 * np.pi ** 2)
    np.testing.assert_allclose(out[:, 3], _lum(t64[y, x]) * scale, rtol=2e-5)


def _filter_ref(params, rad, nrm, dep, var, W, H, alb=None, alim=0.0):
    """Compositor.Filter (compose.cs:101-204) restated per pixel in float64; alb adds the albedo
    guide term exp(-|da|^2 / (2 alim^2)) on surface pixels."""
    radius, strength, smax, lboost, llim, nlim, dlim = params
    rp = radius * np.sqrt(W * H)
    kr = int(np.ceil(rp * 3.0))
    out = rad.copy()
    R = rad.reshape(H, W, 3)
    N = nrm.reshape(H, W, 3)
    D = dep.reshape(H, W)
    V = var.reshape(H, W)
    lum = _lum(R)
    hasN = (N ** 2).sum(-1) > 0
    hasD = ~np.isinf(D)
    for cy in range(H):
        for cx in range(W):
            lb = 1.0 + lum[cy, cx] * lboost
            dw = min(V[cy, cx] * lb * strength, smax)
            if dw < 1e-4:
                continue
            wsum, rsum = 1.0, R[cy, cx].copy()
            for ky in range(-kr, kr + 1):
                for kx in range(-kr, kr + 1):
                    if (kx, ky) == (0, 0) or kx * kx + ky * ky > kr * kr:
                        continue
                    nx, ny = cx + kx, cy + ky
                    if not (0 <= nx < W and 0 <= ny < H):
                        continue
                    if hasN[cy, cx] != hasN[ny, nx] or hasD[cy, cx] != hasD[ny, nx]:
                        continue
                    w = dw * np.exp(-(kx * kx + ky * ky) / (2 * rp * rp))
                    if hasN[cy, cx]:
                        w *= np.exp(-((N[cy, cx] - N[ny, nx]) ** 2).sum() / (2 * nlim * nlim))
                    if hasD[cy, cx]:
                        w *= np.exp(-(D[cy, cx] - D[ny, nx]) ** 2 / (2 * dlim * dlim))
                    if alb is not None and hasN[cy, cx]:
                        A = alb.reshape(H, W, 3)
                        w *= np.exp(-((A[cy, cx] - A[ny, nx]) ** 2).sum() / (2 * alim * alim))
                    ld = max(0.0, lum[ny, nx] - lum[cy, cx])
                    w *= np.exp(-ld * ld / (2 * llim * llim))
                    wsum += w
                    rsum = rsum + R[ny, nx] * w
            out[cy * W + cx] = rsum / wsum
    return out


@pytest.mark.parametrize("guided", [False, True])
def test_bilateral_filter_oracle_vs_float64_restatement(oracle_lib, guided):
    rng = np.random.default_rng(2)
    W, H = 14, 11
    rad = rng.uniform(0, 2, (W * H, 3)).astype(np.float32)
    rad[5] *= 40.0  # a firefly
    nrm = rng.normal(size=(W * H, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[:W] = 0.0  # a row without surfaces: no normal, no depth
    nrm = nrm.astype(np.float32)
    dep = rng.uniform(1, 3, W * H).astype(np.float32)
    dep[:W] = np.inf
    var = rng.uniform(0, 0.3, W * H).astype(np.float32)
    var[7] = 0.0  # early out
    params = np.array([0.15, 2.0, 0.9, 0.5, 0.6, 0.5, 0.4], np.float32)
    alb = rng.uniform(0, 1, (W * H, 3)).astype(np.float32) if guided else None
    got, _ = O.bilateral(params, rad, nrm, dep, var, W, H, threads=2, albedo=alb, albedo_limit=0.3)
    ref = _filter_ref(params.astype(np.float64), rad.astype(np.float64), nrm.astype(np.float64),
                      dep.astype(np.float64), var.astype(np.float64), W, H,
                      None if alb is None else alb.astype(np.float64), 0.3)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-6)
    assert np.array_equal(got[7], rad[7])
