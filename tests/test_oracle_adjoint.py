"""Pins of the oracle's adjoint (gradient) renderer.

- The BSDF partial derivatives against central differences of OpaqueEval.
- Albedo gradients against central finite differences of the float64 forward render with common
  random numbers: for metallic 0 or 1 and no Russian roulette, sampling does not depend on albedo,
  so the two agree to finite-difference accuracy.
- Roughness and sky-texel gradients move the sampling too, so the check is statistical: over
  independent seeds, the mean adjoint gradient and the mean finite difference of the rendered
  image must agree within 4 standard errors.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sharpray_amd import scenes as S


def test_bsdf_derivatives_match_central_differences(oracle_lib):
    L = oracle_lib.lib("f64")
    rng = np.random.default_rng(0)
    m = 300
    v = rng.normal(size=(m, 3)); v[:, 1] = np.abs(v[:, 1]) + 0.1
    l = rng.normal(size=(m, 3)); l[:, 1] = np.abs(l[:, 1]) + 0.1
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    l /= np.linalg.norm(l, axis=1, keepdims=True)
    vl = np.concatenate([v, l], 1).astype(np.float32)
    n = np.array([0, 1, 0], np.float32)
    t = np.array([1, 0, 0, 1], np.float32)
    for rough, metal in [(0.5, 0.0), (0.3, 1.0), (0.8, 0.4)]:
        mat = np.array([0.7, 0.4, 0.2, rough, metal, 0.0, 1.5, 0, 0, 0], np.float32)
        g = np.empty((m, 6), np.float32)
        L.oc_surface_eval_grad(O.f(mat), O.f(n), O.f(n), O.f(t), O.f(vl), m, O.f(g))

        def ev(mt):
            o = np.empty((m, 4), np.float32)
            L.oc_surface_eval(O.f(mt), O.f(n), O.f(n), O.f(t), O.f(vl), m, O.f(o))
            return o[:, :3].astype(np.float64)

        e = 1e-3
        hi, lo = mat.copy(), mat.copy()
        hi[3] += e
        lo[3] -= e
        fdr = (ev(hi) - ev(lo)) / (2 * e)
        scale = max(1.0, float(np.abs(fdr).max()))
        assert np.abs(fdr - g[:, 3:]).max() <= 1e-4 * scale
        for c in range(3):
            hi, lo = mat.copy(), mat.copy()
            hi[c] += e
            lo[c] -= e
            fda = ((ev(hi) - ev(lo)) / (2 * e))[:, c]
            assert np.abs(fda - g[:, c]).max() <= 1e-4 * max(1.0, float(np.abs(fda).max()))


@pytest.fixture(scope="module")
def small():
    spec = S.config_adjoint_small(32, 24, 16, 2)
    return spec, O.OracleScene(spec, precision="f64")


def _loss(osc, p, delta, seed):
    osc.set_params(p)
    img = osc.render(rng_policy=1, rng_seed=seed)["radiance"].astype(np.float64)
    return float((img * delta).sum())


def test_albedo_gradients_match_crn_finite_differences(small):
    spec, osc = small
    rng = np.random.default_rng(1)
    delta = rng.normal(size=(spec.width * spec.height, 3)).astype(np.float32)
    p0 = osc.params()
    g = osc.adjoint(delta, seed=5)
    idx = [o * 4 + c for o in range(len(spec.objects)) for c in range(3)]
    gmax = float(np.abs(g[idx]).max())
    for i in idx:
        eps = 1e-3
        hi, lo = p0.copy(), p0.copy()
        hi[i] += eps
        lo[i] -= eps
        fd = (_loss(osc, hi, delta, 5) - _loss(osc, lo, delta, 5)) / (2 * eps)
        assert abs(g[i] - fd) <= 2e-3 * max(abs(fd), 0.05 * gmax), (i, g[i], fd)
    osc.set_params(p0)


def test_roughness_and_sky_gradients_unbiased(small):
    spec, osc = small
    spec2 = S.config_adjoint_small(32, 24, 64, 2)
    osc2 = O.OracleScene(spec2, precision="f64")
    delta = np.ones((spec2.width * spec2.height, 3), np.float32)
    p0 = osc2.params()
    nObj = len(spec2.objects)
    g0 = osc2.adjoint(delta, seed=99)
    env = np.argsort(-np.abs(g0[nObj * 4:]))[:5] + nObj * 4
    params = [o * 4 + 3 for o in range(nObj)] + list(env)
    K = 16
    adj = np.zeros((K, len(params)))
    fd = np.zeros((K, len(params)))
    for k in range(K):
        osc2.set_params(p0)
        g = osc2.adjoint(delta, seed=200 + k)
        for j, i in enumerate(params):
            eps = 0.02 if i < nObj * 4 else 0.25 * max(abs(float(p0[i])), 0.05)
            hi, lo = p0.copy(), p0.copy()
            hi[i] += eps
            lo[i] -= eps
            adj[k, j] = g[i]
            fd[k, j] = (_loss(osc2, hi, delta, 200 + k) - _loss(osc2, lo, delta, 200 + k)) / (2 * eps)
    osc2.set_params(p0)
    se = np.sqrt(adj.var(0, ddof=1) / K + fd.var(0, ddof=1) / K)
    z = (adj.mean(0) - fd.mean(0)) / np.maximum(se, 1e-12)
    print("z scores", np.round(z, 2))
    assert np.all(np.abs(z) < 4.0), z
