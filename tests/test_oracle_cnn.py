"""Pins of the denoiser's CPU restatement (oracle/cnn_ref.py) against direct loops and closed forms."""
import numpy as np

from oracle import cnn_ref as C


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3, 65504.0], np.float32)
    expect = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, np.float32(3.0e-3), 65536.0], np.float32)
    got = C.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2 ** -6) and got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8 and got[5] == expect[5]


def test_conv3x3_matches_direct_loops():
    rng = np.random.default_rng(0)
    h, w, cin, cout = 5, 7, 6, 4
    x = C.bf16(rng.normal(size=(h, w, cin)).astype(np.float32))
    wt = C.bf16(rng.normal(size=(cout, cin, 3, 3)).astype(np.float32))
    b = rng.normal(size=cout).astype(np.float32)
    got = C.conv3x3(x, wt, b, relu=True)
    ref = np.zeros((h, w, cout))
    for y in range(h):
        for xx in range(w):
            for co in range(cout):
                s = float(b[co])
                for ky in range(3):
                    for kx in range(3):
                        yy, xk = y + ky - 1, xx + kx - 1
                        if 0 <= yy < h and 0 <= xk < w:
                            s += float(np.dot(x[yy, xk].astype(np.float64), wt[co, :, ky, kx].astype(np.float64)))
                ref[y, xx, co] = max(s, 0.0)
    ulp = np.maximum(np.abs(ref), 1e-30) * 2.0 ** -7
    assert np.all(np.abs(got - ref) <= ulp + 1e-6)


def test_pool_and_up():
    x = C.bf16(np.arange(5 * 3 * 2, dtype=np.float32).reshape(5, 3, 2))
    p = C.pool(x)
    assert p.shape == (3, 2, 2)
    # (0,0): mean of rows 0-1, cols 0-1; (2,1): last row and column repeated.
    assert np.allclose(p[0, 0], (x[0, 0] + x[0, 1] + x[1, 0] + x[1, 1]) / 4)
    assert np.allclose(p[2, 1], x[4, 2], rtol=2 ** -7)
    u = C.up(p, 5, 3)
    assert u.shape == (5, 3, 2) and np.array_equal(u[4, 2], p[2, 1]) and np.array_equal(u[1, 1], p[0, 0])


def _box_loop(img, d):
    h, w, _ = img.shape
    ref = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            acc = 0
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    acc = acc + img[min(max(y + d * dy, 0), h - 1), min(max(x + d * dx, 0), w - 1)]
            ref[y, x] = acc / 25
    return ref


def test_kpn_apply_identity_and_box():
    """One-hot logits on the centre tap of every stage are the identity; zero logits make each
    stage an edge-clamped 5x5 box filter with taps 1, 2, 4 and 8 pixels apart, applied in order."""
    rng = np.random.default_rng(1)
    rad = rng.uniform(0, 3, (9, 11, 3)).astype(np.float32)
    logits = np.full((9, 11, 128), -30.0, np.float32)
    logits[..., [12, 37, 62, 87]] = 30.0
    assert np.allclose(C.kpn_apply(logits, rad), rad, rtol=1e-6)
    stages = C.kpn_cascade(np.zeros((9, 11, 128), np.float32), rad)
    assert len(stages) == 5
    ref = rad.astype(np.float64)
    for st in range(4):
        ref = _box_loop(ref, 1 << st)
        assert np.allclose(stages[st + 1], ref, rtol=1e-5), st
    # Only stage 1's kernel shifted: a one-hot tap at (dy, dx) = (-1, 2) moves the image by
    # (-2, 4) pixels (edge-clamped) and the other stages stay the identity.
    logits[..., 37] = -30.0
    logits[..., 25 + 1 * 5 + 4] = 30.0
    out = C.kpn_apply(logits, rad)
    ys, xs = np.clip(np.arange(9) - 2, 0, 8), np.clip(np.arange(11) + 4, 0, 10)
    assert np.allclose(out, rad[ys][:, xs], rtol=1e-6)


def test_kpn_cascade_backward_matches_finite_differences():
    """The cascade's logit gradient against central differences of a float64 weighted sum of its
    output, for every one of the 100 logits at a few pixels."""
    rng = np.random.default_rng(4)
    h, w = 7, 9
    rad = rng.uniform(0, 2, (h, w, 3))
    logits = rng.normal(0, 1.5, (h, w, 128))
    g = rng.normal(size=(h, w, 3))
    dl = C.kpn_cascade_back(logits, rad, g)

    def f(lg):
        return float((C.kpn_apply(lg, rad) * g).sum())

    for (y, x) in [(0, 0), (3, 4), (6, 8), (2, 7)]:
        for c in range(100):
            eps = 1e-6
            hi, lo = logits.copy(), logits.copy()
            hi[y, x, c] += eps
            lo[y, x, c] -= eps
            fd = (f(hi) - f(lo)) / (2 * eps)
            assert abs(fd - dl[y, x, c]) <= 1e-6 * max(1.0, abs(fd)), (y, x, c, fd, dl[y, x, c])


def test_weight_count_matches_library():
    import sharpray_amd as A
    n = sum(co * (ca + cb) * 9 + co for ca, cb, co, *_ in C.LAYERS)
    assert n == A.cnn_weight_count() == A.KpnDenoiser.random_weights(0).size
    assert [tuple(l) for l in C.LAYERS] == [tuple(l) for l in A.CNN_LAYERS]


def _example(seed, h=12, w=10):
    rng = np.random.default_rng(seed)
    rad = rng.uniform(0, 2, (h, w, 3))
    alb = rng.uniform(0, 1, (h, w, 3))
    nrm = rng.normal(size=(h, w, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    dep = rng.uniform(1, 4, (h, w))
    dep[0, :3] = np.inf
    target = rng.uniform(0, 2, (h, w, 3))
    return rad, alb, nrm, dep, target


def test_backward_matches_finite_differences_in_float64():
    """The reverse-mode restatement (float64, no rounding) against central differences of the
    float64 loss, for a sample of weights and biases in every layer."""
    import sharpray_amd as A
    rad, alb, nrm, dep, target = _example(0)
    blob = A.KpnDenoiser.random_weights(3).astype(np.float64)
    rng = np.random.default_rng(1)
    blob += rng.normal(0, 0.01, blob.size)  # non-zero biases
    loss0, grads = C.backward(blob, rad, alb, nrm, dep, target, rnd=C.exact)
    flat = C.flatten(grads)

    def f(b):
        _, out = C.run(b, rad, alb, nrm, dep, rnd=C.exact)
        return C.loss(out, target)

    assert abs(f(blob) - loss0) <= 1e-12 * max(1.0, loss0)
    k = 0
    picks = []
    for cin_a, cin_b, cout, *_ in C.LAYERS:
        nw = cout * (cin_a + cin_b) * 9
        picks += list(k + rng.choice(nw, 3, replace=False)) + [k + nw + int(rng.integers(cout))]
        k += nw + cout
    for i in picks:
        eps = 2e-5  # truncation O(eps^2) stays far below the tolerance; smaller eps loses digits
        hi, lo = blob.copy(), blob.copy()
        hi[i] += eps
        lo[i] -= eps
        fd = (f(hi) - f(lo)) / (2 * eps)
        assert abs(fd - flat[i]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(flat).max()), (i, fd, flat[i])
