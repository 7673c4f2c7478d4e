"""Float32 CPU restatement of the kernel-predicting U-Net denoiser.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the gfx950 bf16-MFMA denoiser; the
product never imports it. No counterpart in the reference (its denoiser is the bilateral
Compositor.Filter, compose.cs:101-204).

Arithmetic contract shared with the GPU: weights and every activation are rounded to bfloat16
(round to nearest even); a 3x3 convolution sums products in float32 and adds the bias, then ReLU,
then rounds; 2x2 average pooling adds ((a + b) + c) + d in float32, scales by 0.25 and rounds,
repeating the last row / column at odd sizes; upsampling is nearest; the final 25 logits are
soft-maxed (max-subtracted) per group of 25 and weight, in cascade, the edge-clamped 5x5
neighbourhoods (taps 1, 2, 4, 8 pixels apart) of the noisy radiance and of each stage's output.
Only the float32 summation order of the convolutions differs from the GPU's.
"""
import numpy as np

# (cinA, cinB, cout, mode, srcA, srcB, level, relu); mode 0 same resolution, 1 pool srcA,
# 2 concat(up(srcA), srcB); source -1 = input features.
LAYERS = [
    (10, 0, 32, 0, -1, -1, 0, True), (32, 0, 32, 0, 0, -1, 0, True),
    (32, 0, 48, 1, 1, -1, 1, True), (48, 0, 48, 0, 2, -1, 1, True),
    (48, 0, 64, 1, 3, -1, 2, True), (64, 0, 64, 0, 4, -1, 2, True),
    (64, 0, 96, 1, 5, -1, 3, True), (96, 0, 96, 0, 6, -1, 3, True),
    (96, 0, 128, 1, 7, -1, 4, True), (128, 0, 128, 0, 8, -1, 4, True),
    (128, 96, 96, 2, 9, 7, 3, True), (96, 0, 96, 0, 10, -1, 3, True),
    (96, 64, 64, 2, 11, 5, 2, True), (64, 0, 64, 0, 12, -1, 2, True),
    (64, 48, 48, 2, 13, 3, 1, True), (48, 0, 48, 0, 14, -1, 1, True),
    (48, 32, 32, 2, 15, 1, 0, True), (32, 0, 32, 0, 16, -1, 0, True),
    (32, 0, 100, 0, 17, -1, 0, False),
]
STAGES = 4  # kernel cascade: four 5x5 kernels, taps 1, 2, 4 and 8 pixels apart


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def exact(x):
    """No rounding: the float64 mode used to pin the backward pass by finite differences."""
    return np.asarray(x, np.float64)


def split_weights(blob, rnd=bf16):
    """[(w (cout, cin, 3, 3), b (cout,))] per layer, weights rounded by rnd (bf16 by default)."""
    out, k = [], 0
    for cin_a, cin_b, cout, *_ in LAYERS:
        cin = cin_a + cin_b
        w = np.asarray(blob[k:k + cout * cin * 9]).reshape(cout, cin, 3, 3)
        k += cout * cin * 9
        b = np.asarray(blob[k:k + cout])
        k += cout
        out.append((rnd(w), b.astype(np.float32) if rnd is bf16 else exact(b)))
    assert k == len(blob)
    return out


def features(rad, alb, nrm, dep, rnd=bf16):
    """(h, w, 10) input features from (h, w, 3) radiance, albedo, normals and (h, w) depth."""
    f = np.concatenate([np.log1p(np.maximum(rad, 0.0)), alb, nrm,
                        np.where(np.isinf(dep), 0.0, 1.0 / (1.0 + dep))[..., None]], axis=-1)
    return rnd(f.astype(np.float32) if rnd is bf16 else f)


def _cols(x):
    h, wd, cin = x.shape
    xp = np.zeros((h + 2, wd + 2, cin), x.dtype)
    xp[1:-1, 1:-1] = x
    return np.concatenate([xp[ky:ky + h, kx:kx + wd] for ky in range(3) for kx in range(3)], axis=-1)


def conv3x3(x, w, b, relu, rnd=bf16):
    """Zero-padded 3x3 cross-correlation of x (h, w, cin) with w (cout, cin, 3, 3), float32 sums."""
    h, wd, cin = x.shape
    wm = w.transpose(2, 3, 1, 0).reshape(9 * cin, -1)  # (tap, cin) x cout
    y = _cols(x).reshape(h * wd, 9 * cin) @ wm + b
    if relu:
        y = np.maximum(y, 0.0)
    y = y.reshape(h, wd, -1)
    return rnd(y.astype(np.float32) if rnd is bf16 else y)


def pool(x, rnd=bf16):
    h, wd, _ = x.shape
    h2, w2 = (h + 1) // 2, (wd + 1) // 2
    y0 = np.arange(h2) * 2
    x0 = np.arange(w2) * 2
    y1 = np.minimum(y0 + 1, h - 1)
    x1 = np.minimum(x0 + 1, wd - 1)
    s = ((x[y0][:, x0] + x[y0][:, x1]) + x[y1][:, x0]) + x[y1][:, x1]
    return rnd((s * np.float32(0.25)).astype(np.float32) if rnd is bf16 else s * 0.25)


def up(x, h, wd):
    return x[np.arange(h) // 2][:, np.arange(wd) // 2]


def staged_input(i, inputs, rnd=bf16):
    """The input of layer i's convolution (after pooling or upsampling + concatenation)."""
    cin_a, cin_b, cout, mode, sa, sb, level, relu = LAYERS[i]
    a = inputs[sa]
    if mode == 0:
        return a
    if mode == 1:
        return pool(a, rnd)
    skip = inputs[sb]
    return np.concatenate([up(a, skip.shape[0], skip.shape[1]), skip], axis=-1)


def layer_forward(i, wb, inputs, rnd=bf16):
    """Layer i from its input activations (inputs: dict index -> (h, w, c); -1 = features)."""
    w, b = wb[i]
    return conv3x3(staged_input(i, inputs, rnd), w, b, LAYERS[i][7], rnd)


def _stage_weights(logits, stage):
    dt = np.float64 if logits.dtype == np.float64 else np.float32
    l = logits[..., 25 * stage:25 * stage + 25].astype(dt)
    e = np.exp(l - l.max(-1, keepdims=True))
    return e / e.sum(-1, keepdims=True)


def _taps(img, stage):
    h, wd, _ = img.shape
    d = 1 << stage
    ys, xs = np.arange(h), np.arange(wd)
    return [(np.clip(ys + d * (k // 5 - 2), 0, h - 1), np.clip(xs + d * (k % 5 - 2), 0, wd - 1)) for k in range(25)]


def kpn_stage(logits, img, stage):
    dt = np.float64 if logits.dtype == np.float64 else np.float32
    w = _stage_weights(logits, stage)
    acc = np.zeros(img.shape, dt)
    for k, (ys, xs) in enumerate(_taps(img, stage)):
        acc += w[..., k:k + 1] * img[ys][:, xs]
    return acc


def kpn_apply(logits, rad):
    """The three cascade stages; returns the final image."""
    return kpn_cascade(logits, rad)[-1]


def kpn_cascade(logits, rad):
    """[input, after stage 0, after stage 1, after stage 2]."""
    out = [rad]
    for st in range(STAGES):
        out.append(kpn_stage(logits, out[-1], st))
    return out


def kpn_cascade_back(logits, rad, g):
    """Gradient of the loss with respect to the 75 logits given g = dL/d(final image)."""
    imgs = kpn_cascade(logits.astype(np.float64), np.asarray(rad, np.float64))
    dl = np.zeros(logits.shape[:2] + (25 * STAGES,))
    for st in range(STAGES - 1, -1, -1):
        src = imgs[st]
        w = _stage_weights(logits.astype(np.float64), st)
        taps = _taps(src, st)
        dw = np.stack([(g * src[ys][:, xs]).sum(-1) for ys, xs in taps], -1)
        dl[..., 25 * st:25 * st + 25] = w * (dw - (w * dw).sum(-1, keepdims=True))
        gin = np.zeros_like(src)
        for k, (ys, xs) in enumerate(taps):
            np.add.at(gin, (ys[:, None], xs[None, :]), w[..., k:k + 1] * g)
        g = gin
    return dl


def run(blob, rad, alb, nrm, dep, rnd=bf16):
    """Whole network: returns (list of layer activations, denoised radiance (h, w, 3))."""
    wb = split_weights(blob, rnd)
    acts = {-1: features(rad, alb, nrm, dep, rnd)}
    for i in range(len(LAYERS)):
        acts[i] = layer_forward(i, wb, acts, rnd)
    return acts, kpn_apply(acts[len(LAYERS) - 1], rad.astype(np.float32) if rnd is bf16 else exact(rad))


def loss(out, target):
    """mean (out - t)^2 / (t^2 + 0.01) over pixels and channels."""
    return float(np.mean((out - target) ** 2 / (target ** 2 + 0.01)))


def backward(blob, rad, alb, nrm, dep, target, rnd=bf16):
    """Loss and its gradient with respect to every layer's (w, b), by reverse-mode differentiation
    of the network above. With rnd = bf16 the gradients are rounded where the GPU rounds them
    (the logit gradient, every layer's masked output gradient and every data gradient)."""
    wb = split_weights(blob, rnd)
    acts = {-1: features(rad, alb, nrm, dep, rnd)}
    staged = {}
    for i in range(len(LAYERS)):
        staged[i] = staged_input(i, acts, rnd)
        w, b = wb[i]
        acts[i] = conv3x3(staged[i], w, b, LAYERS[i][7], rnd)
    r = rad.astype(np.float32) if rnd is bf16 else exact(rad)
    last = len(LAYERS) - 1
    out = kpn_cascade(acts[last].astype(np.float64), r.astype(np.float64))[-1]
    t = np.asarray(target, np.float64)
    n = out.size
    g = 2.0 * (out - t) / (t ** 2 + 0.01) / n
    dl = kpn_cascade_back(acts[last], r, g)
    gm = (lambda v: bf16(v.astype(np.float32))) if rnd is bf16 else exact
    dact = {i: None for i in range(-1, len(LAYERS))}
    dact[last] = gm(dl)
    grads = [None] * len(LAYERS)
    for i in range(last, -1, -1):
        cin_a, cin_b, cout, mode, sa, sb, level, relu = LAYERS[i]
        dy = dact[i].astype(np.float64)
        if relu:
            dy = dy * (acts[i] > 0)
        dy = gm(dy).astype(np.float64)
        x = staged[i].astype(np.float64)
        hh, ww, cin = x.shape
        dwm = _cols(x).reshape(hh * ww, 9 * cin).T @ dy.reshape(hh * ww, cout)  # (tap, cin) x cout
        grads[i] = (dwm.reshape(3, 3, cin, cout).transpose(3, 2, 0, 1), dy.sum((0, 1)))
        if i == 0:
            break
        w = wb[i][0].astype(np.float64)
        wt = w[:, :, ::-1, ::-1].transpose(1, 0, 2, 3)  # flipped, (cin, cout, 3, 3)
        dx = gm(conv3x3(dy, wt, np.zeros(cin), False, exact)).astype(np.float64)
        def add(k, v):
            dact[k] = v if dact[k] is None else dact[k] + v
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
            da = np.zeros((lh, lw, cin_a))
            np.add.at(da, (np.arange(hh)[:, None] // 2, np.arange(ww)[None, :] // 2), dx[..., :cin_a])
            add(sa, da)
            add(sb, dx[..., cin_a:])
    lossv = float(np.mean((out - t) ** 2 / (t ** 2 + 0.01)))
    return lossv, grads


def flatten(grads):
    """Per-layer (dW, db) in the weight blob's order."""
    return np.concatenate([np.concatenate([gw.reshape(-1), gb.reshape(-1)]) for gw, gb in grads])
