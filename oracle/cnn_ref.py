"""Float32 CPU restatement of the kernel-predicting U-Net denoiser.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the gfx950 bf16-MFMA denoiser; the
product never imports it. No counterpart in the reference (its denoiser is the bilateral
Compositor.Filter, compose.cs:101-204).

Arithmetic contract shared with the GPU: weights and every activation are rounded to bfloat16
(round to nearest even); a 3x3 convolution sums products in float32 and adds the bias, then ReLU,
then rounds; 2x2 average pooling adds ((a + b) + c) + d in float32, scales by 0.25 and rounds,
repeating the last row / column at odd sizes; upsampling is nearest; the final 25 logits are
soft-maxed (max-subtracted) and weight the edge-clamped 5x5 neighbourhood of the noisy radiance.
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
    (32, 0, 25, 0, 17, -1, 0, False),
]


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def split_weights(blob):
    """[(w (cout, cin, 3, 3), b (cout,))] per layer, weights rounded to bf16."""
    out, k = [], 0
    for cin_a, cin_b, cout, *_ in LAYERS:
        cin = cin_a + cin_b
        w = np.asarray(blob[k:k + cout * cin * 9], np.float32).reshape(cout, cin, 3, 3)
        k += cout * cin * 9
        b = np.asarray(blob[k:k + cout], np.float32)
        k += cout
        out.append((bf16(w), b))
    assert k == len(blob)
    return out


def features(rad, alb, nrm, dep):
    """(h, w, 10) input features from (h, w, 3) radiance, albedo, normals and (h, w) depth."""
    f = np.concatenate([np.log1p(np.maximum(rad, 0.0)), alb, nrm,
                        np.where(np.isinf(dep), 0.0, 1.0 / (1.0 + dep))[..., None]], axis=-1)
    return bf16(f.astype(np.float32))


def conv3x3(x, w, b, relu):
    """Zero-padded 3x3 cross-correlation of x (h, w, cin) with w (cout, cin, 3, 3), float32 sums."""
    h, wd, cin = x.shape
    xp = np.zeros((h + 2, wd + 2, cin), np.float32)
    xp[1:-1, 1:-1] = x
    cols = np.concatenate([xp[ky:ky + h, kx:kx + wd] for ky in range(3) for kx in range(3)], axis=-1)
    wm = w.transpose(2, 3, 1, 0).reshape(9 * cin, -1)  # (tap, cin) x cout
    y = cols.reshape(h * wd, 9 * cin) @ wm + b
    if relu:
        y = np.maximum(y, 0.0)
    return bf16(y.reshape(h, wd, -1).astype(np.float32))


def pool(x):
    h, wd, _ = x.shape
    h2, w2 = (h + 1) // 2, (wd + 1) // 2
    y0 = np.arange(h2) * 2
    x0 = np.arange(w2) * 2
    y1 = np.minimum(y0 + 1, h - 1)
    x1 = np.minimum(x0 + 1, wd - 1)
    s = ((x[y0][:, x0] + x[y0][:, x1]) + x[y1][:, x0]) + x[y1][:, x1]
    return bf16((s * np.float32(0.25)).astype(np.float32))


def up(x, h, wd):
    return x[np.arange(h) // 2][:, np.arange(wd) // 2]


def layer_forward(i, wb, inputs):
    """Layer i from its input activations (inputs: dict index -> (h, w, c); -1 = features)."""
    cin_a, cin_b, cout, mode, sa, sb, level, relu = LAYERS[i]
    a = inputs[sa]
    if mode == 0:
        x = a
    elif mode == 1:
        x = pool(a)
    else:
        skip = inputs[sb]
        x = np.concatenate([up(a, skip.shape[0], skip.shape[1]), skip], axis=-1)
    w, b = wb[i]
    return conv3x3(x, w, b, relu)


def kpn_apply(logits, rad):
    h, wd, _ = rad.shape
    l = logits[..., :25].astype(np.float32)
    e = np.exp(l - l.max(-1, keepdims=True))
    s = np.zeros((h, wd), np.float32)
    acc = np.zeros((h, wd, 3), np.float32)
    ys, xs = np.arange(h), np.arange(wd)
    for k in range(25):
        dy, dx = k // 5 - 2, k % 5 - 2
        nb = rad[np.clip(ys + dy, 0, h - 1)][:, np.clip(xs + dx, 0, wd - 1)]
        s += e[..., k]
        acc += e[..., k:k + 1] * nb
    return acc / s[..., None]


def run(blob, rad, alb, nrm, dep):
    """Whole network: returns (list of layer activations, denoised radiance (h, w, 3))."""
    wb = split_weights(blob)
    acts = {-1: features(rad, alb, nrm, dep)}
    for i in range(len(LAYERS)):
        acts[i] = layer_forward(i, wb, acts)
    return acts, kpn_apply(acts[len(LAYERS) - 1], rad.astype(np.float32))
