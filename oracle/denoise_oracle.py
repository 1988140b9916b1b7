"""CPU oracle of the learned denoiser (include/pt_denoise.h).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the HIP network's
forward pass, loss and parameter gradients.  The denoiser has no counterpart
in the reference, so this restatement is pinned by the mathematics instead:
its analytic backward pass is checked against central finite differences in
float64 (tests/test_denoise.py), which a flipped tap, a transposed weight, a
misrouted skip or pool gradient would fail.

Network (the same 11 layers as pt_denoise.hip): 3x3 same-padding
convolutions with bias, ReLU except the last; 2x2 average pooling between
encoder levels; decoder layers read [upsample2(deeper output), encoder output]
concatenated.  Parameters are one flat float array, per layer W[co][tap][ci]
(tap = 3*(dy+1) + (dx+1), ci over the first input's then the second input's
channels) followed by b[co].

With rounding=True the forward pass rounds to bf16 where the device does
(weights, the input features, every hidden output, pooled tensors, the
gradients fed to each backward layer), so the two agree to accumulation
order; with rounding=False everything is float64 and differentiable.
"""
import numpy as np

# (input a, input b or None, cin_a, cin_b, cout, relu, upsample a, level)
LAYERS = [
    ("x", None, 9, 0, 32, True, False, 0),
    ("y0", None, 32, 0, 32, True, False, 0),
    ("p1", None, 32, 0, 48, True, False, 1),
    ("p2", None, 48, 0, 64, True, False, 2),
    ("p3", None, 64, 0, 80, True, False, 3),
    ("p4", None, 80, 0, 96, True, False, 4),
    ("y5", "y4", 96, 80, 80, True, True, 3),
    ("y6", "y3", 80, 64, 64, True, True, 2),
    ("y7", "y2", 64, 48, 48, True, True, 1),
    ("y8", "y1", 48, 32, 32, True, True, 0),
    ("y9", None, 32, 0, 3, False, False, 0),
]
LOG_CLAMP = 12.0
RELMSE_EPS = 0.01


def parameter_count():
    return sum(co * 9 * (ca + cb) + co for _, _, ca, cb, co, _, _, _ in LAYERS)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(params):
    out, o = [], 0
    for _, _, ca, cb, co, _, _, _ in LAYERS:
        nw = co * 9 * (ca + cb)
        out.append((params[o:o + nw].reshape(co, 9, ca + cb), params[o + nw:o + nw + co]))
        o += nw + co
    return out


def conv(x, w, b):
    """3x3 same convolution, x (n,h,w,ci), w (co,9,ci) -> (n,h,w,co)."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, h, wd, w.shape[0]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :] @ w[:, t, :].T
    return y + b


def conv_backward(x, w, gy):
    """Gradients of conv with respect to its input, weights and bias."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    g2 = gy.reshape(-1, gy.shape[-1])
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        xs = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :].reshape(-1, x.shape[-1])
        gw[:, t, :] = g2.T @ xs
        gxp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :] += gy @ w[:, t, :]
    return gxp[:, 1:-1, 1:-1, :], gw, g2.sum(axis=0)


def pool(x):
    return 0.25 * (((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2])


def pool_backward(g):
    return 0.25 * g.repeat(2, axis=1).repeat(2, axis=2)


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up_backward(g):
    return g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]


def input_features(features, rounding=True):
    """(n,h,w,9) linear radiance, albedo, normal -> network input (log1p radiance)."""
    f = np.asarray(features, np.float64)
    rad = np.log1p(np.maximum(f[..., :3], 0).astype(np.float32)).astype(np.float64)
    x = np.concatenate([rad, f[..., 3:9]], axis=-1)
    return bf16(x) if rounding else x


def forward(params, features, rounding=True):
    """Raw network output (n,h,w,3) and the cache for backward."""
    rnd = bf16 if rounding else (lambda v: v)
    layers = unpack(np.asarray(params, np.float64))
    t = {"x": input_features(features, rounding)}
    cache = []
    for k, ((a, b, ca, cb, co, relu, up_a, level), (w, bias)) in enumerate(zip(LAYERS, layers)):
        if k in (2, 3, 4, 5):
            t[f"p{k - 1}"] = rnd(pool(t[f"y{k - 1}"]))
        xa = up(t[a]) if up_a else t[a]
        xin = np.concatenate([xa, t[b]], axis=-1) if b else xa
        y = conv(xin, rnd(w), bias)
        if relu:
            y = rnd(np.maximum(y, 0))
        cache.append((xin, rnd(w)))
        t[f"y{k}"] = y
    return t["y10"], (t, cache)


def loss(out, x_log, target):
    """The device loss and its gradient with respect to the raw output."""
    y = out + x_log
    t = np.maximum(np.asarray(target, np.float64), 0)
    lt = np.log1p(t)
    yc = np.minimum(y, LOG_CLAMP)
    e = np.expm1(yc)
    den = t * t + RELMSE_EPS
    n = y.size
    value = (np.abs(y - lt) + (e - t) ** 2 / den).sum() / n
    grad = np.sign(y - lt) + np.where(y < LOG_CLAMP, 2 * (e - t) / den * (e + 1), 0)
    return value, grad / n


def backward(params, state, gout, rounding=True):
    """Flat parameter gradient for dL/d(out) = gout."""
    rnd = bf16 if rounding else (lambda v: v)
    t, cache = state
    grads = [None] * len(LAYERS)
    gy = {f"y{k}": np.zeros_like(t[f"y{k}"]) for k in range(10)}
    for k in range(len(LAYERS) - 1, -1, -1):
        a, b, ca, cb, co, relu, up_a, level = LAYERS[k]
        if k == len(LAYERS) - 1:
            g = rnd(gout)
        else:
            g = rnd(np.where(t[f"y{k}"] > 0, gy[f"y{k}"], 0))
        xin, w = cache[k]
        gx, gw, gb = conv_backward(xin, w, g)
        grads[k] = np.concatenate([gw.reshape(-1), gb])
        if k == 0:
            break
        if b:
            ga = gx[..., :ca]
            gy[a] += up_backward(ga) if up_a else ga
            gy[b] += gx[..., ca:]
        elif a.startswith("p"):
            gy[f"y{int(a[1:])}"] += pool_backward(gx)
        else:
            gy[a] += gx
    return np.concatenate(grads)


def denoise(params, features, rounding=True):
    """Denoised linear radiance (n,h,w,3)."""
    out, (t, _) = forward(params, features, rounding)
    y = np.minimum(out + t["x"][..., :3], LOG_CLAMP)
    return np.maximum(np.expm1(y), 0)


def loss_and_gradients(params, features, targets, rounding=True):
    out, state = forward(params, features, rounding)
    value, gout = loss(out, state[0]["x"][..., :3], targets)
    return value, backward(params, state, gout, rounding)
