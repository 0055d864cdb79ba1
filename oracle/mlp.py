"""TEST INFRASTRUCTURE ONLY: FP64 CPU forward/backward of the learned constitutive surrogate's MLP.

The surrogate has no counterpart in the reference. This module is the checker of the GPU training and inference
kernels (libfecx.so, FECX_MLP_J2 / FECX_MLP_NH), used only by tests/. It is pinned by central finite differences
of its own loss (tests/test_surrogate_cpu.py).

Network: standardized features -> [W_l h + b_l, GELU (tanh form)] hidden layers -> linear output. Loss: mean
over samples and outputs of the squared error of the standardized outputs. Parameters in canonical order: per layer
W (out x in, row-major) then b.

emulate=True reproduces the kernels' BF16 roundings (round-to-nearest-even), so GPU and CPU differ only by the FP32
vs FP64 accumulation:
  - the weights, the layer inputs h_l and the GELU derivatives;
  - the output-side gradients dz_l, both where the weight gradient reads them and where the backward GEMM does.
"""
import numpy as np

SHAPES = {0: (13, 128, 128, 128, 13), 1: (9, 256, 256, 15)}


def bf16(a):
    """Round to bfloat16 (round to nearest, ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def unflatten(arch, w):
    d = SHAPES[arch]
    out, o = [], 0
    for l in range(len(d) - 1):
        W = np.asarray(w[o:o + d[l + 1] * d[l]], dtype=np.float64).reshape(d[l + 1], d[l])
        o += d[l + 1] * d[l]
        b = np.asarray(w[o:o + d[l + 1]], dtype=np.float64)
        o += d[l + 1]
        out.append((W, b))
    assert o == len(w)
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def features(arch, raw):
    """Raw inputs -> features: J2 {sym(H) in the order 11,21,31,22,32,33, state (7)}; NH H."""
    raw = np.asarray(raw, dtype=np.float64)
    if arch == 0:
        H = raw[:, :9]
        eps = np.stack([H[:, 0], 0.5 * (H[:, 3] + H[:, 1]), 0.5 * (H[:, 6] + H[:, 2]), H[:, 4],
                        0.5 * (H[:, 7] + H[:, 5]), H[:, 8]], axis=1)
        return np.concatenate([eps, raw[:, 9:16]], axis=1)
    return raw[:, :9].copy()


GELU_U = 2.0 * np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def gelu(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) = z sigma(2 sqrt(2/pi) (z + 0.044715 z^3))."""
    return 0.5 * z * (1.0 + np.tanh(0.5 * GELU_U * (z + GELU_C * z ** 3)))


def gelu_prime(z):
    s = 0.5 * (1.0 + np.tanh(0.5 * GELU_U * (z + GELU_C * z ** 3)))
    return s + z * s * (1.0 - s) * GELU_U * (1.0 + 3.0 * GELU_C * z * z)


def forward(arch, w, scaling, raw, emulate=False):
    """Standardized outputs (n, dout) and the cache for backward."""
    in_shift, in_scale, _, _ = [np.asarray(s, dtype=np.float64) for s in scaling]
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(arch, w)
    x = (features(arch, raw) - in_shift) / in_scale
    if emulate:
        x = x.astype(np.float32).astype(np.float64)
    h = rnd(x)
    hs, gps = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + b
        if l == len(layers) - 1:
            return z, (layers, hs, gps)
        gps.append(rnd(gelu_prime(z)))
        h = rnd(gelu(z))
        hs.append(h)


def loss_and_grad(arch, w, scaling, raw, target, emulate=False):
    _, _, out_shift, out_scale = [np.asarray(s, dtype=np.float64) for s in scaling]
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (layers, hs, gps) = forward(arch, w, scaling, raw, emulate)
    t = (np.asarray(target, dtype=np.float64) - out_shift) / out_scale
    n, dout = t.shape
    loss = float(np.sum((y - t) ** 2) / (n * dout))
    dz = 2.0 * (y - t) / (n * dout)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dzr = rnd(dz)
        grads[l] = (dzr.T @ hs[l], dzr.sum(axis=0))
        if l > 0:
            dz = (dzr @ rnd(W)) * gps[l - 1]
    return loss, flatten(grads)


def loss_only(arch, w, scaling, raw, target):
    _, _, out_shift, out_scale = [np.asarray(s, dtype=np.float64) for s in scaling]
    y, _ = forward(arch, w, scaling, raw, False)
    t = (np.asarray(target, dtype=np.float64) - out_shift) / out_scale
    return float(np.mean((y - t) ** 2))
