"""The surrogate network's FP64 CPU reference (oracle/mlp.py) pinned by central finite differences of its own loss,
and its BF16 emulation checked against the plain FP64 evaluation it approximates."""
import numpy as np
import pytest


def _problem(mlpref, arch, n=7, seed=3):
    g = np.random.default_rng(seed)
    d = mlpref.SHAPES[arch]
    nparams = sum(d[l + 1] * d[l] + d[l + 1] for l in range(len(d) - 1))
    w = np.concatenate([np.concatenate([g.normal(0, np.sqrt(2.0 / d[l]), d[l + 1] * d[l]), g.normal(0, 0.1, d[l + 1])])
                        for l in range(len(d) - 1)])
    assert w.size == nparams
    raw_in = 16 if arch == 0 else 9
    raw = g.normal(0, 1, (n, raw_in))
    tgt = g.normal(0, 1, (n, d[-1]))
    scaling = (g.uniform(-0.2, 0.2, d[0]), g.uniform(0.5, 2.0, d[0]), g.uniform(-0.2, 0.2, d[-1]),
               g.uniform(0.5, 2.0, d[-1]))
    return w, scaling, raw, tgt


@pytest.mark.parametrize("arch", [0, 1])
def test_reference_gradient_central_differences(mlpref, arch):
    w, scaling, raw, tgt = _problem(mlpref, arch)
    _, grad = mlpref.loss_and_grad(arch, w, scaling, raw, tgt)
    g = np.random.default_rng(11)
    d = mlpref.SHAPES[arch]
    # coordinates in every layer's W and b
    idx, o = [], 0
    for l in range(len(d) - 1):
        idx += list(o + g.integers(0, d[l + 1] * d[l], 6)) + list(o + d[l + 1] * d[l] + g.integers(0, d[l + 1], 3))
        o += d[l + 1] * d[l] + d[l + 1]
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mlpref.loss_only(arch, wp, scaling, raw, tgt) - mlpref.loss_only(arch, wm, scaling, raw, tgt)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_bf16_rounding_is_round_to_nearest_even(mlpref):
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0], dtype=np.float64)
    r = mlpref.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3 2^-8 is a tie -> even (1 + 2^-6 ... i.e. 1 + 4 2^-8)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2.0 ** -8 and r[3] == 1.0 and r[4] == -2.5 and r[5] == 0.0


@pytest.mark.parametrize("arch", [0, 1])
def test_emulated_forward_close_to_fp64(mlpref, arch):
    w, scaling, raw, _ = _problem(mlpref, arch, n=64)
    y, _ = mlpref.forward(arch, w, scaling, raw, emulate=False)
    ye, _ = mlpref.forward(arch, w, scaling, raw, emulate=True)
    assert np.max(np.abs(ye - y)) <= 0.05 * np.max(np.abs(y))
