"""Learned constitutive surrogate on libfecx.so (no counterpart in the reference, whose materials are energy functions
differentiated by AD inside the element loop, test/mechanics/TestMechanicsCommon.jl:140-222).

A small MLP (GELU hidden layers, BF16 MFMA with FP32 accumulation) is evaluated batched over quadrature points and
trained on the GPU by the library's own forward/backward kernels on samples the plain material model generates. In an
assembly with a surrogate physics (J2Surrogate, NeoHookeanSurrogate) every quadrature point's network output is
checked against the material equations; points that fail fall back to the plain model.

    FECX_MLP_J2: 13 -> 128 -> 128 -> 128 -> 13   raw in {H (9), state (7)}, raw out {sigma (6), d plastic (6), d eqps}
    FECX_MLP_NH: 9 -> 256 -> 256 -> 15           raw in H (9), raw out {P (9), C^-1 - I (6)}
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .assembler import J2Plasticity, NeoHookean, SparseMatrixAssembler, _host_ptr, _ptr

MLP_J2, MLP_NH = _lib.MLP_J2, _lib.MLP_NH
SHAPES = {MLP_J2: (13, 128, 128, 128, 13), MLP_NH: (9, 256, 256, 15)}
RAW_IN = {MLP_J2: 16, MLP_NH: 9}
RAW_OUT = {MLP_J2: 13, MLP_NH: 15}


class J2Surrogate(J2Plasticity):
    """J2 plasticity whose material update is the FECX_MLP_J2 network, with residual check and plain fallback."""
    kind = _lib.J2_SURROGATE


class NeoHookeanSurrogate(NeoHookean):
    """Neo-Hookean whose stress and tangent come from the FECX_MLP_NH network, with residual check and fallback."""
    kind = _lib.NH_SURROGATE


def _f32(t):
    assert t.dtype == torch.float32 and t.is_contiguous() and t.is_cuda
    return _ptr(t)


class Surrogate:
    """The network held by an assembler's libfecx context."""

    def __init__(self, asm: SparseMatrixAssembler, arch, seed=0):
        self.asm, self.arch = asm, arch
        self.ctx = asm.ctx
        self.ctx.call("fecx_surrogate_init", ctypes.c_int(arch), ctypes.c_uint64(seed))
        n = ctypes.c_int64(0)
        self.ctx.call("fecx_surrogate_num_params", ctypes.byref(n))
        self.nparams = n.value

    @property
    def shape(self):
        return SHAPES[self.arch]

    def params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        self.ctx.call("fecx_surrogate_get_params", _host_ptr(w))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparams
        self.ctx.call("fecx_surrogate_set_params", _host_ptr(w))

    def scaling(self):
        d = self.shape
        a = [np.zeros(d[0], np.float32), np.zeros(d[0], np.float32), np.zeros(d[-1], np.float32),
             np.zeros(d[-1], np.float32)]
        self.ctx.call("fecx_surrogate_get_scaling", *[_host_ptr(x) for x in a])
        return tuple(a)

    def set_scaling(self, in_shift, in_scale, out_shift, out_scale):
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (in_shift, in_scale, out_shift, out_scale)]
        self.ctx.call("fecx_surrogate_set_scaling", *[_host_ptr(x) for x in a])

    def fit_scaling(self, X, T):
        self.ctx.call("fecx_surrogate_fit_scaling", _f32(X), _f32(T), ctypes.c_int64(X.shape[0]))

    def forward(self, X):
        Y = torch.empty((X.shape[0], RAW_OUT[self.arch]), dtype=torch.float32, device=X.device)
        self.ctx.call("fecx_surrogate_forward", _f32(X), ctypes.c_int64(X.shape[0]), _f32(Y))
        return Y

    def gradient(self, X, T):
        g = torch.empty(self.nparams, dtype=torch.float32, device=X.device)
        loss = ctypes.c_double(0.0)
        self.ctx.call("fecx_surrogate_gradient", _f32(X), _f32(T), ctypes.c_int64(X.shape[0]), ctypes.byref(loss),
                      _f32(g))
        return loss.value, g

    def train(self, X, T, epochs, batch=4096, lr=1e-3, seed=0):
        losses = np.zeros(epochs, dtype=np.float64)
        self.ctx.call("fecx_surrogate_train", _f32(X), _f32(T), ctypes.c_int64(X.shape[0]), ctypes.c_int(epochs),
                      ctypes.c_int(batch), ctypes.c_float(lr), ctypes.c_uint64(seed), _host_ptr(losses))
        return losses

    def set_tolerance(self, tol):
        self.ctx.call("fecx_surrogate_set_tolerance", ctypes.c_float(tol))

    def freeze(self, on=True):
        """Reuse the last pass's per-point network/plain decisions (a fixed residual function during a solve)."""
        self.ctx.call("fecx_surrogate_freeze", ctypes.c_int(1 if on else 0))

    def enable(self, on=True):
        """Off: assemblies with the surrogate physics use the plain model."""
        self.ctx.call("fecx_surrogate_enable", ctypes.c_int(1 if on else 0))

    def fallbacks(self):
        nf, npts = ctypes.c_int64(0), ctypes.c_int64(0)
        self.ctx.call("fecx_surrogate_fallbacks", ctypes.byref(nf), ctypes.byref(npts))
        return nf.value, npts.value


def material_update(asm, kind, X):
    """Batched material update of n raw inputs (plain kind J2/NEOHOOKEAN, or the surrogate kinds with check and
    fallback); uses the properties of the assembler's material."""
    arch = MLP_J2 if kind in (_lib.J2, _lib.J2_SURROGATE) else MLP_NH
    Y = torch.empty((X.shape[0], RAW_OUT[arch]), dtype=torch.float32, device=X.device)
    asm.ctx.call("fecx_material_update", ctypes.c_int(kind), _f32(X), ctypes.c_int64(X.shape[0]), _f32(Y))
    return Y


def qp_gradients(asm, U):
    """Displacement gradients (ne, nq, 9) at every quadrature point for the full FP64 field U on the device."""
    ne = asm.var.fspace.mesh.num_elements
    ns, nq = ctypes.c_int(0), ctypes.c_int(0)
    asm.ctx.call("fecx_get_state_size", ctypes.byref(ns), ctypes.byref(nq))
    H = torch.empty((ne, nq.value, 9), dtype=torch.float32, device=U.device)
    asm.ctx.call("fecx_qp_gradients", _ptr(U), _f32(H))
    return H


def j2_generate_samples(asm, npaths, steps=20, max_norm=0.01, seed=0):
    """Training samples of the plain J2 model along random macroscopic strain paths: X (n, 16), T (n, 13)."""
    n = npaths * steps
    X = torch.empty((n, 16), dtype=torch.float32, device=asm.device)
    T = torch.empty((n, 13), dtype=torch.float32, device=asm.device)
    asm.ctx.call("fecx_j2_generate_samples", ctypes.c_int64(npaths), ctypes.c_int(steps), ctypes.c_float(max_norm),
                 ctypes.c_uint64(seed), _f32(X), _f32(T))
    return X, T
