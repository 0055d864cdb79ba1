"""The learned constitutive surrogate on the GPU: MFMA forward and gradients against the FP64 CPU reference
(oracle/mlp.py), the batched plain material updates against the oracle's material point, training on plain-model
samples, the residual check with plain fallback, and the surrogate inside an assembly.

Tolerances:
  forward, vs the BF16-emulating reference : max |dy| <= 1e-2 of the output scale (accumulation order may flip a BF16
                                             rounding of a hidden activation now and then)
  gradient, vs the BF16-emulating reference: relative 2-norm error <= 2e-2
  gradient, vs plain FP64                  : cosine >= 0.99 (BF16 activations and weights)
  plain batched update vs FP64 oracle      : relative max error <= 1e-5 (FP32 arithmetic)
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

J2_PROPS = (200.0, 0.3, 0.25, 2.0)


def _asm(fec, gpu, physics, n=2):
    wl = fec.workloads.j2_hex8(n)
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    p = fec.create_parameters(wl.mesh, asm, physics)
    return wl, asm, p


def _scaling_random(g, d0, dout):
    return (g.uniform(-0.1, 0.1, d0).astype(np.float32), g.uniform(0.5, 2.0, d0).astype(np.float32),
            g.uniform(-0.1, 0.1, dout).astype(np.float32), g.uniform(0.5, 2.0, dout).astype(np.float32))


@pytest.mark.parametrize("arch", [0, 1])
def test_forward_and_gradient_match_reference(fec, mlpref, gpu, arch):
    print(f"[surrogate] forward/gradient arch {arch}", flush=True)
    physics = fec.J2Surrogate(*J2_PROPS) if arch == 0 else fec.NeoHookeanSurrogate(1.0, 0.5)
    _, asm, _ = _asm(fec, gpu, physics)
    net = fec.Surrogate(asm, arch, seed=5)
    g = np.random.default_rng(7)
    d = net.shape
    sc = _scaling_random(g, d[0], d[-1])
    net.set_scaling(*sc)
    n = 300  # not a multiple of 32 or 256: padded tiles and chunks
    raw = g.normal(0, 1, (n, 16 if arch == 0 else 9)).astype(np.float32)
    tgt = g.normal(0, 1, (n, d[-1])).astype(np.float32)
    w = net.params()
    assert w.size == net.nparams
    X, T = torch.from_numpy(raw).to(gpu), torch.from_numpy(tgt).to(gpu)
    Y = net.forward(X).cpu().numpy().astype(np.float64)
    yref, _ = mlpref.forward(arch, w, sc, raw, emulate=True)
    yref = yref * sc[3] + sc[2]
    scale = np.max(np.abs(yref))
    assert np.max(np.abs(Y - yref)) <= 1e-2 * scale, np.max(np.abs(Y - yref)) / scale
    loss, grad = net.gradient(X, T)
    grad = grad.cpu().numpy().astype(np.float64)
    lref, gref = mlpref.loss_and_grad(arch, w, sc, raw, tgt, emulate=True)
    l64, g64 = mlpref.loss_and_grad(arch, w, sc, raw, tgt, emulate=False)
    assert abs(loss - lref) <= 1e-3 * lref
    err = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    cos = float(grad @ g64 / (np.linalg.norm(grad) * np.linalg.norm(g64)))
    print(f"[surrogate] arch {arch}: loss {loss:.6f} (ref {lref:.6f}, fp64 {l64:.6f}) grad rel err {err:.2e} "
          f"cos(fp64) {cos:.6f}", flush=True)
    assert err <= 2e-2
    assert cos >= 0.99
    # set_params round trip and deterministic gradients
    net.set_params(w)
    loss2, grad2 = net.gradient(X, T)
    assert loss2 == loss and torch.equal(grad2.cpu(), torch.from_numpy(grad.astype(np.float32)))


def test_plain_batched_updates_match_oracle(fec, orc, gpu):
    print("[surrogate] plain batched J2 / NH vs oracle", flush=True)
    _, asm, _ = _asm(fec, gpu, fec.J2Plasticity(*J2_PROPS))
    X, T = fec.surrogate.j2_generate_samples(asm, 64, steps=20, max_norm=0.01, seed=3)
    Y = fec.surrogate.material_update(asm, fec._lib.J2, X)
    # the generator and the batched update run the same model (in two kernels)
    assert (Y - T).abs().max().item() <= 1e-6 * T.abs().max().item()
    Xh, Yh = X.cpu().numpy().astype(np.float64), Y.cpu().numpy().astype(np.float64)
    worst = 0.0
    for r in range(0, Xh.shape[0], 37):
        P, _, _, sn = orc.material_point(orc.J2, J2_PROPS, 3, 3, Xh[r, :9].reshape(3, 3), state_old=Xh[r, 9:16])
        sig = np.array([P[0], P[3], P[6], P[4], P[7], P[8]])
        ref = np.concatenate([sig, sn - Xh[r, 9:16]])
        worst = max(worst, np.max(np.abs(Yh[r] - ref)) / max(np.max(np.abs(sig)), 1e-3))
    assert worst <= 1e-5, worst
    # neo-Hookean
    _, asm2, _ = _asm(fec, gpu, fec.NeoHookean(1.0, 0.5))
    g = np.random.default_rng(1)
    H = (0.15 * g.normal(0, 1, (200, 9))).astype(np.float32)
    Yn = fec.surrogate.material_update(asm2, fec._lib.NEOHOOKEAN, torch.from_numpy(H).to(gpu)).cpu().numpy()
    for r in range(0, 200, 13):
        P, _, _, _ = orc.material_point(orc.NEOHOOKEAN, (1.0, 0.5), 3, 3, H[r].astype(np.float64).reshape(3, 3))
        F = np.eye(3) + H[r].astype(np.float64).reshape(3, 3)
        Ci = np.linalg.inv(F.T @ F)
        ref = np.concatenate([P, [Ci[0, 0] - 1, Ci[1, 0], Ci[2, 0], Ci[1, 1] - 1, Ci[2, 1], Ci[2, 2] - 1]])
        assert np.max(np.abs(Yn[r] - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref))), r


def test_training_and_fallback(fec, gpu):
    """Trained on plain-model samples, the network passes the residual check on most held-out samples; with a random
    network (almost) every sample falls back; fallback samples are the plain model's output."""
    print("[surrogate] J2 training", flush=True)
    _, asm, _ = _asm(fec, gpu, fec.J2Surrogate(*J2_PROPS))
    X, T = fec.surrogate.j2_generate_samples(asm, 4000, steps=20, max_norm=0.01, seed=1)
    Xv, Tv = fec.surrogate.j2_generate_samples(asm, 500, steps=20, max_norm=0.01, seed=2)
    net = fec.Surrogate(asm, fec.surrogate.MLP_J2, seed=0)
    net.fit_scaling(X, T)
    net.set_tolerance(0.05)
    Y0 = fec.surrogate.material_update(asm, fec._lib.J2_SURROGATE, Xv)
    fb0, n0 = net.fallbacks()
    losses = net.train(X, T, epochs=12, batch=2048, lr=2e-3, seed=0)
    print(f"[surrogate] losses {np.array2string(losses, precision=5)}", flush=True)
    assert losses[-1] < 0.1 * losses[0]
    Y1 = fec.surrogate.material_update(asm, fec._lib.J2_SURROGATE, Xv)
    fb1, n1 = net.fallbacks()
    print(f"[surrogate] fallbacks untrained {fb0}/{n0}, trained {fb1}/{n1}", flush=True)
    assert n0 == n1 == Xv.shape[0]
    assert fb1 < fb0
    # accepted samples: stress close to the plain model's
    plain = fec.surrogate.material_update(asm, fec._lib.J2, Xv)
    err = (Y1[:, :6] - plain[:, :6]).abs().max().item() / plain[:, :6].abs().max().item()
    print(f"[surrogate] max stress error of the surrogate update {err:.3e}", flush=True)
    assert err < 0.2
    # everything falls back under an impossible tolerance: the plain model's output
    net.set_tolerance(1e-12)
    Yf = fec.surrogate.material_update(asm, fec._lib.J2_SURROGATE, Xv)
    assert net.fallbacks()[0] == Xv.shape[0]
    assert (Yf - plain).abs().max().item() <= 1e-6 * plain.abs().max().item()


def test_surrogate_assembly_all_fallback_equals_plain(fec, gpu):
    """With every quadrature point falling back, the surrogate assembly (gradient pass, batched material pass, element
    kernel reading the per-point data) gives the plain assembly's residual and Jacobian (the same FP32 model in other
    kernels: equal to FP32 rounding)."""
    print("[surrogate] assembly all-fallback", flush=True)
    for plain_phys, surr_phys, arch in ((fec.J2Plasticity(*J2_PROPS), fec.J2Surrogate(*J2_PROPS), 0),
                                        (fec.NeoHookean(1.0, 0.5), fec.NeoHookeanSurrogate(1.0, 0.5), 1)):
        out = []
        for phys in (plain_phys, surr_phys):
            wl, asm, p = _asm(fec, gpu, phys, n=3)
            U = (wl.U if arch == 0 else 0.05 * wl.U / max(np.abs(wl.U).max(), 1e-30)).copy()
            Uu = torch.from_numpy(U[asm.dof.unknown_dofs - 1]).to(gpu)
            if phys is surr_phys:
                net = fec.Surrogate(asm, arch, seed=1)
                net.set_tolerance(1e-12)
            fec.assemble_vector_and_stiffness_(asm, Uu, p)
            out.append((fec.residual(asm).cpu().numpy().copy(), asm.stiffness_storage.cpu().numpy().copy()))
            if phys is surr_phys:
                nfb, npts = net.fallbacks()
                assert nfb == npts == wl.mesh.num_elements * 8
        for a, b in zip(out[0], out[1]):
            assert np.max(np.abs(a - b)) <= 1e-6 * np.max(np.abs(a)), np.max(np.abs(a - b)) / np.max(np.abs(a))


def test_newton_with_surrogate_reaches_plain_tolerance(fec, gpu):
    """J2 load steps solved by Newton + CG twice: with the plain model, and with the trained surrogate under the
    schedule (decide, freeze, hand over to the plain model when the surrogate iterations stall). Both reach the same
    relative residual tolerance on every step and the same solution."""
    print("[surrogate] Newton plain vs surrogate", flush=True)
    _, asm0, _ = _asm(fec, gpu, fec.J2Surrogate(*J2_PROPS))
    X, T = fec.surrogate.j2_generate_samples(asm0, 4000, steps=20, max_norm=0.01, seed=1)
    net0 = fec.Surrogate(asm0, fec.surrogate.MLP_J2, seed=0)
    net0.fit_scaling(X, T)
    net0.train(X, T, epochs=12, batch=2048, lr=2e-3, seed=0)
    params, scaling = net0.params(), net0.scaling()
    st = fec.NewtonSolverSettings(max_iters=15, rel_residual_tol=1e-5, abs_residual_tol=0.0, abs_increment_tol=0.0)
    out = {}
    for surr in (False, True):
        holder = {}

        def setup(asm):
            if surr:
                net = fec.Surrogate(asm, fec.surrogate.MLP_J2, seed=0)
                net.set_params(params)
                net.set_scaling(*scaling)
                net.set_tolerance(0.05)
                holder["s"] = fec.solvers.SurrogateSchedule(net)

        def schedule(n, h):
            if "s" in holder:
                holder["s"](n, h)

        asm, p, Uu, hists, _ = fec.solvers.j2_load_steps(fec, 6, steps=6, newton_settings=st, setup=setup,
                                                          surrogate=surr, schedule=schedule)
        out[surr] = (Uu.cpu().numpy().copy(), hists)
        print(f"[surrogate] {'surrogate' if surr else 'plain'}: iterations per step "
              f"{[len(h) for h in hists]}, final rel {[f'{h[-1].rel_R:.1e}' for h in hists]}", flush=True)
        assert all(h[-1].converged for h in hists)
    d = np.linalg.norm(out[True][0] - out[False][0]) / np.linalg.norm(out[False][0])
    print(f"[surrogate] solution difference {d:.2e}", flush=True)
    assert d < 1e-5
