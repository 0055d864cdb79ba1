"""Parity at the configurations' full sizes, through properties that do not depend on the size (the FP64 oracle is
checked at bounded sizes elsewhere):
  - the Jacobian is symmetric: x . (K y) = y . (K x) for random x, y (all three physics here are hyperelastic or
    linear), checked with the library's SpMV on the assembled CSR;
  - Poisson: K 1 = 0 on an all-free mesh (constants are in the null space); the residual is affine in u:
    R(u1 + u2) = R(u1) + R(u2) - R(0);
  - the matrix-free action and the assembled action both equal K v (SpMV of the assembled Jacobian);
  - assembly and the assembled action are deterministic: repeated calls give the same bits (no atomics).
Tolerances (relative to the norms involved): symmetry 1e-12 (FP64 SpMV of an exactly symmetric matrix, summation
order only), null space 1e-6 (FP32 element sums), linearity 1e-6, actions 2e-6."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _asm(fec, wl):
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=0)
    p = fec.create_parameters(wl.mesh, asm, wl.physics)
    return asm, p


def _spmv(fec, asm, vals, x):
    y = torch.empty_like(x)
    asm.ctx.call("fecx_csr_spmv", fec.assembler._ptr(vals), fec.assembler._ptr(x), fec.assembler._ptr(y))
    return y


@pytest.mark.parametrize("name,n", [("poisson_hex8", 128), ("neohookean_hex8", 96), ("neohookean_tet10", 80)])
def test_full_size_properties(fec, gpu, name, n):
    print(f"[fullsize] {name} {n}", flush=True)
    wl = fec.workloads.CONFIGS[name](n)
    asm, p = _asm(fec, wl)
    Uu = torch.from_numpy(wl.U[asm.dof.unknown_dofs - 1]).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    K = asm.stiffness_storage.clone()
    R = fec.residual(asm).clone()
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    assert torch.equal(K, asm.stiffness_storage) and torch.equal(R, fec.residual(asm)), "not deterministic"
    g = torch.Generator(device=gpu)
    g.manual_seed(1)
    x = torch.randn(Uu.numel(), dtype=torch.float64, device=gpu, generator=g)
    y = torch.randn(Uu.numel(), dtype=torch.float64, device=gpu, generator=g)
    Kx, Ky = _spmv(fec, asm, K, x), _spmv(fec, asm, K, y)
    a, b = float(x @ Ky), float(y @ Kx)
    print(f"[fullsize] symmetry {a:.12e} vs {b:.12e}", flush=True)
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))
    # actions: matrix-free and assembled equal K v
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, x, p)
    mf = fec.hvp(asm, x).clone()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, x, p)
    am = fec.hvp(asm, x).clone()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, x, p)
    assert torch.equal(am, fec.hvp(asm, x)), "assembled action not deterministic"
    nk = float(torch.linalg.norm(Kx))
    e_mf, e_am = float(torch.linalg.norm(mf - Kx)) / nk, float(torch.linalg.norm(am - Kx)) / nk
    print(f"[fullsize] action vs K v: matrix-free {e_mf:.2e}, assembled {e_am:.2e}", flush=True)
    assert e_mf <= 2e-6 and e_am <= 2e-6
    if name == "poisson_hex8":
        ones = torch.ones_like(x)
        K1 = _spmv(fec, asm, K, ones)
        null = float(torch.linalg.norm(K1)) / float(torch.linalg.norm(Kx))
        print(f"[fullsize] |K 1| / |K x| = {null:.2e}", flush=True)
        assert null <= 1e-6
        z = torch.zeros_like(Uu)
        fec.assemble_vector_(asm, fec.residual, z, p)
        R0 = fec.residual(asm).clone()
        fec.assemble_vector_(asm, fec.residual, x, p)
        Rx = fec.residual(asm).clone()
        fec.assemble_vector_(asm, fec.residual, y, p)
        Ry = fec.residual(asm).clone()
        fec.assemble_vector_(asm, fec.residual, x + y, p)
        Rxy = fec.residual(asm).clone()
        lin = float(torch.linalg.norm(Rxy - Rx - Ry + R0)) / float(torch.linalg.norm(Rxy))
        print(f"[fullsize] residual linearity defect {lin:.2e}", flush=True)
        assert lin <= 1e-6
