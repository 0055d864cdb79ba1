"""GPU linear algebra on the assembled CSR Jacobian and the Newton loop (src/Solvers.jl:95-220): SpMV and CG against
scipy on the same CSR, the 64 x 64 quad4 Poisson problem solved end to end on the GPU against its analytic solution, and Newton
on J2 load steps.

Tolerances: SpMV relative 1e-14 (FP64, different summation order); CG relative error vs a direct solve <= 1e-7
(stopping rule ||r|| <= 1e-10 ||b||); 64 x 64 Poisson max nodal error vs the analytic solution <= 1e-3 (Q1 discretisation error at
h = 1/64 is ~2e-4)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _poisson(fec, n):
    wl = fec.workloads.poisson_quad4(n)
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=0)
    wl.mesh.nodeset_nodes["_bc"] = wl.dirichlet_dofs
    bcs = [fec.DirichletBC("u", lambda X, t: 0.0, nset_name="_bc")]
    p = fec.create_parameters(wl.mesh, asm, wl.physics, dirichlet_bcs=bcs)
    return wl, asm, p


def test_spmv_and_cg_match_scipy(fec, gpu):
    import scipy.sparse.linalg as spla
    print("[solvers] spmv / cg vs scipy", flush=True)
    wl, asm, p = _poisson(fec, 32)
    Uu = fec.create_unknowns(asm)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    K = fec.stiffness(asm)
    A = K.to_scipy()
    g = np.random.default_rng(0)
    x = g.normal(0, 1, A.shape[0])
    y = torch.zeros_like(Uu)
    asm.ctx.call("fecx_csr_spmv", fec.assembler._ptr(K.values), fec.assembler._ptr(torch.from_numpy(x).to(gpu)),
                 fec.assembler._ptr(y))
    yr = A @ x
    assert np.max(np.abs(y.cpu().numpy() - yr)) <= 1e-14 * np.max(np.abs(yr)) * 10
    b = fec.residual(asm).clone()
    ls = fec.IterativeLinearSolver(asm, "cg", fec.IterativeLinearSolverSettings(rtol=1e-10, atol=0.0))
    xg = torch.zeros_like(b)
    it, res = ls.cg(K.values, b, xg)
    xs = spla.spsolve(A.tocsc(), b.cpu().numpy())
    err = np.linalg.norm(xg.cpu().numpy() - xs) / np.linalg.norm(xs)
    print(f"[solvers] cg iterations {it}, residual {res:.3e}, rel error vs direct {err:.3e}", flush=True)
    assert 0 < it < 1000 and err <= 1e-7


def test_poisson_quad4_64_newton_cg_on_gpu(fec, gpu):
    """Poisson on 64 x 64 quad4 (f = 2 pi^2 sin sin, u = 0 on the boundary) assembled and solved on the GPU: one Newton step
    of a linear problem, CG to 1e-12; the nodal solution against sin(pi x) sin(pi y). The residual's floor is FP32
    element rounding: about 4e-5 of the initial residual here (nodal residuals are cancelling sums of O(1) element
    contributions), so Newton stops at a relative residual of 1e-4."""
    print("[solvers] Poisson 64x64 quad4 on the GPU", flush=True)
    wl, asm, p = _poisson(fec, 64)
    Uu = fec.create_unknowns(asm)
    ls = fec.IterativeLinearSolver(asm, "cg", fec.IterativeLinearSolverSettings(rtol=1e-12, atol=0.0))
    newton = fec.NewtonSolver(ls, fec.NewtonSolverSettings(max_iters=3, rel_residual_tol=1e-4))
    hist = newton.solve(Uu, p)
    print(f"[solvers] {[(h.iteration, h.norm_R, h.rel_R, h.cg_iters) for h in hist]}", flush=True)
    X = wl.mesh.coords
    exact = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
    u = np.zeros(wl.mesh.num_nodes)
    u[asm.dof.unknown_dofs - 1] = Uu.cpu().numpy()
    err = np.max(np.abs(u - exact))
    print(f"[solvers] Poisson 64x64 max nodal error {err:.3e}", flush=True)
    assert err <= 1e-3
    assert hist[-1].converged


def test_newton_j2_load_steps(fec, gpu):
    """Quasi-static J2 with the boundary driven along the strain path: every load step's Newton iteration reduces the
    residual by 1e-5 within 10 iterations. FP32 element kernels bound the attainable relative residual: with the
    load-step predictor the first residual is small and the floor lies at 1-3e-6 of it."""
    print("[solvers] Newton on J2 load steps", flush=True)
    st = fec.NewtonSolverSettings(max_iters=10, rel_residual_tol=1e-5, abs_residual_tol=0.0, abs_increment_tol=0.0)
    log = lambda k, h: print(f"[solvers] step {k}: " + ", ".join(f"{r.rel_R:.1e}/{r.cg_iters}" for r in h), flush=True)
    asm, p, Uu, hists, _ = fec.solvers.j2_load_steps(fec, 6, steps=6, newton_settings=st, log=log)
    for h in hists:
        assert h[-1].converged, [(r.rel_R, r.norm_R) for r in h]
    ns, nq = 7, 8
    state = np.zeros(6 ** 3 * nq * ns, dtype=np.float32)
    asm.ctx.call("fecx_get_state_new", state.ctypes.data_as(__import__("ctypes").c_void_p))
    eqps = state.reshape(-1, nq, ns)[..., 6]
    print(f"[solvers] plastic points {np.mean(eqps > 0):.3f}, max eqps {eqps.max():.3e}", flush=True)
    assert eqps.max() > 0.0  # the path goes plastic
