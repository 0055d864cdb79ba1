"""Plane-strain mechanics on quad4 meshes with two displacement fields: the reference's Mechanics(PlaneStrain())
(test/mechanics/TestMechanics.jl, test/mechanics/TestMechanicsCommon.jl; src/Formulations.jl:318-427).

CPU (oracle pins): the 2D oracle against the pinned 3D oracle on the one-element-thick hex8 slab extruded from the
same quad4 mesh. With in-plane displacements equal on both layers and zero normal displacement the slab is in plane
strain, so its residual summed over a node's two layers, and its stiffness summed over both layers' node pairs, equal
the 2D ones (thickness 1: the 2 x 2 x 2 rule integrates the constant z-direction exactly); and a linear displacement
field leaves zero residual at interior nodes (patch test).
GPU: residual, CSR Jacobian, both actions and the mass matrix against the oracle within 1e-6 on a structured quad4
mesh and on the reference's mechanics_coarse.g; and the reference's mechanics regression (nset_3 fixed, nset_1 with
displ_y = 1e-3 t) solved by Newton + CG on the GPU against the oracle's direct solve."""
import os

import numpy as np
import pytest
import torch

from test_gpu_parity import rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
QUADS = os.path.join(ROOT, "tests", "golden", "mechanics_coarse.g")
PROPS = [1e3, 10.0e9, 1.0e9]  # rho, K, G: create_properties(::Mechanics), TestMechanicsCommon.jl:7-12


def _jittered_quads(fec, n, seed=0):
    m = fec.StructuredMesh("quad", (0.0, 0.0), (1.0, 1.0), (n + 1, n + 1))
    g = np.random.default_rng(seed)
    X = m.coords.copy()
    inner = (X[:, 0] > 1e-12) & (X[:, 0] < 1 - 1e-12) & (X[:, 1] > 1e-12) & (X[:, 1] < 1 - 1e-12)
    X[inner] += g.uniform(-0.15 / n, 0.15 / n, (inner.sum(), 2))
    m.coords = X
    return m


def _extrude(X2, conn2):
    """One layer of hex8 over the quads, nodes of layer z at z = 0 and 1: node a + nn z."""
    nn = X2.shape[0]
    X3 = np.vstack([np.hstack([X2, np.zeros((nn, 1))]), np.hstack([X2, np.ones((nn, 1))])])
    conn3 = np.hstack([conn2, conn2 + nn]).astype(np.int64)
    return X3, conn3


def test_plane_strain_oracle_equals_slab(fec, orc):
    m = _jittered_quads(fec, 5)
    X2, conn2 = m.coords, m.conn.astype(np.int64)
    nn = X2.shape[0]
    g = np.random.default_rng(1)
    u2 = 1e-3 * g.normal(0, 1, (nn, 2))
    R2 = orc.assemble_residual(orc.QUAD4, X2, conn2, 2, orc.LINEAR_ELASTIC, PROPS, u2.ravel())
    X3, conn3 = _extrude(X2, conn2)
    u3 = np.zeros((2 * nn, 3))
    u3[:, :2] = np.vstack([u2, u2])
    R3 = orc.assemble_residual(orc.HEX8, X3, conn3, 3, orc.LINEAR_ELASTIC, PROPS, u3.ravel()).reshape(2 * nn, 3)
    R3s = R3[:nn, :2] + R3[nn:, :2]
    assert np.abs(R3s.ravel() - R2).max() <= 1e-12 * np.abs(R2).max()
    # stiffness: element storage (NDOF x NDOF per element) summed over the slab's two node layers
    K2 = orc.assemble_stiffness_storage(orc.QUAD4, X2, conn2, 2, orc.LINEAR_ELASTIC, PROPS, u2.ravel())
    K3 = orc.assemble_stiffness_storage(orc.HEX8, X3, conn3, 3, orc.LINEAR_ELASTIC, PROPS, u3.ravel())
    ne = conn2.shape[0]
    K2 = np.asarray(K2).reshape(ne, 4, 2, 4, 2)
    K3 = np.asarray(K3).reshape(ne, 8, 3, 8, 3)
    K3s = sum(K3[:, za * 4:(za + 1) * 4, :2, zb * 4:(zb + 1) * 4, :2] for za in (0, 1) for zb in (0, 1))
    # element storage is in the reference's orientation (the transpose of the pattern block), the same for both
    assert np.abs(K3s - K2).max() <= 1e-12 * np.abs(K2).max()
    # patch test: a linear field u = B x gives zero residual at the interior nodes
    B = np.array([[1e-3, 2e-4], [-5e-4, 3e-3]])
    R = orc.assemble_residual(orc.QUAD4, X2, conn2, 2, orc.LINEAR_ELASTIC, PROPS, (X2 @ B.T).ravel()).reshape(nn, 2)
    inner = (X2[:, 0] > 1e-12) & (X2[:, 0] < 1 - 1e-12) & (X2[:, 1] > 1e-12) & (X2[:, 1] < 1 - 1e-12)
    assert np.abs(R[inner]).max() <= 1e-9 * np.abs(R).max()


def _asm(fec, mesh, gpu, bcs=()):
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    u = fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(u, sparse_matrix_type="csr", device=gpu.index or 0)
    p = fec.create_parameters(mesh, asm, fec.Mechanics(fec.PlaneStrain()), dirichlet_bcs=list(bcs))
    return asm, p


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["structured 12^2 jittered", "mechanics_coarse.g"])
def test_plane_strain_operators_match_oracle(fec, orc, gpu, which):
    print(f"[plane strain] {which}", flush=True)
    mesh = _jittered_quads(fec, 12) if which.startswith("structured") else fec.exodus.UnstructuredMesh(QUADS)
    nn = mesh.num_nodes
    left = np.nonzero(np.isclose(mesh.coords[:, 0], 0.0))[0] + 1
    mesh.nodeset_nodes["_left"] = left
    bcs = [fec.DirichletBC("displ_x", lambda X, t: 0.0, nset_name="_left"),
           fec.DirichletBC("displ_y", lambda X, t: 0.0, nset_name="_left")]
    asm, p = _asm(fec, mesh, gpu, bcs)
    unknown = asm.dof.unknown_dofs
    _, d2u = orc.update_dofs(2 * nn, asm.dof.dirichlet_dofs)
    g = np.random.default_rng(3)
    U = 1e-3 * g.normal(0, 1, 2 * nn)
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Vv = np.zeros(2 * nn)
    Vv[unknown - 1] = g.normal(0, 1, len(unknown))
    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)
    X, conn = mesh.coords, mesh.conn
    R_ref = orc.assemble_residual(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U)[unknown - 1]
    rp, cv, vals = orc.stiffness_csr(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U, d2u, len(unknown))
    KVa_ref = orc.assemble_action(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U, Vv)[unknown - 1]
    KVm_ref = orc.assemble_action(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U, Vv, matrix_free=True)[unknown - 1]
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = fec.stiffness(asm)
    assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
    errs = {"R": rel(R, R_ref), "K": rel(K.values.cpu().numpy(), vals)}
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    errs["KVa"] = rel(fec.hvp(asm, Vu).cpu().numpy(), KVa_ref)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    errs["KVm"] = rel(fec.hvp(asm, Vu).cpu().numpy(), KVm_ref)
    fec.assemble_mass_(asm, fec.mass, Uu, p)
    M = fec.mass(asm)
    _, _, mv = orc.mass_csr(orc.QUAD4, X, conn, 2, PROPS[0], d2u, len(unknown))
    errs["M"] = rel(M.values.cpu().numpy(), mv)
    print(f"[plane strain] {which}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < 1e-6 for v in errs.values()), errs


@pytest.mark.gpu
def test_reference_mechanics_regression(fec, orc, gpu):
    """test/mechanics/TestMechanics.jl on mechanics_coarse.g: displ_x and displ_y fixed on nset_3, displ_x fixed and
    displ_y = 1e-3 t on nset_1, one load step to t = 1, Newton + CG (the reference's Krylov cg). The linear problem's
    solution against the oracle's assembled system solved directly."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    print("[plane strain] reference mechanics regression", flush=True)
    mesh = fec.exodus.UnstructuredMesh(QUADS)
    bcs = [fec.DirichletBC("displ_x", lambda X, t: 0.0, nset_name="nset_3"),
           fec.DirichletBC("displ_y", lambda X, t: 0.0, nset_name="nset_3"),
           fec.DirichletBC("displ_x", lambda X, t: 0.0, nset_name="nset_1"),
           fec.DirichletBC("displ_y", lambda X, t: 1e-3 * t, nset_name="nset_1")]
    asm, p = _asm(fec, mesh, gpu, bcs)
    p.t = 1.0
    p.update_bc_values()
    Uu = fec.create_unknowns(asm)
    ls = fec.IterativeLinearSolver(asm, "cg", fec.IterativeLinearSolverSettings(rtol=1e-12, atol=0.0))
    hist = fec.NewtonSolver(ls, fec.NewtonSolverSettings(max_iters=5, rel_residual_tol=1e-6)).solve(Uu, p)
    print(f"[plane strain] Newton iterations {len(hist)}, rel {hist[-1].rel_R:.1e}", flush=True)
    # oracle: K u_free = -R(u with the boundary values, zero free values)
    nn = mesh.num_nodes
    unknown = asm.dof.unknown_dofs
    dd = asm.dof.dirichlet_dofs
    _, d2u = orc.update_dofs(2 * nn, dd)
    U0 = np.zeros(2 * nn)
    nset1 = mesh.nodeset_nodes["nset_1"]
    U0[2 * (nset1 - 1) + 1] = 1e-3  # nset_1 (top edge) and nset_3 (bottom edge) share no node
    X, conn = mesh.coords, mesh.conn
    R0 = orc.assemble_residual(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U0)[unknown - 1]
    rp, cv, vals = orc.stiffness_csr(orc.QUAD4, X, conn, 2, orc.LINEAR_ELASTIC, PROPS, U0, d2u, len(unknown))
    A = sp.csr_matrix((vals, cv - 1, rp - 1), shape=(len(unknown), len(unknown)))
    u_ref = -spla.spsolve(A.tocsc(), R0)
    err = np.linalg.norm(Uu.cpu().numpy() - u_ref) / np.linalg.norm(u_ref)
    print(f"[plane strain] solution rel error vs oracle direct solve {err:.2e}", flush=True)
    assert hist[-1].converged and err < 1e-6
