"""The assemble_matrix_free_action_full! contract (src/assemblers/MatrixAction.jl:99-149), replaying
test/TestAssemblers.jl:341-430 on the test Mechanics physics (hex8, whole-node Dirichlet on two faces):
  (a) the full-DOF action equals the assembled full Jacobian times v_full at the free rows, v_full nonzero at the
      Dirichlet slots (the K_{f,BC} v_BC cross term the free-DOF entry point drops), rtol 1e-6;
  (b) with the Dirichlet slots of v_full zeroed it reproduces the free-DOF entry point;
  (c) the cross term is nonzero;
and the scratch field's Dirichlet slots are reset to zero after the call (MatrixAction.jl:136-148), so a following
free-DOF action does not see the v_BC values."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mesh_and_bcs(fec, n=4):
    wl = fec.workloads.neohookean_hex8(n)
    mesh = wl.mesh
    X = mesh.coords
    nodes = np.flatnonzero(np.isclose(X[:, 0], 0.0) | np.isclose(X[:, 2], 1.0)) + 1
    mesh.nodeset_nodes["_fixed"] = nodes
    return mesh, nodes


def _assembler(fec, mesh, gpu, bcs):
    var = fec.VectorFunction(fec.FunctionSpace(mesh, "H1Field", "Lagrange"), "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    dbcs = [fec.DirichletBC(name, lambda X, t: 0.0, nset_name="_fixed") for name in var.names] if bcs else []
    p = fec.create_parameters(mesh, asm, fec.Mechanics(), dirichlet_bcs=dbcs)
    return asm, p


def test_matrix_free_action_full_contract(fec, gpu):
    print("[mf full] hex8 Mechanics 4^3", flush=True)
    mesh, nodes = _mesh_and_bcs(fec)
    asm_ref, p_ref = _assembler(fec, mesh, gpu, bcs=False)
    asm, p = _assembler(fec, mesh, gpu, bcs=True)
    N = len(asm.dof)
    free = asm.dof.unknown_dofs - 1
    bc = asm.dof.dirichlet_dofs - 1
    assert len(bc) == 3 * len(nodes) and len(free) + len(bc) == N
    # reference: the no-BC assembler's full K times v_full
    U_ref = fec.create_unknowns(asm_ref)
    fec.assemble_stiffness_(asm_ref, fec.stiffness, U_ref, p_ref)
    K_full = fec.stiffness(asm_ref).to_scipy()
    v_full_h = np.arange(1, N + 1, dtype=np.float64) / N
    Kv_free_ref = (K_full @ v_full_h)[free]
    Uu = fec.create_unknowns(asm)
    fec.assemble_vector_(asm, fec.residual, Uu, p)  # populates p.field (BC values, unknowns)
    U_full = p.field.clone()
    v_full = torch.from_numpy(v_full_h).to(gpu)
    # (a)
    fec.assemble_matrix_free_action_full_(asm, fec.stiffness_action, U_full, v_full, p)
    Kv_free_mf = fec.hvp(asm, Uu).cpu().numpy().copy()
    err = np.abs(Kv_free_mf - Kv_free_ref).max() / np.abs(Kv_free_ref).max()
    print(f"[mf full] (a) full action vs K_full v_full: {err:.2e}", flush=True)
    assert np.allclose(Kv_free_mf, Kv_free_ref, rtol=1e-6, atol=1e-6 * np.abs(Kv_free_ref).max())
    # the scratch field's Dirichlet slots were reset
    assert not p.hvp_scratch_field[torch.from_numpy(bc).to(gpu)].any()
    # (b)
    v0_h = v_full_h.copy()
    v0_h[bc] = 0.0
    Vu = torch.from_numpy(v_full_h[free]).to(gpu)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    Kv_legacy = fec.hvp(asm, Uu).cpu().numpy().copy()
    fec.assemble_matrix_free_action_full_(asm, fec.stiffness_action, U_full, torch.from_numpy(v0_h).to(gpu), p)
    Kv_zero_bc = fec.hvp(asm, Uu).cpu().numpy().copy()
    assert np.array_equal(Kv_zero_bc, Kv_legacy)
    # (c)
    assert not np.allclose(Kv_free_mf, Kv_zero_bc, rtol=0, atol=0)
    # a free-DOF action right after a full call with nonzero v_BC: unaffected by those values
    fec.assemble_matrix_free_action_full_(asm, fec.stiffness_action, U_full, v_full, p)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    assert np.array_equal(fec.hvp(asm, Uu).cpu().numpy(), Kv_legacy)


def test_matrix_free_action_full_validates_inputs(fec, gpu):
    mesh, _ = _mesh_and_bcs(fec, 2)
    asm, p = _assembler(fec, mesh, gpu, bcs=True)
    N = len(asm.dof)
    good = torch.zeros(N, dtype=torch.float64, device=gpu)
    for bad in (torch.zeros(N - 1, dtype=torch.float64, device=gpu), torch.zeros(N, dtype=torch.float32, device=gpu),
                torch.zeros(N, dtype=torch.float64), torch.zeros(2 * N, dtype=torch.float64, device=gpu)[::2]):
        with pytest.raises(fec.FECXError):
            fec.assemble_matrix_free_action_full_(asm, fec.stiffness_action, bad, good, p)
        with pytest.raises(fec.FECXError):
            fec.assemble_matrix_free_action_full_(asm, fec.stiffness_action, good, bad, p)
