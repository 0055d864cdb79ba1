"""The allocation-free entry point fecx_assemble_vector_matrix_noalloc (no counterpart in the reference: for callers
that cannot allocate device memory) against fecx_assemble_vector_matrix through the C ABI: the same two-phase kernels
with the element-private outputs in the library's per-device scratch array, so the residual and the CSR values must
agree bit for bit (the structured-hex8 fused kernel is switched off on these contexts so that both calls take the
two-phase path). Calls in a row on the same GPU, each on a smaller mesh than the one before (the scratch keeps the
previous call's contents), then a mesh too large for the scratch, which is refused before any launch; and the fused
design, which needs no scratch, on a mesh beyond it."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ctx_for(fec, wl, gpu):
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    p = fec.create_parameters(wl.mesh, asm, wl.physics)
    asm.ctx.call("fecx_set_scatter", ctypes.c_int(1))  # FECX_SCATTER_TWO_PHASE
    U = torch.from_numpy(wl.U).to(gpu)
    fec.assemble_vector_(asm, fec.residual, U, p)  # the full field p.field
    return asm, p


def _both(fec, asm, p):
    P = fec.assembler._ptr
    R1 = torch.full_like(asm.residual_storage, float("nan"))
    K1 = torch.full_like(asm.stiffness_storage, float("nan"))
    asm.ctx.call("fecx_assemble_vector_matrix", P(p.field), P(R1), P(K1))
    R2 = torch.full_like(R1, float("nan"))
    K2 = torch.full_like(K1, float("nan"))
    asm.ctx.call("fecx_assemble_vector_matrix_noalloc", P(p.field), P(R2), P(K2))
    return R1.cpu().numpy(), K1.cpu().numpy(), R2.cpu().numpy(), K2.cpu().numpy()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_allocating_entry_point(fec, gpu, order):
    W = fec.workloads
    cases = [("neohookean_hex8 16^3", W.neohookean_hex8(16)), ("neohookean_tet10 4^3x6", W.neohookean_tet10(4)),
             ("poisson_hex8 6^3", W.poisson_hex8(6))]
    sizes = []
    for name, wl in cases:  # each call smaller than the one before, all on one GPU (one scratch array)
        print(f"[noalloc] {name}", flush=True)
        asm, p = _ctx_for(fec, wl, gpu)
        R1, K1, R2, K2 = _both(fec, asm, p)
        assert np.isfinite(R1).all() and np.isfinite(K1).all()
        assert np.array_equal(R1, R2) and np.array_equal(K1, K2), name
        sizes.append(wl.mesh.num_elements * (wl.mesh.conn.shape[1] * wl.nf) ** 2)
    assert sizes[0] > sizes[1] > sizes[2]


def test_noalloc_fused_beyond_the_scratch(fec, gpu):
    """A structured hex8 mesh whose element storage would not fit the scratch (64,000 elements x 1,392 bytes), on
    the fused design: no element-private outputs, so it runs, and agrees bit for bit with the allocating entry point on
    the same design; then a smaller mesh in a row."""
    W = fec.workloads
    for name, wl in (("neohookean_hex8 40^3", W.neohookean_hex8(40)), ("neohookean_hex8 12^3", W.neohookean_hex8(12))):
        print(f"[noalloc] fused {name}", flush=True)
        asm, p = _ctx_for(fec, wl, gpu)
        asm.ctx.call("fecx_set_scatter", ctypes.c_int(0))  # FECX_SCATTER_FUSED
        R1, K1, R2, K2 = _both(fec, asm, p)
        assert np.isfinite(R1).all() and np.isfinite(K1).all()
        assert np.array_equal(R1, R2) and np.array_equal(K1, K2), name


def test_noalloc_with_per_component_dirichlet(fec, gpu):
    """A Dirichlet set fixing single components (the dof-level CSR formed by a gather after the kernels): the
    allocation-free call agrees with the allocating one on both designs."""
    wl = fec.workloads.neohookean_hex8(8)
    mesh = wl.mesh
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    X = mesh.coords
    bcs = []
    for d, name in enumerate(var.names):  # roller planes: component d fixed on the plane x_d = 0
        mesh.nodeset_nodes[f"_r{d}"] = np.nonzero(np.isclose(X[:, d], 0.0))[0] + 1
        bcs.append(fec.DirichletBC(name, lambda X, t: 0.0, nset_name=f"_r{d}"))
    p = fec.create_parameters(mesh, asm, wl.physics, dirichlet_bcs=bcs)
    Uu = torch.from_numpy(wl.U[asm.dof.unknown_dofs - 1]).to(gpu)
    fec.assemble_vector_(asm, fec.residual, Uu, p)  # the full field p.field
    for design in (0, 1):  # FECX_SCATTER_FUSED, FECX_SCATTER_TWO_PHASE
        print(f"[noalloc] per-component Dirichlet, design {design}", flush=True)
        asm.ctx.call("fecx_set_scatter", ctypes.c_int(design))
        R1, K1, R2, K2 = _both(fec, asm, p)
        assert np.isfinite(R1).all() and np.isfinite(K1).all()
        assert np.array_equal(R1, R2) and np.array_equal(K1, K2), design


def test_noalloc_refuses_a_mesh_beyond_the_scratch(fec, gpu):
    wl = fec.workloads.neohookean_hex8(40)  # 64,000 elements x 1,392 bytes > 64 MiB
    asm, p = _ctx_for(fec, wl, gpu)
    P = fec.assembler._ptr
    with pytest.raises(fec.FECXError) as ei:
        asm.ctx.call("fecx_assemble_vector_matrix_noalloc", P(p.field), P(asm.residual_storage),
                     P(asm.stiffness_storage))
    assert ei.value.code == -1 and "scratch" in str(ei.value)
