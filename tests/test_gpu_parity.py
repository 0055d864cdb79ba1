"""GPU parity: residual, CSR Jacobian, assembled and matrix-free actions from libfecx.so (FP32 element kernels,
FP64 reductions) against the FP64 oracle on the same seeded meshes and fields.

Tolerance: 1e-6 relative error against FP64 for every output, in the 2-norm over all entries and in the max-norm
(largest entry error over the largest entry) -- the bound the FP32 element kernels are held to. FP32 element
arithmetic (unit roundoff 6e-8) summed over <= 24 elements x 8 quadrature points lands near 1e-7 at these sizes. The
sparsity pattern is compared bit-exactly."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {"R": 1e-6, "K": 1e-6, "KV": 1e-6}


def rel(a, b):
    """max of the relative 2-norm and max-norm errors"""
    e2 = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    em = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) if len(b) else 0.0
    return max(e2, em)


def _kind(orc, mesh):
    return {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8, "TET10": orc.TET10}[mesh.element_type]


def _setup(fec, wl, gpu):
    mesh = wl.mesh
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    bcs = []
    if len(wl.dirichlet_dofs):
        # Dirichlet on the given dofs with value 0 (all components of the listed nodes)
        nodes = np.unique((wl.dirichlet_dofs - 1) // wl.nf + 1)
        mesh.nodeset_nodes["_bc"] = nodes
        bcs = [fec.DirichletBC(name, lambda X, t: 0.0, nset_name="_bc") for name in var.names]
    p = fec.create_parameters(mesh, asm, wl.physics, dirichlet_bcs=bcs)
    return asm, p


def _oracle_all(orc, wl, U, Vv, unknown, d2u, st_old=None):
    mesh = wl.mesh
    k = _kind(orc, mesh)
    props = wl.physics.props()
    R = orc.assemble_residual(k, mesh.coords, mesh.conn, wl.nf, wl.physics.kind, props, U, st_old,
                              None if st_old is None else np.zeros_like(st_old))
    rp, cv, vals = orc.stiffness_csr(k, mesh.coords, mesh.conn, wl.nf, wl.physics.kind, props, U, d2u, len(unknown),
                                     st_old)
    KVa = orc.assemble_action(k, mesh.coords, mesh.conn, wl.nf, wl.physics.kind, props, U, Vv, st_old)
    KVm = orc.assemble_action(k, mesh.coords, mesh.conn, wl.nf, wl.physics.kind, props, U, Vv, st_old,
                              matrix_free=True)
    return R[unknown - 1], (rp, cv, vals), KVa[unknown - 1], KVm[unknown - 1]


CASES = ["poisson_quad4", "poisson_hex8", "poisson_tet10", "mechanics_hex8", "neohookean_hex8", "neohookean_tet10",
         "j2_hex8"]


def _workload(fec, name):
    W = fec.workloads
    if name == "poisson_quad4":
        return W.poisson_quad4(12)
    if name == "poisson_hex8":
        return W.poisson_hex8(6)
    if name == "poisson_tet10":
        wl = W.neohookean_tet10(2)
        return W.Workload("poisson_tet10", wl.mesh, fec.Poisson("sines", 1.0), 1,
                          2 * fec.rng.uniform(0, 1, wl.mesh.num_nodes) - 1, fec.boundary_nodes(wl.mesh))
    if name == "mechanics_hex8":
        wl = W.neohookean_hex8(5)
        return W.Workload("mechanics_hex8", wl.mesh, fec.Mechanics(), 3, 0.01 * wl.U)
    if name == "neohookean_hex8":
        return W.neohookean_hex8(6)
    if name == "neohookean_tet10":
        return W.neohookean_tet10(3)
    if name == "mechanics_tet10":
        wl = W.neohookean_tet10(3)
        return W.Workload("mechanics_tet10", wl.mesh, fec.Mechanics(), 3, 0.01 * wl.U)
    if name == "j2_hex8":
        return W.j2_hex8(4, dirichlet=True)
    raise KeyError(name)


@pytest.mark.parametrize("name", CASES)
def test_operators_match_oracle(fec, orc, gpu, name):
    _operators_vs_oracle(fec, orc, gpu, name, fp64_geometry=False)


@pytest.mark.parametrize("name", ["poisson_tet10", "mechanics_tet10", "neohookean_tet10"])
def test_operators_match_oracle_fp64_geometry(fec, orc, gpu, name):
    """tet10 with the FP64 element geometry, field gradient and stress (fecx_set_fp64_geometry): the same outputs
    against the oracle, the same bounds, and the same bit-for-bit agreement of fused and separate calls."""
    _operators_vs_oracle(fec, orc, gpu, name, fp64_geometry=True)


def _operators_vs_oracle(fec, orc, gpu, name, fp64_geometry):
    print(f"[gpu parity] {name}" + (" fp64 geometry" if fp64_geometry else ""), flush=True)
    wl = _workload(fec, name)
    asm, p = _setup(fec, wl, gpu)
    if fp64_geometry:
        asm.ctx.call("fecx_set_fp64_geometry", ctypes.c_int(1))
    unknown = asm.dof.unknown_dofs
    N = wl.nf * wl.mesh.num_nodes
    _, d2u = orc.update_dofs(N, asm.dof.dirichlet_dofs)
    U = wl.U.copy()
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Vv = np.zeros(N)
    Vv[unknown - 1] = fec.rng.normal(3, 9, len(unknown))
    st_old = None
    if wl.physics.kind == orc.J2:
        nq = 8
        st_old = np.zeros((wl.mesh.num_elements, nq, 7))
        g = fec.rng.normal(1, 11, st_old.size).reshape(st_old.shape)
        st_old[..., :6] = 2e-3 * g[..., :6]
        st_old[..., 6] = 2e-3 * np.abs(g[..., 6])
        st_old = st_old.ravel()
        asm.ctx.call("fecx_set_state_old", st_old.astype(np.float32).ctypes.data_as(__import__("ctypes").c_void_p))
        st_old = st_old.astype(np.float32).astype(np.float64)
    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)
    R_ref, (rp, cv, vals), KVa_ref, KVm_ref = _oracle_all(orc, wl, U, Vv, unknown, d2u, st_old)

    fec.assemble_vector_(asm, fec.residual, Uu, p)
    R = fec.residual(asm).cpu().numpy()
    fec.assemble_stiffness_(asm, fec.stiffness, Uu, p)
    K = fec.stiffness(asm)
    Kvals = K.values.cpu().numpy().copy()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    KVa = fec.hvp(asm, Vu).cpu().numpy().copy()
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    KVm = fec.hvp(asm, Vu).cpu().numpy().copy()

    assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
    errs = {"R": rel(R, R_ref), "K": rel(Kvals, vals), "KVa": rel(KVa, KVa_ref), "KVm": rel(KVm, KVm_ref)}
    print(f"[gpu parity] {name}{' fp64 geometry' if fp64_geometry else ''}: " +
          " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["R"] < TOL["R"] and errs["K"] < TOL["K"]
    assert errs["KVa"] < TOL["KV"] and errs["KVm"] < TOL["KV"]

    # one fused pass gives the same numbers as the two separate calls, bit for bit (no atomics anywhere)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    assert np.array_equal(fec.residual(asm).cpu().numpy(), R)
    assert np.array_equal(asm.stiffness_storage.cpu().numpy(), Kvals)
    if fp64_geometry:  # the allocation-free entry point runs the same element kernel
        P = fec.assembler._ptr
        R2 = torch.full_like(asm.residual_storage, float("nan"))
        K2 = torch.full_like(asm.stiffness_storage, float("nan"))
        asm.ctx.call("fecx_assemble_vector_matrix_noalloc", P(p.field), P(R2), P(K2))
        R1 = torch.full_like(R2, float("nan"))
        K1 = torch.full_like(K2, float("nan"))
        asm.ctx.call("fecx_assemble_vector_matrix", P(p.field), P(R1), P(K1))
        assert torch.equal(R1, R2) and torch.equal(K1, K2)

    if st_old is not None:
        ns = np.zeros(st_old.size, dtype=np.float32)
        fec.assemble_vector_(asm, fec.residual, Uu, p)
        asm.ctx.call("fecx_get_state_new", ns.ctypes.data_as(__import__("ctypes").c_void_p))
        st_ref = np.zeros_like(st_old)
        k = _kind(orc, wl.mesh)
        orc.assemble_residual(k, wl.mesh.coords, wl.mesh.conn, 3, orc.J2, wl.physics.props(), U, st_old, st_ref)
        assert np.abs(ns - st_ref).max() < 1e-6 * max(1.0, np.abs(st_ref).max()) + 1e-9
        assert (st_ref[6::7] > st_old[6::7]).any(), "the J2 case should yield somewhere"


def test_numeric_error_flag(fec, gpu):
    """A folded element (negative Jacobian) is reported as FECX_ERR_NUMERIC, not silently assembled."""
    wl = fec.workloads.neohookean_hex8(3)
    mesh = wl.mesh
    mesh.coords = mesh.coords.copy()
    mesh.coords[21] += np.array([0.6, 0.6, 0.6])  # interior node pushed through its neighbours
    asm, p = _setup(fec, wl, gpu)
    Uu = torch.from_numpy(np.zeros(3 * mesh.num_nodes)).to(gpu)
    with pytest.raises(fec.FECXError) as ei:
        fec.assemble_vector_(asm, fec.residual, Uu, p)
    assert ei.value.code == -5


def test_numeric_error_flag_recovers(fec, gpu):
    """The error flag is cleared on the device only after a call that found it set: on one context a bad call
    (u_x = -2x, so F_xx = -1 and J < 0 at every point) raises, the next good call succeeds, a bad one raises again."""
    wl = fec.workloads.neohookean_hex8(3)
    asm, p = _setup(fec, wl, gpu)
    N = 3 * wl.mesh.num_nodes
    bad = np.zeros(N)
    bad[0::3] = -2.0 * wl.mesh.coords[:, 0]
    Ub = torch.from_numpy(bad).to(gpu)
    Ug = torch.zeros(N, dtype=torch.float64, device=gpu)
    for U, ok in [(Ub, False), (Ug, True), (Ug, True), (Ub, False), (Ub, False), (Ug, True)]:
        if ok:
            fec.assemble_vector_and_stiffness_(asm, U, p)
            assert float(torch.linalg.norm(fec.residual(asm))) == 0.0  # undeformed: no residual
        else:
            with pytest.raises(fec.FECXError) as ei:
                fec.assemble_vector_and_stiffness_(asm, U, p)
            assert ei.value.code == -5


@pytest.mark.parametrize("name", ["neohookean_hex8", "poisson_hex8"])
def test_pipelined_chunks_bit_identical(fec, gpu, name):
    """Pipelined matrix assembly (element and reduction pieces on two streams) gives the same bits as one piece. The
    two-phase design is forced (structured hex8 meshes otherwise take the fused kernel, which has no pieces), and the
    element-kernel launch count shows that the pieces were taken. tet10 meshes are not cases: their heavy vertex rows
    form a second row bucket, and pipelining applies to one-bucket meshes only."""
    print(f"[gpu chunks] {name}", flush=True)
    wl = _workload(fec, name)
    asm, p = _setup(fec, wl, gpu)
    asm.ctx.call("fecx_set_scatter", ctypes.c_int(1))  # FECX_SCATTER_TWO_PHASE
    U = wl.U.copy()
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Uu = torch.from_numpy(U[asm.dof.unknown_dofs - 1]).to(gpu)
    out = []
    for chunks in (1, 5, 1000):
        fec.set_chunks(asm, chunks)
        fec.reset_profile(asm)
        fec.set_profiling(asm, True)
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
        fec.set_profiling(asm, False)
        launches = fec.profile(asm)["element"][0]
        assert (launches == 1) if chunks == 1 else (launches > 1), (chunks, launches)
        out.append((fec.residual(asm).cpu().numpy().copy(), asm.stiffness_storage.cpu().numpy().copy()))
    for R, K in out[1:]:
        assert np.array_equal(R, out[0][0]) and np.array_equal(K, out[0][1])


FANS = {"fan_tet10_nh": (32, 165), "fan11_tet10_nh": (11, 60), "fan13_tet10_nh": (13, 70)}  # ring points, blocks


@pytest.mark.parametrize("case", ["one_hex8_nh", "one_tet10_poisson", "all_dirichlet", "one_free_node",
                                  "fan_tet10_nh", "fan11_tet10_nh", "fan13_tet10_nh"])
def test_edge_meshes(fec, orc, gpu, case):
    """Smallest meshes and degenerate DOF sets: one element, every dof constrained (empty system), one unknown; and a
    vertex row wider than one reduction sweep."""
    print(f"[gpu edge] {case}", flush=True)
    W = fec.workloads
    if case == "one_hex8_nh":
        mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (2, 2, 2))
        wl = W.Workload(case, mesh, fec.NeoHookean(1.0, 0.5), 3, 0.05 * np.sin(np.arange(24.0)))
    elif case == "one_tet10_poisson":
        v = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.0], [0.1, 0.2, 0.9]])
        edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
        X = np.vstack([v] + [0.5 * (v[a] + v[b]) for a, b in edges])
        mesh = fec.TetMesh(X, np.arange(1, 11, dtype=np.int64)[None, :])
        wl = W.Workload(case, mesh, fec.Poisson("sines", 1.0), 1, np.cos(np.arange(10.0)))
    elif case in FANS:
        # 32 ring points: a 165-block centre row (slot-owner kernel, 128 lanes, two sweeps); 11: 60 blocks and 22
        # incidences (scatter kernel, one staging pass); 13: 70 blocks and 26 incidences (scatter kernel, two passes)
        mesh = fec.fan_tet10_mesh(FANS[case][0])
        wl = W.Workload(case, mesh, fec.NeoHookean(1.0, 0.5), 3, 0.02 * np.sin(np.arange(3.0 * mesh.num_nodes)))
    else:
        mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (3, 3, 3))
        dd = np.arange(1, 28, dtype=np.int64) if case == "all_dirichlet" else np.delete(np.arange(1, 28), 13)
        wl = W.Workload(case, mesh, fec.Poisson("constant", 1.0), 1, np.linspace(-1, 1, 27), dd)
    asm, p = _setup(fec, wl, gpu)
    unknown = asm.dof.unknown_dofs
    N = wl.nf * mesh.num_nodes
    _, d2u = orc.update_dofs(N, asm.dof.dirichlet_dofs)
    U = wl.U.copy()
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Vv = np.zeros(N)
    Vv[unknown - 1] = 1.0 + np.arange(len(unknown))
    R_ref, (rp, cv, vals), KVa_ref, KVm_ref = _oracle_all(orc, wl, U, Vv, unknown, d2u)
    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy()
    K = fec.stiffness(asm)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    KV = fec.hvp(asm, Vu).cpu().numpy()
    assert len(R) == len(unknown) and K.values.numel() == len(vals)
    assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
    if case.startswith("fan"):
        assert np.diff(rp).max() // 3 == FANS[case][1]
    if len(unknown):
        assert rel(R, R_ref) < TOL["R"] and rel(K.values.cpu().numpy(), vals) < TOL["K"]
        assert rel(KV, KVm_ref) < TOL["KV"]


@pytest.mark.parametrize("dirichlet", ["none", "random components"])
def test_merged_row_launch_bit_identical(fec, gpu, dirichlet):
    """tet10 with 3 fields: the edge and vertex rows reduced in one launch (fecx_set_merged_rows 1, the default) and
    in two give the same residual and CSR values bit for bit."""
    print(f"[gpu parity] merged rows, {dirichlet}", flush=True)
    wl = fec.workloads.neohookean_tet10(4)
    mesh = wl.mesh
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    U = wl.U.copy()
    bcs = []
    if dirichlet != "none":
        g = np.random.default_rng(7)
        dd = np.sort(g.choice(3 * mesh.num_nodes, int(0.2 * 3 * mesh.num_nodes), replace=False) + 1)
        for comp, name in enumerate(var.names):
            nodes = (dd[(dd - 1) % 3 == comp] - 1) // 3 + 1
            mesh.nodeset_nodes[f"_m{comp}"] = nodes
            bcs.append(fec.DirichletBC(name, lambda X, t: 0.0, nset_name=f"_m{comp}"))
        U = 0.2 * U
        U[dd - 1] = 0.0
    p = fec.create_parameters(mesh, asm, wl.physics, dirichlet_bcs=bcs)
    Uu = torch.from_numpy(U[asm.dof.unknown_dofs - 1]).to(gpu)
    out = []
    for on in (1, 0, 1):
        asm.ctx.call("fecx_set_merged_rows", ctypes.c_int(on))
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
        out.append((fec.residual(asm).cpu().numpy().copy(), fec.stiffness(asm).values.cpu().numpy().copy()))
    for R, K in out[1:]:
        assert np.isfinite(K).all() and np.array_equal(R, out[0][0]) and np.array_equal(K, out[0][1])
