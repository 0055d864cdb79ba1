"""Dirichlet sets that constrain some but not all components of a node (nf > 1): a BC on one component, as
DirichletBC("displ_x", ...) alone builds (src/bcs/DirichletBCs.jl:49-98), e.g. the roller (symmetry) conditions of a
mechanics model. The reference's reduced pattern then has rows and columns of single dofs (_update_dofs!,
SparsityPatterns.jl:53-117 on the unknown dofs), not whole nodes.

CPU: the host pattern and the COO->CSR map equal the oracle's restatement bit for bit.
GPU: residual, CSR values, assembled and matrix-free actions and the mass matrix against the oracle within 1e-6 (2-norm
and max-norm), on every scatter design that applies; SpMV and CG on the dof-level CSR against scipy."""
import ctypes

import numpy as np
import pytest
import torch

from test_gpu_parity import rel, _kind, _oracle_all


def _roller_dofs(mesh, nf=3):
    """x fixed on x = 0, y on y = 0, z on z = 0 (symmetry planes), plus all three components on x = 1."""
    X = mesh.coords
    dofs = []
    for d in range(3):
        nodes = np.nonzero(np.isclose(X[:, d], X[:, d].min()))[0]
        dofs.append(nf * nodes + d + 1)
    far = np.nonzero(np.isclose(X[:, 0], X[:, 0].max()))[0]
    for d in range(3):
        dofs.append(nf * far + d + 1)
    return np.unique(np.concatenate(dofs)).astype(np.int64)


def _random_dofs(nn, nf, frac, seed):
    g = np.random.default_rng(seed)
    return np.sort(g.choice(nf * nn, int(frac * nf * nn), replace=False) + 1).astype(np.int64)


def _cases(fec):
    W = fec.workloads
    hexm = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (6, 6, 6))
    tet = W.neohookean_tet10(2).mesh
    quad = fec.StructuredMesh("quad", (0, 0), (1, 1), (7, 7))
    return [("hex8 rollers", hexm, 3, _roller_dofs(hexm)),
            ("hex8 random", hexm, 3, _random_dofs(hexm.num_nodes, 3, 0.2, 1)),
            ("tet10 random", tet, 3, _random_dofs(tet.num_nodes, 3, 0.15, 2)),
            ("quad4 nf=2 random", quad, 2, _random_dofs(quad.num_nodes, 2, 0.25, 3))]


@pytest.mark.parametrize("idx", range(4))
def test_partial_pattern_matches_oracle(fec, orc, idx):
    name, mesh, nf, dir_dofs = _cases(fec)[idx]
    hp = fec.HostPattern(mesh, nf, dir_dofs)
    N = nf * mesh.num_nodes
    unknown, d2u = orc.update_dofs(N, dir_dofs)
    assert np.array_equal(hp.unknown_dofs, unknown)
    Is, Js, slots, perm = orc.pattern(nf, mesh.conn, d2u)
    rowptr, colval, _, cmap = orc.sparse_csr(len(unknown), Is, Js, want_map=True)
    rp, cv = hp.pattern(base=1)
    assert np.array_equal(rp, rowptr) and np.array_equal(cv, colval), name
    full = -np.ones(mesh.conn.shape[0] * (mesh.conn.shape[1] * nf) ** 2, dtype=np.int64)
    full[slots - 1] = cmap
    assert np.array_equal(hp.coo_map(), full), name
    # some node is split: the dof-level pattern is not a node-block pattern
    split = np.zeros(mesh.num_nodes, dtype=int)
    np.add.at(split, (dir_dofs - 1) // nf, 1)
    assert ((split > 0) & (split < nf)).any()


def _setup(fec, mesh, nf, physics, dir_dofs, gpu):
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    bcs = []
    for comp, name in enumerate(var.names):
        nodes = (dir_dofs[(dir_dofs - 1) % nf == comp] - 1) // nf + 1
        if len(nodes):
            mesh.nodeset_nodes[f"_bc{comp}"] = nodes
            bcs.append(fec.DirichletBC(name, lambda X, t: 0.0, nset_name=f"_bc{comp}"))
    p = fec.create_parameters(mesh, asm, physics, dirichlet_bcs=bcs)
    assert np.array_equal(asm.dof.dirichlet_dofs, dir_dofs)
    return asm, p


GPU_CASES = ["neohookean_hex8 rollers", "neohookean_hex8 random", "j2_hex8 rollers", "neohookean_tet10 random"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", GPU_CASES)
def test_partial_dirichlet_operators_match_oracle(fec, orc, gpu, name):
    print(f"[partial] {name}", flush=True)
    W = fec.workloads
    wname, bc = name.split()
    wl = {"neohookean_hex8": lambda: W.neohookean_hex8(6), "j2_hex8": lambda: W.j2_hex8(5),
          "neohookean_tet10": lambda: W.neohookean_tet10(2)}[wname]()
    mesh = wl.mesh
    dir_dofs = _roller_dofs(mesh) if bc == "rollers" else _random_dofs(mesh.num_nodes, 3, 0.2, 5)
    asm, p = _setup(fec, mesh, 3, wl.physics, dir_dofs, gpu)
    unknown = asm.dof.unknown_dofs
    N = 3 * mesh.num_nodes
    _, d2u = orc.update_dofs(N, dir_dofs)
    U = wl.U.copy()
    U[dir_dofs - 1] = 0.0
    Vv = np.zeros(N)
    Vv[unknown - 1] = fec.rng.normal(4, 9, len(unknown))
    st_old = None
    if wl.physics.kind == orc.J2:
        g = fec.rng.normal(1, 11, mesh.num_elements * 8 * 7).reshape(mesh.num_elements, 8, 7)
        st = np.concatenate([2e-3 * g[..., :6], 2e-3 * np.abs(g[..., 6:])], axis=2).ravel().astype(np.float32)
        asm.ctx.call("fecx_set_state_old", st.ctypes.data_as(ctypes.c_void_p))
        st_old = st.astype(np.float64)
    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)
    R_ref, (rp, cv, vals), KVa_ref, KVm_ref = _oracle_all(orc, wl, U, Vv, unknown, d2u, st_old)

    structured = mesh.element_type == "HEX8"
    designs = [0, 1] if structured else [1]  # FECX_SCATTER_FUSED, FECX_SCATTER_TWO_PHASE
    out = {}
    for design in designs:
        asm.ctx.call("fecx_set_scatter", ctypes.c_int(design))
        path = ctypes.c_int(-1)
        asm.ctx.call("fecx_get_scatter_path", ctypes.byref(path))
        assert path.value == design
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
        R = fec.residual(asm).cpu().numpy().copy()
        K = fec.stiffness(asm)
        Kv = K.values.cpu().numpy().copy()
        assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
        errs = {"R": rel(R, R_ref), "K": rel(Kv, vals)}
        print(f"[partial] {name} design {design}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
        assert errs["R"] < 1e-6 and errs["K"] < 1e-6
        out[design] = (R, Kv)
    if len(designs) == 2:  # the fused and the two-phase design agree
        assert np.abs(out[0][1] - out[1][1]).max() <= 1e-12 * np.abs(out[1][1]).max()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    ea = rel(fec.hvp(asm, Vu).cpu().numpy(), KVa_ref)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    em = rel(fec.hvp(asm, Vu).cpu().numpy(), KVm_ref)
    print(f"[partial] {name}: assembled action {ea:.2e}, matrix-free action {em:.2e}", flush=True)
    assert ea < 1e-6 and em < 1e-6


@pytest.mark.gpu
def test_partial_dirichlet_mass_spmv_cg(fec, orc, gpu):
    """Mass matrix on the dof-level pattern against the oracle; SpMV and Jacobi-CG on the dof-level CSR against
    scipy (the solver's dof-level path)."""
    import scipy.sparse.linalg as spla
    print("[partial] mass, spmv, cg", flush=True)
    wl = fec.workloads.neohookean_hex8(6)
    mesh = wl.mesh
    dir_dofs = _roller_dofs(mesh)
    asm, p = _setup(fec, mesh, 3, wl.physics, dir_dofs, gpu)
    unknown = asm.dof.unknown_dofs
    _, d2u = orc.update_dofs(3 * mesh.num_nodes, dir_dofs)
    Uu = torch.from_numpy(wl.U[unknown - 1] * 0.2).to(gpu)
    rho = fec.density(wl.physics)
    fec.assemble_mass_(asm, fec.mass, Uu, p)
    M = fec.mass(asm)
    rp, cv, mv = orc.mass_csr(orc.HEX8, mesh.coords, mesh.conn, 3, rho, d2u, len(unknown))
    assert np.array_equal(M.rowptr.cpu().numpy() + 1, rp) and np.array_equal(M.colval.cpu().numpy() + 1, cv)
    assert rel(M.values.cpu().numpy(), mv) < 1e-6

    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    K = fec.stiffness(asm)
    A = K.to_scipy()
    assert A.shape == (len(unknown), len(unknown))
    x = np.random.default_rng(0).normal(0, 1, A.shape[0])
    y = torch.zeros_like(Uu)
    asm.ctx.call("fecx_csr_spmv", fec.assembler._ptr(K.values), fec.assembler._ptr(torch.from_numpy(x).to(gpu)),
                 fec.assembler._ptr(y))
    yr = A @ x
    assert np.max(np.abs(y.cpu().numpy() - yr)) <= 1e-13 * np.max(np.abs(yr))
    b = fec.residual(asm).clone()
    ls = fec.IterativeLinearSolver(asm, "cg", fec.IterativeLinearSolverSettings(rtol=1e-10, atol=0.0))
    xg = torch.zeros_like(b)
    it, res = ls.cg(K.values, b, xg)
    xs = spla.spsolve(A.tocsc(), b.cpu().numpy())
    err = np.linalg.norm(xg.cpu().numpy() - xs) / np.linalg.norm(xs)
    print(f"[partial] cg iterations {it}, rel error vs direct {err:.2e}", flush=True)
    assert 0 < it < 2000 and err <= 1e-7
