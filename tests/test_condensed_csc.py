"""Condensed mode (use_condensed = true) and the CSC output (sparse_matrix_type = :csc, the reference's default).

Reference: src/assemblers/Utils.jl:53-167 (the constraint adjustments), SparseMatrixAssembler.jl:66,126-133,246-251,
Assemblers.jl:310-388 (the accessors that apply them), SparsityPatterns.jl:301-308 (_csc_matrix = sparse!), and the
consistency contract of test/TestAssemblers.jl:78-277 (every assembler result agrees across csc/csr and condensed or
not).

CPU part: the oracle's CSC and condensed adjustments are pinned by what the mathematics fixes -- the CSC of A is the CSR
of A^T with the same sums, scipy's COO -> CSC, the penalty 1e6 tr(A)/n of the unadjusted matrix, the free-dof block
of the condensed matrix equal to the reduced matrix, and a condensed solve equal to the reduced solve with zero at the
constrained dofs. GPU part: the library against that oracle, relative error bound 1e-6 (2-norm and max-norm) on
values, patterns bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

TOL = 1e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _kind(orc, mesh):
    return {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8, "TET10": orc.TET10}[mesh.element_type]


def _poisson_quad4(fec, n=10):
    wl = fec.workloads.poisson_quad4(n)
    wl.U = 2 * fec.rng.uniform(0, 7, wl.mesh.num_nodes) - 1
    return wl


def _nh_hex8_partial(fec, n=4):
    """neo-Hookean hex8 with Dirichlet dofs on single components: x fixed on the x = 0 face, y and z on y = 0."""
    wl = fec.workloads.neohookean_hex8(n)
    X = wl.mesh.coords
    fx = np.flatnonzero(np.isclose(X[:, 0], 0.0)) + 1
    fy = np.flatnonzero(np.isclose(X[:, 1], 0.0)) + 1
    d = np.concatenate([3 * (fx - 1) + 1, 3 * (fy - 1) + 2, 3 * (fy - 1) + 3])
    wl.dirichlet_dofs = np.unique(d).astype(np.int64)
    return wl


def _storage(orc, wl, U):
    m = wl.mesh
    return orc.assemble_stiffness_storage(_kind(orc, m), m.coords, m.conn, wl.nf, wl.physics.kind, wl.physics.props(),
                                          U)


def _full_csr(orc, wl, storage):
    N = wl.nf * wl.mesh.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    return orc.sparse_csr_rows(wl.nf, wl.mesh.conn, d2u, N, storage)


# ------------------------------------------------------------------ oracle pins (CPU)
@pytest.mark.parametrize("case", ["poisson_quad4", "nh_hex8_partial"])
def test_oracle_csc_is_transposed_csr_and_scipy(fec, orc, case):
    import scipy.sparse as sp
    wl = _poisson_quad4(fec) if case == "poisson_quad4" else _nh_hex8_partial(fec)
    N = wl.nf * wl.mesh.num_nodes
    unk, d2u = orc.update_dofs(N, wl.dirichlet_dofs)
    st = _storage(orc, wl, wl.U)
    Is, Js, slots, _ = orc.pattern(wl.nf, wl.mesh.conn, d2u)
    V = orc.gather_slots(slots, st)
    cp, rv, nzv = orc.sparse_csc(len(unk), Is, Js, V)
    rp, cv, vals, _ = orc.sparse_csr(len(unk), Js, Is, V)  # CSR of the transpose: same sums in the same order
    assert np.array_equal(cp, rp) and np.array_equal(rv, cv) and np.array_equal(nzv, vals)
    ref = sp.coo_matrix((V, (Is - 1, Js - 1)), shape=(len(unk),) * 2).tocsc()
    ref.sort_indices()
    assert np.array_equal(ref.indptr, cp - 1) and np.array_equal(ref.indices, rv - 1)
    assert rel(nzv, ref.data) < 1e-14


@pytest.mark.parametrize("case", ["poisson_quad4", "nh_hex8_partial"])
def test_oracle_condensed_adjustments(fec, orc, case):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    wl = _poisson_quad4(fec) if case == "poisson_quad4" else _nh_hex8_partial(fec)
    N = wl.nf * wl.mesh.num_nodes
    dd = wl.dirichlet_dofs
    c = orc.constraint_storage(N, dd)
    st = _storage(orc, wl, wl.U)
    rp, cv, vals = _full_csr(orc, wl, st)
    A0 = sp.csr_matrix((vals, cv - 1, rp - 1), shape=(N, N))
    penalty = 1e6 * sum(A0.diagonal().tolist()) / N  # serial sum in row order, as tr() of a CSR matrix
    adj = orc.adjust_matrix(rp, cv, vals.copy(), c)
    A = sp.csr_matrix((adj, cv - 1, rp - 1), shape=(N, N))
    con, free = dd - 1, np.setdiff1d(np.arange(N), dd - 1)
    # constrained rows: zero but the diagonal, which holds the penalty; free rows as assembled
    Ac = A[con].toarray()
    assert np.array_equal(Ac[np.arange(len(con)), con], np.full(len(con), penalty))
    Ac[np.arange(len(con)), con] = 0.0
    assert not Ac.any()
    assert (A[free] != A0[free]).nnz == 0
    # the free block is the reduced (non-condensed) matrix, bit for bit
    unk, d2u = orc.update_dofs(N, dd)
    rpr, cvr, valr = orc.sparse_csr_rows(wl.nf, wl.mesh.conn, d2u, len(unk), st)
    Ar = sp.csr_matrix((valr, cvr - 1, rpr - 1), shape=(len(unk),) * 2)
    assert (A[free][:, free] != Ar).nnz == 0
    # CSC form: constrained columns zeroed instead (the transpose for this symmetric matrix)
    adjc = orc.adjust_matrix(rp, cv, vals.copy(), c)  # same compressed arrays read as CSC columns
    Acsc = sp.csc_matrix((adjc, cv - 1, rp - 1), shape=(N, N))
    assert abs(Acsc - A.T).max() <= 1e-12 * abs(A0).max()
    # condensed solve = reduced solve with zeros at the constrained dofs (homogeneous BCs)
    R = orc.assemble_residual(_kind(orc, wl.mesh), wl.mesh.coords, wl.mesh.conn, wl.nf, wl.physics.kind,
                              wl.physics.props(), wl.U)
    b = orc.adjust_vector(R.copy(), c)
    assert np.array_equal(b[con], np.zeros(len(con))) and np.array_equal(b[free], R[free])
    x = spla.spsolve(A.tocsc(), b)
    xr = spla.spsolve(Ar.tocsc(), R[free])
    assert np.abs(x[con]).max() == 0.0
    assert rel(x[free], xr) < 1e-9
    v = np.cos(np.arange(N))
    Av = orc.adjust_action((A0 @ v).copy(), c, v)
    assert np.array_equal(Av[con], v[con]) and rel(Av[free], (A0 @ v)[free]) == 0.0


def test_host_condensed_pattern_matches_oracle(fec, orc):
    """The library's host pattern in condensed mode is the full pattern, whatever the Dirichlet set (per-component
    ones included): bit-exact against the oracle's unreduced sparse!."""
    wl = _nh_hex8_partial(fec, 3)
    N = 3 * wl.mesh.num_nodes
    hp = fec.HostPattern(wl.mesh, 3, wl.dirichlet_dofs, condensed=True)
    rp, cv = hp.pattern(1)
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    Is, Js, _, _ = orc.pattern(3, wl.mesh.conn, d2u)
    rpo, cvo, _, _ = orc.sparse_csr(N, Is, Js)
    assert np.array_equal(rp, rpo) and np.array_equal(cv, cvo)
    assert np.array_equal(hp.unknown_dofs, np.setdiff1d(np.arange(1, N + 1), wl.dirichlet_dofs))
    # the same per-component set without condensed mode: the reference's dof-level reduced pattern
    hr = fec.HostPattern(wl.mesh, 3, wl.dirichlet_dofs, condensed=False)
    unknown, d2u = orc.update_dofs(N, wl.dirichlet_dofs)
    Is, Js, _, _ = orc.pattern(3, wl.mesh.conn, d2u)
    rpo, cvo, _, _ = orc.sparse_csr(len(unknown), Is, Js)
    rp, cv = hr.pattern(1)
    assert np.array_equal(rp, rpo) and np.array_equal(cv, cvo)


# ------------------------------------------------------------------ GPU parity
def _gpu_setup(fec, wl, gpu, sp_type, condensed):
    mesh = wl.mesh
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type=sp_type, use_condensed=condensed, device=gpu.index or 0)
    bcs = []
    if len(wl.dirichlet_dofs):
        for comp, name in enumerate(var.names):
            nodes = (wl.dirichlet_dofs[(wl.dirichlet_dofs - 1) % wl.nf == comp] - 1) // wl.nf + 1
            if len(nodes):
                mesh.nodeset_nodes[f"_bc{comp}"] = nodes
                bcs.append(fec.DirichletBC(name, lambda X, t: 0.0, nset_name=f"_bc{comp}"))
    p = fec.create_parameters(mesh, asm, wl.physics, dirichlet_bcs=bcs)
    assert np.array_equal(asm.dof.dirichlet_dofs, wl.dirichlet_dofs)
    return asm, p


@pytest.mark.gpu
@pytest.mark.parametrize("sp_type", ["csr", "csc"])
@pytest.mark.parametrize("case", ["poisson_quad4", "nh_hex8_partial"])
def test_condensed_matches_oracle(fec, orc, gpu, case, sp_type):
    print(f"[condensed] {case} {sp_type}", flush=True)
    wl = _poisson_quad4(fec) if case == "poisson_quad4" else _nh_hex8_partial(fec)
    asm, p = _gpu_setup(fec, wl, gpu, sp_type, True)
    N = wl.nf * wl.mesh.num_nodes
    dd = wl.dirichlet_dofs
    c = orc.constraint_storage(N, dd)
    U = wl.U.copy()
    U[dd - 1] = 0.0
    Vf = np.cos(np.arange(N))
    Uu = torch.from_numpy(U).to(gpu)  # condensed: full-length unknown vectors
    Vu = torch.from_numpy(Vf).to(gpu)
    assert fec.create_unknowns(asm).numel() == N
    k = _kind(orc, wl.mesh)
    props = wl.physics.props()
    R_ref = orc.adjust_vector(orc.assemble_residual(k, wl.mesh.coords, wl.mesh.conn, wl.nf, wl.physics.kind, props, U),
                              c)
    rp, cv, vals = _full_csr(orc, wl, _storage(orc, wl, U))
    K_ref = orc.adjust_matrix(rp, cv, vals, c)
    # the scratch field V holds Vu at the unknown dofs only (update_field_unknowns! writes those), zero at the
    # Dirichlet dofs; hvp(asm, v) then sets the constrained entries to v
    Vfield = Vf.copy()
    Vfield[dd - 1] = 0.0
    KV_ref = orc.assemble_action(k, wl.mesh.coords, wl.mesh.conn, wl.nf, wl.physics.kind, props, U, Vfield)

    fec.assemble_vector_(asm, fec.residual, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    fec.assemble_stiffness_(asm, fec.stiffness, Uu, p)
    K = fec.stiffness(asm)
    K2 = fec.stiffness(asm)  # a second accessor call returns the same matrix (the adjustment is applied once)
    ptr, idx = (K.rowptr, K.colval) if sp_type == "csr" else (K.colptr, K.rowval)
    assert np.array_equal(ptr.cpu().numpy() + 1, rp) and np.array_equal(idx.cpu().numpy() + 1, cv)
    Kv = K.values.cpu().numpy().copy()
    assert np.array_equal(K2.values.cpu().numpy(), Kv)
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    KV = fec.hvp(asm, Vu).cpu().numpy().copy()
    KV_ref_adj = orc.adjust_action(KV_ref.copy(), c, Vf)
    errs = {"R": (rel(R, R_ref), rel_max(R, R_ref)), "K": (rel(Kv, K_ref), rel_max(Kv, K_ref)),
            "KV": (rel(KV, KV_ref_adj), rel_max(KV, KV_ref_adj))}
    print(f"[condensed] {case} {sp_type}: " + " ".join(f"{a}={b[0]:.2e}/{b[1]:.2e}" for a, b in errs.items()),
          flush=True)
    for a, (e2, em) in errs.items():
        assert e2 < TOL and em < TOL, (a, e2, em)
    # constrained lines: zero except the penalty on the diagonal, exactly as the oracle forms it from the GPU's trace
    A = (K.to_scipy()).tocsr()
    assert A.shape == (N, N)
    con = dd - 1
    dg = A.diagonal()
    assert np.all(dg[con] == dg[con[0]]) and dg[con[0]] > 0
    assert np.array_equal(R[con], np.zeros(len(con))) and np.array_equal(KV[con], Vf[con])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["poisson_quad4", "nh_hex8_partial"])
def test_condensed_consistent_with_reduced(fec, orc, gpu, case):
    """The reference's consistency contract across modes: the condensed results at the free dofs equal the reduced
    ones (bit for bit: the same element sums in the same order), for the residual, the Jacobian block and the action."""
    print(f"[condensed vs reduced] {case}", flush=True)
    wl = _poisson_quad4(fec) if case == "poisson_quad4" else _nh_hex8_partial(fec)
    N = wl.nf * wl.mesh.num_nodes
    dd = wl.dirichlet_dofs
    if case == "nh_hex8_partial":  # the reduced mode needs whole nodes: constrain every component of those nodes
        nodes = np.unique((dd - 1) // 3)
        dd = np.sort(np.concatenate([3 * nodes + k + 1 for k in range(3)])).astype(np.int64)
        wl.dirichlet_dofs = dd
    U = wl.U.copy()
    U[dd - 1] = 0.0
    free = np.setdiff1d(np.arange(N), dd - 1)
    Vf = np.cos(np.arange(N))
    asm_c, p_c = _gpu_setup(fec, wl, gpu, "csr", True)
    asm_r, p_r = _gpu_setup(fec, wl, gpu, "csr", False)
    Uc, Vc = torch.from_numpy(U).to(gpu), torch.from_numpy(Vf).to(gpu)
    Ur, Vr = torch.from_numpy(U[free]).to(gpu), torch.from_numpy(Vf[free]).to(gpu)
    fec.assemble_vector_and_stiffness_(asm_c, Uc, p_c)
    fec.assemble_vector_and_stiffness_(asm_r, Ur, p_r)
    Rc, Rr = fec.residual(asm_c).cpu().numpy(), fec.residual(asm_r).cpu().numpy()
    assert np.array_equal(Rc[free], Rr)
    Kc, Kr = fec.stiffness(asm_c).to_scipy(), fec.stiffness(asm_r).to_scipy()
    assert (Kc[free][:, free] != Kr).nnz == 0
    fec.assemble_matrix_action_(asm_c, fec.stiffness, Uc, Vc, p_c)
    fec.assemble_matrix_action_(asm_r, fec.stiffness, Ur, Vr, p_r)
    KVc, KVr = fec.hvp(asm_c, Vc).cpu().numpy(), fec.hvp(asm_r, Vr).cpu().numpy()
    # the reduced action uses V = 0 at the Dirichlet dofs, the condensed one the same (update_field_unknowns
    # writes only the unknown dofs of the scratch field)
    assert np.array_equal(KVc[free], KVr)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["poisson_quad4", "nh_hex8_partial"])
def test_csc_matches_oracle_and_is_symmetric(fec, orc, gpu, case):
    """sparse_matrix_type = :csc in the reduced mode against the oracle's sparse! CSC; the CSR values are symmetric
    bit for bit, which is what lets the CSC share them."""
    print(f"[csc] {case}", flush=True)
    wl = _poisson_quad4(fec) if case == "poisson_quad4" else _nh_hex8_partial(fec)
    if case == "nh_hex8_partial":
        nodes = np.unique((wl.dirichlet_dofs - 1) // 3)
        wl.dirichlet_dofs = np.sort(np.concatenate([3 * nodes + k + 1 for k in range(3)])).astype(np.int64)
    asm, p = _gpu_setup(fec, wl, gpu, "csc", False)
    N = wl.nf * wl.mesh.num_nodes
    unk, d2u = orc.update_dofs(N, wl.dirichlet_dofs)
    U = wl.U.copy()
    U[wl.dirichlet_dofs - 1] = 0.0
    st = _storage(orc, wl, U)
    Is, Js, slots, _ = orc.pattern(wl.nf, wl.mesh.conn, d2u)
    cp, rv, nzv = orc.sparse_csc(len(unk), Is, Js, orc.gather_slots(slots, st))
    fec.assemble_stiffness_(asm, fec.stiffness, torch.from_numpy(U[unk - 1]).to(gpu), p)
    K = fec.stiffness(asm)
    assert isinstance(K, fec.CSCMatrix)
    assert np.array_equal(K.colptr.cpu().numpy() + 1, cp) and np.array_equal(K.rowval.cpu().numpy() + 1, rv)
    nz = K.nzval.cpu().numpy()
    e2, em = rel(nz, nzv), rel_max(nz, nzv)
    print(f"[csc] {case}: K={e2:.2e}/{em:.2e}", flush=True)
    assert e2 < TOL and em < TOL
    A = K.to_scipy()
    assert (A != A.T).nnz == 0
