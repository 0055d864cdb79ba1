"""Pins of the FP64 oracle (oracle/oracle.cpp) against the reference's own fixtures and against what the mathematics
fixes: known-answer mesh tests, closed-form element matrices from an independent numpy quadrature, finite differences
of the stored energies, brute-force sparsity, an independent numpy restatement of the pattern, and the reference's
poisson.gold solution. Each check is chosen so a plausible slip (dropped term, wrong sign or index, transposed
operand) fails it."""
import itertools
import os

import numpy as np
import pytest
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ meshes
def test_quad_structured_mesh_kat(orc):
    # test/TestMesh.jl:97-108
    X, conn = orc.structured_mesh(orc.QUAD4, (0.0, 0.0), (1.0, 1.0), (3, 3))
    np.testing.assert_allclose(X.T, [[0.0, 0.5, 1.0, 0.0, 0.5, 1.0, 0.0, 0.5, 1.0],
                                     [0.0, 0.0, 0.0, 0.5, 0.5, 0.5, 1.0, 1.0, 1.0]])
    assert np.array_equal(conn.T, [[1, 4, 2, 5], [2, 5, 3, 6], [5, 8, 6, 9], [4, 7, 5, 8]])


def test_hex_structured_mesh_nodesets(fec):
    # test/TestMesh.jl:78-86 (coordinates of the six hex nodesets)
    m = fec.StructuredMesh("hex", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (3, 3, 3))
    c = m.nodal_coords
    for name, d, v in [("bottom", 1, 0.0), ("right", 0, 1.0), ("front", 2, 1.0), ("top", 1, 1.0),
                       ("left", 0, 0.0), ("back", 2, 0.0)]:
        assert np.allclose(c[d, m.nodeset_nodes[name] - 1], v), name


@pytest.mark.parametrize("kind,counts", [("quad", (3, 3)), ("quad", (6, 6)), ("hex", (3, 3, 3)), ("hex", (5, 5, 5))])
def test_python_mesh_matches_oracle(fec, orc, kind, counts):
    m = fec.StructuredMesh(kind, (0.0,) * len(counts), (1.0,) * len(counts), counts)
    X, conn = orc.structured_mesh(orc.QUAD4 if kind == "quad" else orc.HEX8, (0.0,) * len(counts),
                                  (1.0,) * len(counts), counts)
    assert np.array_equal(m.coords, X)
    assert np.array_equal(m.conn, conn)


def test_tet10_mesh_geometry(fec):
    m = fec.kuhn_tet10_mesh(3, jitter=0.2, seed=0)
    assert m.conn.shape == (3 ** 3 * 6, 10) and m.coords.shape == (7 ** 3, 3)
    assert np.array_equal(np.unique(m.conn), np.arange(1, 7 ** 3 + 1))  # every refined-grid node used
    X = m.coords
    vol = 0.0
    for c in m.conn - 1:
        v = X[c[:4]]
        d = np.linalg.det(np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))
        assert d > 0
        vol += d / 6
        for k, (i, j) in enumerate([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]):
            np.testing.assert_allclose(X[c[4 + k]], 0.5 * (v[i] + v[j]), atol=1e-15)
    assert abs(vol - 1.0) < 1e-12


# ------------------------------------------------------------------ reference elements
_REF_VOL = {0: 4.0, 1: 8.0, 2: 1.0 / 6.0}


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_shape_function_invariants(orc, kind):
    xi, w, N, dN = orc.ref_element(kind)
    np.testing.assert_allclose(N.sum(axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(dN.sum(axis=1), 0.0, atol=1e-13)
    np.testing.assert_allclose(w.sum(), _REF_VOL[kind], rtol=1e-14)
    # linear reproduction: sum_a N_a xi_a = xi and sum_a dN_a xi_a = I
    nodes = {0: np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float),
             1: np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                          [-1, 1, 1]], float)}
    if kind == 2:
        v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
        e = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
        nodes[2] = np.vstack([v, [(v[i] + v[j]) / 2 for i, j in e]])
    P = nodes[kind]
    nd = P.shape[1]
    np.testing.assert_allclose(N @ P, xi[:, :nd], atol=1e-14)
    for q in range(len(w)):
        np.testing.assert_allclose(P.T @ dN[q], np.eye(nd), atol=1e-13)
    # quadrature exactness for quadratics
    if kind == 2:
        assert abs(np.sum(w * xi[:, 0] ** 2) - 1.0 / 60.0) < 1e-15
        assert abs(np.sum(w * xi[:, 0] * xi[:, 1]) - 1.0 / 120.0) < 1e-15
    else:
        assert abs(np.sum(w * xi[:, 0] ** 2) - _REF_VOL[kind] / 3.0) < 1e-14


def _np_q1_stiffness(nd, h=1.0):
    """Independent Q1 Laplace element stiffness on a cube/square of side h, 3-point Gauss per direction."""
    g, gw = np.polynomial.legendre.leggauss(3)
    s = np.array(list(itertools.product([-1, 1], repeat=nd)))[:, ::-1]  # lexicographic, x fastest
    order = [0, 1, 3, 2] if nd == 2 else [0, 1, 3, 2, 4, 5, 7, 6]  # counter-clockwise local order
    s = s[order]
    K = np.zeros((len(s), len(s)))
    for pt, wt in zip(itertools.product(g, repeat=nd), itertools.product(gw, repeat=nd)):
        pt = np.array(pt)
        dN = np.zeros((len(s), nd))
        for a in range(len(s)):
            f = 1 + s[a] * pt
            for j in range(nd):
                dN[a, j] = s[a, j] * np.prod(np.delete(f, j)) / 2 ** nd
        G = dN * (2.0 / h)
        K += np.prod(wt) * (h / 2) ** nd * G @ G.T
    return K


@pytest.mark.parametrize("nd", [2, 3])
def test_poisson_element_stiffness_closed_form(orc, nd):
    kind = orc.QUAD4 if nd == 2 else orc.HEX8
    X, conn = orc.structured_mesh(kind, (0.0,) * nd, (0.5,) * nd, (2,) * nd)
    st = orc.assemble_stiffness_storage(kind, X, conn, 1, orc.POISSON, [0.0, 0.0], np.zeros(len(X)))
    K = st.reshape(2 ** nd, 2 ** nd)
    np.testing.assert_allclose(K, _np_q1_stiffness(nd, 0.5), atol=1e-14)
    if nd == 2:  # textbook Q1 square stiffness, independent of h in 2D
        np.testing.assert_allclose(K, np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1],
                                                [-1, -2, -1, 4]]) / 6.0, atol=1e-14)


# ------------------------------------------------------------------ materials: finite differences
def _psi_nh(F, lam, mu):
    J = np.linalg.det(F)
    return 0.5 * mu * (np.trace(F.T @ F) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2


def _psi_le(H, K, G):
    eps = 0.5 * (H + H.T)
    dev = eps - np.trace(eps) / 3 * np.eye(3)
    return 0.5 * K * np.trace(eps) ** 2 + G * np.sum(dev * dev)


def _fd_grad(f, H, h=1e-6):
    g = np.zeros(9)
    for c in range(9):
        E = np.zeros(9)
        E[c] = h
        g[c] = (f((H.ravel() + E).reshape(3, 3)) - f((H.ravel() - E).reshape(3, 3))) / (2 * h)
    return g


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neohookean_stress_and_tangent_fd(orc, seed):
    lam, mu = 1.0, 0.5
    H = 0.2 * np.random.default_rng(seed).standard_normal((3, 3))
    P, _, A, _ = orc.material_point(orc.NEOHOOKEAN, [lam, mu], 3, 3, H)
    np.testing.assert_allclose(P, _fd_grad(lambda G: _psi_nh(np.eye(3) + G, lam, mu), H), rtol=1e-7, atol=1e-9)
    Afd = np.stack([_fd_grad(lambda G: orc.material_point(orc.NEOHOOKEAN, [lam, mu], 3, 3, G)[0][r], H)
                    for r in range(9)])
    np.testing.assert_allclose(A, Afd, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(A, A.T, atol=1e-13)
    P0, _, A0, _ = orc.material_point(orc.NEOHOOKEAN, [lam, mu], 3, 3, np.zeros((3, 3)))
    assert np.abs(P0).max() < 1e-15
    # small-strain limit: A(0) = lambda I(x)I + mu (d_ik d_jl + d_il d_jk)
    I = np.eye(3)
    lin = lam * np.einsum("ij,kl->ijkl", I, I) + mu * (np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I))
    np.testing.assert_allclose(A0, lin.reshape(9, 9), atol=1e-14)


def test_linear_elastic_stress_and_tangent_fd(orc):
    K, G = 10.0, 1.0
    H = 0.01 * np.random.default_rng(3).standard_normal((3, 3))
    P, _, A, _ = orc.material_point(orc.LINEAR_ELASTIC, [1e3, K, G], 3, 3, H)
    np.testing.assert_allclose(P, _fd_grad(lambda X: _psi_le(X, K, G), H), rtol=1e-6, atol=1e-10)
    Afd = np.stack([_fd_grad(lambda X: orc.material_point(orc.LINEAR_ELASTIC, [1e3, K, G], 3, 3, X)[0][r], H)
                    for r in range(9)])
    np.testing.assert_allclose(A, Afd, rtol=1e-6, atol=1e-8)


J2P = [200.0, 0.3, 0.25, 2.0]


def test_j2_elastic_predictor_and_yield_consistency(orc):
    E, nu, sy, Hh = J2P
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    rng = np.random.default_rng(4)
    # elastic: stress = lam tr(eps) I + 2 mu eps, state unchanged
    H = 1e-4 * rng.standard_normal((3, 3))
    P, _, A, sn = orc.material_point(orc.J2, J2P, 3, 3, H)
    eps = 0.5 * (H + H.T)
    np.testing.assert_allclose(P.reshape(3, 3), lam * np.trace(eps) * np.eye(3) + 2 * mu * eps, atol=1e-12)
    assert np.all(sn == 0)
    # plastic: the returned state sits on the yield surface
    for s in range(5):
        H = 0.01 * rng.standard_normal((3, 3))
        so = np.concatenate([1e-3 * rng.standard_normal(6), [abs(1e-3 * rng.standard_normal())]])
        P, _, A, sn = orc.material_point(orc.J2, J2P, 3, 3, H, state_old=so)
        S = P.reshape(3, 3)
        dev = S - np.trace(S) / 3 * np.eye(3)
        f = np.linalg.norm(dev) - np.sqrt(2 / 3) * (sy + Hh * sn[6])
        if sn[6] > so[6]:
            assert abs(f) < 1e-12
            # plastic strain increment is deviatoric and along the stress deviator
            dep = np.array([[sn[0] - so[0], sn[1] - so[1], sn[2] - so[2]], [sn[1] - so[1], sn[3] - so[3], sn[4] - so[4]],
                            [sn[2] - so[2], sn[4] - so[4], sn[5] - so[5]]])
            assert abs(np.trace(dep)) < 1e-14
            np.testing.assert_allclose(dep / np.linalg.norm(dep), dev / np.linalg.norm(dev), atol=1e-10)
            assert abs(np.linalg.norm(dep) - np.sqrt(3 / 2) * (sn[6] - so[6])) < 1e-14
        else:
            assert f <= 1e-12


def test_j2_consistent_tangent_fd(orc):
    rng = np.random.default_rng(5)
    H = 0.006 * rng.standard_normal((3, 3))
    so = np.zeros(7)
    P, _, A, sn = orc.material_point(orc.J2, J2P, 3, 3, H, state_old=so)
    assert sn[6] > 0  # plastic step
    Afd = np.stack([_fd_grad(lambda X: orc.material_point(orc.J2, J2P, 3, 3, X, state_old=so)[0][r], H, h=1e-8)
                    for r in range(9)])
    np.testing.assert_allclose(A, Afd, rtol=1e-5, atol=1e-4)


def test_j2_uniaxial_closed_form(orc):
    # proportional deviatoric strain path e = t*diag(1,-1/2,-1/2): |dev sigma| = sqrt(2/3)(sy + H alpha) and
    # alpha follows from the radial return in closed form
    E, nu, sy, Hh = J2P
    mu = E / (2 * (1 + nu))
    t = 0.01
    H = t * np.diag([1.0, -0.5, -0.5])
    P, _, _, sn = orc.material_point(orc.J2, J2P, 3, 3, H, state_old=np.zeros(7))
    s_tr = 2 * mu * t * np.sqrt(1.5)
    dg = (s_tr - np.sqrt(2 / 3) * sy) / (2 * mu + 2 * Hh / 3)
    assert abs(sn[6] - np.sqrt(2 / 3) * dg) < 1e-15
    assert abs(P[0] - P[4] - (s_tr - 2 * mu * dg) * np.sqrt(1.5)) < 1e-12


# ------------------------------------------------------------------ pattern and sparse!
def _np_pattern(nf, conn, d2u):
    """Independent numpy restatement of SparsityPatterns.jl:193-229."""
    ne, nnpe = conn.shape
    ndof = nf * nnpe
    dc = (nf * (conn[:, :, None] - 1) + np.arange(nf)[None, None, :] + 1).reshape(ne, ndof)
    I = np.repeat(dc, ndof, axis=1).ravel()
    J = np.tile(dc, (1, ndof)).ravel()
    slots = np.arange(1, I.size + 1)
    rI, rJ = d2u[I - 1], d2u[J - 1]
    keep = (rI > 0) & (rJ > 0)
    Is, Js, slots = rI[keep], rJ[keep], slots[keep]
    perm = np.argsort((Is << 32) | Js, kind="stable") + 1
    return Is, Js, slots, perm


@pytest.mark.parametrize("case", ["quad_bc", "hex_nf1", "hex_nf3", "tet10_nf3"])
def test_pattern_vs_numpy_and_brute_force(fec, orc, case):
    if case == "quad_bc":
        wl = fec.workloads.poisson_quad4(4)
        mesh, nf, dd = wl.mesh, 1, wl.dirichlet_dofs
    elif case == "hex_nf1":
        mesh, nf, dd = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (4, 4, 4)), 1, np.zeros(0, np.int64)
    elif case == "hex_nf3":
        mesh, nf = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (3, 4, 3)), 3
        b = fec.boundary_nodes(mesh)[:5]
        dd = np.sort(np.concatenate([3 * (b - 1) + d + 1 for d in range(3)]))
    else:
        mesh, nf, dd = fec.kuhn_tet10_mesh(2), 3, np.zeros(0, np.int64)
    N = nf * mesh.num_nodes
    unknown, d2u = orc.update_dofs(N, dd)
    assert len(unknown) + len(dd) == N and np.all(d2u[dd - 1] == -1)
    Is, Js, slots, perm = orc.pattern(nf, mesh.conn, d2u)
    nI, nJ, nslots, nperm = _np_pattern(nf, mesh.conn, d2u)
    assert np.array_equal(Is, nI) and np.array_equal(Js, nJ) and np.array_equal(slots, nslots)
    assert np.array_equal(perm, nperm)
    V = np.random.default_rng(0).standard_normal(Is.size)
    rowptr, colval, vals, cmap = orc.sparse_csr(len(unknown), Is, Js, V, want_map=True)
    ref = sp.coo_matrix((V, (Is - 1, Js - 1)), shape=(len(unknown),) * 2).tocsr()
    ref.sort_indices()
    assert np.array_equal(rowptr - 1, ref.indptr) and np.array_equal(colval - 1, ref.indices)
    np.testing.assert_allclose(vals, ref.data, rtol=1e-13, atol=1e-13)
    # the map sends each COO entry to the CSR slot of its (I, J)
    assert np.array_equal(colval[cmap] , Js)
    rows = np.searchsorted(rowptr - 1, cmap, side="right")
    assert np.array_equal(rows, Is)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_structured_nnz_closed_form(fec, orc, n):
    for nf in (1, 3):
        mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (n, n, n))
        _, d2u = orc.update_dofs(nf * mesh.num_nodes, np.zeros(0, np.int64))
        Is, Js, _, _ = orc.pattern(nf, mesh.conn, d2u)
        rp, _, _, _ = orc.sparse_csr(nf * mesh.num_nodes, Is, Js)
        assert rp[-1] - 1 == (3 * n - 2) ** 3 * nf * nf


# ------------------------------------------------------------------ assembled operators
def test_poisson_patch_and_nullspace(fec, orc):
    for kind, mesh in [(orc.QUAD4, fec.StructuredMesh("quad", (0, 0), (1, 1), (5, 5))),
                       (orc.HEX8, fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (4, 4, 4))),
                       (orc.TET10, fec.kuhn_tet10_mesh(2))]:
        X = mesh.coords
        nd = X.shape[1]
        U = 0.3 + X @ np.arange(1, nd + 1)  # affine
        R = orc.assemble_residual(kind, X, mesh.conn, 1, orc.POISSON, [0.0, 0.0], U)
        interior = np.setdiff1d(np.arange(1, len(X) + 1), fec.boundary_nodes(mesh)) - 1
        assert np.abs(R[interior]).max() < 1e-12
        _, d2u = orc.update_dofs(len(X), np.zeros(0, np.int64))
        rp, cv, vals = orc.stiffness_csr(kind, X, mesh.conn, 1, orc.POISSON, [0.0, 0.0], U, d2u, len(X))
        Kd = sp.csr_matrix((vals, cv - 1, rp - 1)).toarray()
        np.testing.assert_allclose(Kd.sum(axis=1), 0.0, atol=1e-12)
        np.testing.assert_allclose(Kd, Kd.T, atol=1e-14)
        # residual is K u - f for Poisson (linear): R(U) - R(0) = K U
        R0 = orc.assemble_residual(kind, X, mesh.conn, 1, orc.POISSON, [0.0, 1.0], np.zeros(len(X)))
        R1 = orc.assemble_residual(kind, X, mesh.conn, 1, orc.POISSON, [0.0, 1.0], U)
        np.testing.assert_allclose(R1 - R0, Kd @ U, atol=1e-12)


@pytest.mark.parametrize("phys", ["le", "nh"])
def test_mechanics_patch_and_tangent_fd(fec, orc, phys):
    mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (3, 3, 3))
    X = mesh.coords
    kind, props = (orc.LINEAR_ELASTIC, [1e3, 10.0, 1.0]) if phys == "le" else (orc.NEOHOOKEAN, [1.0, 0.5])
    Fm = np.array([[0.02, 0.01, -0.03], [0.0, -0.01, 0.02], [0.01, 0.0, 0.015]])
    U = (X @ Fm.T).ravel()
    R = orc.assemble_residual(orc.HEX8, X, mesh.conn, 3, kind, props, U)
    interior = np.setdiff1d(np.arange(1, len(X) + 1), fec.boundary_nodes(mesh)) - 1
    assert np.abs(R.reshape(-1, 3)[interior]).max() < 1e-13
    # tangent = dR/dU by central differences; matrix-free and assembled actions = K v
    U = U + 0.01 * np.random.default_rng(1).standard_normal(U.size)
    N = U.size
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    rp, cv, vals = orc.stiffness_csr(orc.HEX8, X, mesh.conn, 3, kind, props, U, d2u, N)
    Kd = sp.csr_matrix((vals, cv - 1, rp - 1)).toarray()
    v = np.random.default_rng(2).standard_normal(N)
    h = 1e-6
    fd = (orc.assemble_residual(orc.HEX8, X, mesh.conn, 3, kind, props, U + h * v) -
          orc.assemble_residual(orc.HEX8, X, mesh.conn, 3, kind, props, U - h * v)) / (2 * h)
    np.testing.assert_allclose(Kd @ v, fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())
    for mf in (False, True):
        KV = orc.assemble_action(orc.HEX8, X, mesh.conn, 3, kind, props, U, v, matrix_free=mf)
        np.testing.assert_allclose(KV, Kd @ v, rtol=1e-12, atol=1e-12 * np.abs(KV).max())


def test_poisson_quad4_64_cg_analytic(fec, orc):
    wl = fec.workloads.poisson_quad4(64)
    mesh = wl.mesh
    N = mesh.num_nodes
    unknown, d2u = orc.update_dofs(N, wl.dirichlet_dofs)
    R = orc.assemble_residual(orc.QUAD4, mesh.coords, mesh.conn, 1, orc.POISSON, wl.physics.props(), np.zeros(N))
    rp, cv, vals = orc.stiffness_csr(orc.QUAD4, mesh.coords, mesh.conn, 1, orc.POISSON, wl.physics.props(),
                                     np.zeros(N), d2u, len(unknown))
    x, it, rel = orc.cg(rp, cv, vals, -R[unknown - 1], rtol=1e-12)
    assert rel < 1e-12
    X = mesh.coords[unknown - 1]
    exact = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
    assert np.abs(x - exact).max() < 1e-3  # O(h^2) nodal error at h = 1/64


def test_poisson_gold(fec, orc):
    """The reference's regression gold (test/poisson/poisson.gold, 128 x 128 quad4) reproduced by the oracle's
    residual + CSR + CG path on the structured mesh of the same geometry (nodes matched by coordinates)."""
    g = np.load(os.path.join(HERE, "golden", "poisson_gold_subset.npz"))
    wl = fec.workloads.poisson_quad4(128)
    mesh = wl.mesh
    N = mesh.num_nodes
    unknown, d2u = orc.update_dofs(N, wl.dirichlet_dofs)
    R = orc.assemble_residual(orc.QUAD4, mesh.coords, mesh.conn, 1, orc.POISSON, wl.physics.props(), np.zeros(N))
    rp, cv, vals = orc.stiffness_csr(orc.QUAD4, mesh.coords, mesh.conn, 1, orc.POISSON, wl.physics.props(),
                                     np.zeros(N), d2u, len(unknown))
    x, it, rel = orc.cg(rp, cv, vals, -R[unknown - 1], rtol=1e-13)
    u = np.zeros(N)
    u[unknown - 1] = x
    idx = np.rint(g["x"] * 128).astype(int) + 129 * np.rint(g["y"] * 128).astype(int)
    np.testing.assert_allclose(mesh.coords[idx, 0], g["x"], atol=1e-14)
    np.testing.assert_allclose(mesh.coords[idx, 1], g["y"], atol=1e-14)
    assert np.abs(u[idx] - g["u"]).max() < 1e-10


@pytest.mark.parametrize("amp", [0.05, 0.1])
def test_field_seed_rule(fec, amp):
    """The neo-Hookean field seeds are the first seeds whose sine field keeps det(I + grad u) > 0.2 (checked here on
    a coarse grid for the seeds below the chosen one, on the full grid for the chosen one)."""
    W = fec.workloads
    s = W.FIELD_SEED[amp]
    assert W.sine_field_min_det(amp, s) > 0.2
    for t in range(0, s, max(1, s // 40)):
        a = W.sine_field_coefficients(t)
        if np.abs(a[3]).max() * amp * 4 * np.pi > 0.9:
            continue  # the k = 4 term alone folds the field
        assert W.sine_field_min_det(amp, t, pts=41) <= 0.2, t


# ------------------------------------------------------------------ mass, lumped mass, diagonals
def _single_tet10(v):
    """One straight-sided tet10 (Exodus TETRA10 node order) on vertices v (4 x 3)."""
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    X = np.vstack([v] + [0.5 * (v[a] + v[b]) for a, b in edges])
    return X, np.arange(1, 11, dtype=np.int64)[None, :]


@pytest.mark.parametrize("nd", [2, 3])
def test_mass_element_closed_form(orc, nd):
    # bilinear / trilinear element on a box: 2x2(x2) Gauss is exact for N_a N_b, and
    # M_ab = rho V / 6^nd prod_d (2 if a and b share coordinate d else 1)
    lo, hi = (0.0,) * nd, (2.0, 3.0, 0.5)[:nd]
    kind = orc.QUAD4 if nd == 2 else orc.HEX8
    X, conn = orc.structured_mesh(kind, lo, hi, (2,) * nd)
    rho, V = 1.7, float(np.prod(hi))
    for nf in (1, nd):
        st = orc.assemble_mass_storage(kind, X, conn, nf, rho)
        ndof = conn.shape[1] * nf
        M = st.reshape(ndof, ndof)
        xl = X[conn[0] - 1]
        same = (xl[:, None, :] == xl[None, :, :])
        Mn = rho * V / 6.0 ** nd * np.prod(np.where(same, 2.0, 1.0), axis=2)
        np.testing.assert_allclose(M, np.kron(Mn, np.eye(nf)), rtol=1e-13, atol=1e-15)


def test_tet10_lumped_mass_closed_form(orc):
    # integral of a quadratic tet's vertex function is -V/20, of an edge function V/5 (the degree-2 rule is exact)
    v = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 3.0, 0.0], [0.0, 0.0, 1.5]])
    X, conn = _single_tet10(v)
    V = 1.5
    m = orc.assemble_lumped_mass(orc.TET10, X, conn, 1, 2.0)
    np.testing.assert_allclose(m, 2.0 * V * np.array([-1 / 20] * 4 + [1 / 5] * 6), rtol=1e-13)


@pytest.mark.parametrize("case", ["quad4", "hex8_nf3", "tet10_nf3"])
def test_lumped_is_row_sum_and_diagonals(fec, orc, case):
    # the reference's lumped-mass contract (test/TestAssemblers.jl:432-521): row sums of the consistent mass, total
    # mass rho V per component; the diagonal assemblies equal the diagonals of the sparse! CSR matrices; the mass
    # action equals CSR x v
    if case == "quad4":
        mesh, nf, kind = fec.StructuredMesh("quad", (0, 0), (1, 2), (5, 5)), 1, orc.QUAD4
    elif case == "hex8_nf3":
        mesh, nf, kind = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (4, 4, 4)), 3, orc.HEX8
    else:
        mesh, nf, kind = fec.kuhn_tet10_mesh(2), 3, orc.TET10
    X, conn = mesh.coords, mesh.conn
    vol = {"quad4": 2.0, "hex8_nf3": 1.0, "tet10_nf3": 1.0}[case]
    rho = 3.0
    N = nf * mesh.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    rp, cv, mv = orc.mass_csr(kind, X, conn, nf, rho, d2u, N)
    M = sp.csr_matrix((mv, cv - 1, rp - 1), shape=(N, N))
    lm = orc.assemble_lumped_mass(kind, X, conn, nf, rho)
    np.testing.assert_allclose(lm, np.asarray(M.sum(axis=1)).ravel(), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(lm.sum(), rho * vol * nf, rtol=1e-12)
    np.testing.assert_allclose(orc.assemble_diagonal(kind, X, conn, nf, 0, [], None, which="mass", rho=rho),
                               M.diagonal(), rtol=1e-13, atol=1e-15)
    v = np.sin(np.arange(N) * 0.37)
    np.testing.assert_allclose(orc.assemble_mass_action(kind, X, conn, nf, rho, v), M @ v, rtol=1e-12, atol=1e-14)
    phys = orc.POISSON if nf == 1 else orc.NEOHOOKEAN
    props = [1.0, 1.0] if nf == 1 else [1.0, 0.5]
    U = 0.02 * np.cos(np.arange(N) * 0.91)
    krp, kcv, kv = orc.stiffness_csr(kind, X, conn, nf, phys, props, U, d2u, N)
    K = sp.csr_matrix((kv, kcv - 1, krp - 1), shape=(N, N))
    np.testing.assert_allclose(orc.assemble_diagonal(kind, X, conn, nf, phys, props, U), K.diagonal(), rtol=1e-13,
                               atol=1e-15)
