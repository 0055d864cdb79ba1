"""GPU parity of the mass-type assemblies against the FP64 oracle: the consistent mass (CSR on the Jacobian's pattern),
the row-sum lumped mass, the consistent-mass diagonal, the mass action and the stiffness diagonal
(src/assemblers/{Matrix,LumpedMass,Diagonal,MatrixAction}.jl with the test physics' mass callbacks,
test/mechanics/TestMechanicsCommon.jl:52-124, test/poisson/TestPoissonCommon.jl:18-41).

Tolerance: 1e-6 relative error (the larger of the 2-norm and max-norm errors; FP32 element arithmetic, FP64 sums),
as in test_gpu_parity.py.
The reference's lumped-mass contract (test/TestAssemblers.jl:432-521) is checked too: the BC-restricted lumped mass is
the full one at the free dofs (bit for bit here: the same kernel and summation order)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = 1e-6


def rel(a, b):
    """max of the relative 2-norm and max-norm errors"""
    e2 = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    em = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) if len(b) else 0.0
    return max(e2, em)


def _case(fec, name):
    W = fec.workloads
    if name == "poisson_quad4":
        return W.poisson_quad4(10)
    if name == "poisson_hex8":
        wl = W.poisson_hex8(5)
        wl.dirichlet_dofs = fec.boundary_nodes(wl.mesh)
        return wl
    if name == "poisson_tet10":
        wl = W.neohookean_tet10(2)
        return W.Workload("poisson_tet10", wl.mesh, fec.Poisson("sines", 1.0), 1,
                          2 * fec.rng.uniform(0, 1, wl.mesh.num_nodes) - 1, fec.boundary_nodes(wl.mesh))
    if name == "mechanics_hex8":
        wl = W.neohookean_hex8(4)
        nodes = wl.mesh.nodeset_nodes["left"]
        dd = np.sort(np.concatenate([3 * (nodes - 1) + d + 1 for d in range(3)]))
        return W.Workload("mechanics_hex8", wl.mesh, fec.Mechanics(rho=7.8e3, K=1.0, G=0.5), 3, 1e-3 * wl.U, dd)
    if name == "neohookean_tet10":
        return W.neohookean_tet10(2)
    if name == "j2_hex8":
        return W.j2_hex8(3, dirichlet=True)
    raise KeyError(name)


def _setup(fec, wl, gpu, with_bcs=True):
    mesh = wl.mesh
    V = fec.FunctionSpace(mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    bcs = []
    if with_bcs and len(wl.dirichlet_dofs):
        nodes = np.unique((wl.dirichlet_dofs - 1) // wl.nf + 1)
        mesh.nodeset_nodes["_bc"] = nodes
        bcs = [fec.DirichletBC(name, lambda X, t: 0.0, nset_name="_bc") for name in var.names]
    p = fec.create_parameters(mesh, asm, wl.physics, dirichlet_bcs=bcs)
    return asm, p


CASES = ["poisson_quad4", "poisson_hex8", "poisson_tet10", "mechanics_hex8", "neohookean_tet10", "j2_hex8"]


@pytest.mark.parametrize("name", CASES)
def test_mass_assemblies_match_oracle(fec, orc, gpu, name):
    print(f"[gpu mass] {name}", flush=True)
    wl = _case(fec, name)
    mesh = wl.mesh
    kind = {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8, "TET10": orc.TET10}[mesh.element_type]
    asm, p = _setup(fec, wl, gpu)
    rho = fec.density(wl.physics)
    unknown = asm.dof.unknown_dofs
    N = wl.nf * mesh.num_nodes
    _, d2u = orc.update_dofs(N, asm.dof.dirichlet_dofs)
    U = wl.U.copy()
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vv = np.zeros(N)
    Vv[unknown - 1] = np.cos(0.7 * np.arange(len(unknown)))
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)

    fec.assemble_mass_(asm, fec.mass, Uu, p)
    M = fec.mass(asm)
    rp, cv, mv = orc.mass_csr(kind, mesh.coords, mesh.conn, wl.nf, rho, d2u, len(unknown))
    assert np.array_equal(M.rowptr.cpu().numpy() + 1, rp) and np.array_equal(M.colval.cpu().numpy() + 1, cv)
    errs = {"M": rel(M.values.cpu().numpy(), mv)}

    fec.assemble_lumped_mass_(asm, fec.lumped_mass, Uu, p)
    lm = fec.lumped_mass(asm).cpu().numpy().copy()
    errs["lumped"] = rel(lm, orc.assemble_lumped_mass(kind, mesh.coords, mesh.conn, wl.nf, rho)[unknown - 1])

    fec.assemble_diagonal_(asm, fec.mass, Uu, p)
    errs["Mdiag"] = rel(fec.diagonal(asm).cpu().numpy(),
                        orc.assemble_diagonal(kind, mesh.coords, mesh.conn, wl.nf, 0, [], None, which="mass",
                                              rho=rho)[unknown - 1])

    fec.assemble_matrix_action_(asm, fec.mass, Uu, Vu, p)
    errs["MV"] = rel(fec.hvp(asm, Vu).cpu().numpy(),
                     orc.assemble_mass_action(kind, mesh.coords, mesh.conn, wl.nf, rho, Vv)[unknown - 1])

    st_old = None
    if wl.physics.kind == orc.J2:
        st_old = np.zeros((mesh.num_elements, 8, 7))
        g = fec.rng.normal(2, 11, st_old.size).reshape(st_old.shape)
        st_old[..., :6] = 2e-3 * g[..., :6]
        st_old[..., 6] = 2e-3 * np.abs(g[..., 6])
        st_old = st_old.ravel().astype(np.float32)
        asm.ctx.call("fecx_set_state_old", st_old.ctypes.data_as(__import__("ctypes").c_void_p))
        st_old = st_old.astype(np.float64)
    fec.assemble_diagonal_(asm, fec.stiffness, Uu, p)
    kd = fec.diagonal(asm).cpu().numpy().copy()
    errs["Kdiag"] = rel(kd, orc.assemble_diagonal(kind, mesh.coords, mesh.conn, wl.nf, wl.physics.kind,
                                                  wl.physics.props(), U, st_old)[unknown - 1])
    fec.assemble_stiffness_(asm, fec.stiffness, Uu, p)
    K = fec.stiffness(asm).to_scipy()
    errs["Kdiag_vs_csr"] = rel(kd, K.diagonal())
    print(f"[gpu mass] {name}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < TOL for v in errs.values()), errs

    # BC-restricted lumped mass = the full one at the free dofs (test/TestAssemblers.jl:432-521)
    if len(asm.dof.dirichlet_dofs):
        asm_full, p_full = _setup(fec, wl, gpu, with_bcs=False)
        fec.assemble_lumped_mass_(asm_full, fec.lumped_mass, torch.zeros(N, dtype=torch.float64, device=gpu), p_full)
        full = fec.lumped_mass(asm_full).cpu().numpy()
        assert np.array_equal(full[unknown - 1], lm)
        np.testing.assert_allclose(full.sum(), rho * wl.nf, rtol=1e-6)  # every mesh here fills the unit square/cube
