"""The scatter designs of structured hex8 meshes: the fused kernel (fused_hex8_kernel: residual + CSR Jacobian in one
kernel, each CSR row summed from the element matrices of two element layers kept on chip), the two-phase path, FP64
atomics and element colouring (scatter_hex8_kernel).

CPU: fecx_build_pattern detects the StructuredMesh numbering (src/meshes/StructuredMesh.jl:85-131), including unequal
node counts per direction, and rejects a mesh whose connectivity does not follow it.
GPU: every design against the two-phase path on the same context (fecx_set_scatter) -- the same element matrices (the
same element code) and the same FP64 contributions in another order, so the results agree to FP64 rounding (2e-15
relative, max-norm) -- and the fused results against the FP64 oracle at 1e-6 (2-norm and max-norm); tile and z-range
edges are covered by sizes that are not multiples of the 7 x 3 node tiles, Dirichlet nodes by J2 with a clamped
boundary."""
import ctypes

import numpy as np
import pytest
import torch


def _errs(a, b):
    return max(float(np.linalg.norm(a - b) / np.linalg.norm(b)), float(np.abs(a - b).max() / np.abs(b).max()))


def _brick(fec, counts):
    """A hex8 brick with node counts (Nx, Ny, Nz), numbered as StructuredMesh numbers it (node i + Nx (j + Ny k),
    elements ex outer, ey, ez inner, local order StructuredMesh.jl:116-123) with exact coordinates for any counts
    (the reference's generator itself requires equal counts, StructuredMesh.jl:438-440)."""
    Nx, Ny, Nz = counts
    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    coords = np.stack([i.ravel() / (Nx - 1), j.ravel() / (Ny - 1), k.ravel() / (Nz - 1)], axis=1)
    ex, ey, ez = (a.ravel() for a in np.meshgrid(np.arange(Nx - 1), np.arange(Ny - 1), np.arange(Nz - 1),
                                                  indexing="ij"))
    node = lambda a, b, c: a + Nx * (b + Ny * c) + 1
    conn = np.stack([node(ex, ey, ez), node(ex + 1, ey, ez), node(ex + 1, ey + 1, ez), node(ex, ey + 1, ez),
                     node(ex, ey, ez + 1), node(ex + 1, ey, ez + 1), node(ex + 1, ey + 1, ez + 1),
                     node(ex, ey + 1, ez + 1)], axis=1).astype(np.int64)
    return fec.BlockMesh(coords, [("block_1", conn)], "HEX8")


FUSED, TWO_PHASE, ATOMIC, COLOURED, AUTO = 0, 1, 2, 3, 4


def _applies(fec, ctx):
    """1 if a residual + Jacobian call of the context takes the fused kernel"""
    out = ctypes.c_int(-1)
    ctx.call("fecx_get_scatter_path", ctypes.byref(out))
    return 1 if out.value == FUSED else 0


def test_structured_detection(fec):
    for counts in ((5, 5, 5), (9, 6, 4), (3, 8, 2)):
        mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), counts) if len(set(counts)) == 1 else _brick(fec, counts)
        hp = fec.HostPattern(mesh, 3)
        props = np.array([1.0, 0.5])
        hp.ctx.call("fecx_set_material", ctypes.c_int(2), props.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(2))
        assert _applies(fec, hp.ctx) == 1, counts
        hp.ctx.call("fecx_set_scatter", ctypes.c_int(TWO_PHASE))
        assert _applies(fec, hp.ctx) == 0
    # the same elements in another order are not the structured numbering
    mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (4, 4, 4))
    perm = np.arange(mesh.num_elements)[::-1]
    shuffled = fec.BlockMesh(mesh.coords, [("b", mesh.conn[perm])], "HEX8")
    hp = fec.HostPattern(shuffled, 1)
    props = np.array([0.0, 1.0])
    hp.ctx.call("fecx_set_material", ctypes.c_int(0), props.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(2))
    assert _applies(fec, hp.ctx) == 0
    tet = fec.workloads.neohookean_tet10(2).mesh
    hp = fec.HostPattern(tet, 3)
    props = np.array([1.0, 0.5])
    hp.ctx.call("fecx_set_material", ctypes.c_int(2), props.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(2))
    assert _applies(fec, hp.ctx) == 0


def _case(fec, name):
    W = fec.workloads
    if name == "poisson 17x9x12":
        mesh = _brick(fec, (17, 9, 12))
        U = 2 * fec.rng.uniform(0, 1, mesh.num_nodes) - 1
        return W.Workload("p", mesh, fec.Poisson("constant", 1.0), 1, U)
    if name == "neohookean 16^3":
        return W.neohookean_hex8(16)
    if name == "neohookean 20x8x31":
        wl = W.neohookean_hex8(4)
        mesh = _brick(fec, (21, 9, 32))
        return W.Workload("nh", mesh, wl.physics, 3, W._sine_field(mesh.coords, 0.05, 33))
    if name == "mechanics 11^3":
        wl = W.neohookean_hex8(11)
        return W.Workload("le", wl.mesh, fec.Mechanics(rho=1.0, K=3.0, G=1.0), 3, 0.01 * wl.U)
    if name == "j2 9^3 clamped":
        return W.j2_hex8(9, dirichlet=True)
    if name.startswith("neohookean ") and name.count("x") == 2:  # thin and tiny bricks (node counts)
        counts = tuple(int(c) for c in name.split()[1].split("x"))
        wl = W.neohookean_hex8(2)
        mesh = _brick(fec, counts)
        # not the sine field, which vanishes on every node of a one-element-wide brick
        return W.Workload("nh", mesh, wl.physics, 3, 0.02 * np.cos(1.3 * np.arange(3.0 * mesh.num_nodes)))
    if name == "poisson 3x15x2":
        mesh = _brick(fec, (3, 15, 2))
        return W.Workload("p", mesh, fec.Poisson("constant", 1.0), 1, np.cos(np.arange(mesh.num_nodes, dtype=float)))
    raise KeyError(name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["poisson 17x9x12", "neohookean 16^3", "neohookean 20x8x31", "mechanics 11^3",
                                  "j2 9^3 clamped", "neohookean 2x2x2", "neohookean 9x2x2", "neohookean 2x3x17",
                                  "neohookean 16x16x2", "poisson 3x15x2"])
def test_fused_matches_two_phase_and_oracle(fec, orc, gpu, name):
    print(f"[fused] {name}", flush=True)
    wl = _case(fec, name)
    m = wl.mesh
    V = fec.FunctionSpace(m, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    bcs = []
    if len(wl.dirichlet_dofs):
        m.nodeset_nodes["_bc"] = np.unique((wl.dirichlet_dofs - 1) // wl.nf + 1)
        bcs = [fec.DirichletBC(nm, lambda X, t: 0.0, nset_name="_bc") for nm in var.names]
    p = fec.create_parameters(m, asm, wl.physics, dirichlet_bcs=bcs)
    # the default design: fused for the vector-valued materials, two-phase for the scalar Poisson problem
    assert _applies(fec, asm.ctx) == (1 if wl.nf == 3 else 0)
    asm.ctx.call("fecx_set_scatter", ctypes.c_int(FUSED))
    assert _applies(fec, asm.ctx) == 1
    U = wl.U.copy()
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    unk = asm.dof.unknown_dofs
    Uu = torch.from_numpy(U[unk - 1]).to(gpu)
    st = None
    if wl.physics.kind == orc.J2:  # a nonzero old state
        g = fec.rng.normal(1, 11, m.num_elements * 8 * 7).reshape(m.num_elements, 8, 7)
        st = np.concatenate([2e-3 * g[..., :6], 2e-3 * np.abs(g[..., 6:])], axis=2).ravel().astype(np.float32)
        asm.ctx.call("fecx_set_state_old", st.ctypes.data_as(ctypes.c_void_p))
    out = {}
    designs = [FUSED, TWO_PHASE] + ([ATOMIC, COLOURED] if wl.physics.kind in (orc.POISSON, orc.NEOHOOKEAN) else [])
    names = {FUSED: "fused", TWO_PHASE: "reduce", ATOMIC: "atomic", COLOURED: "coloured"}
    for design in designs:
        asm.ctx.call("fecx_set_scatter", ctypes.c_int(design))
        fec.reset_profile(asm)
        fec.set_profiling(asm, True)
        fec.assemble_vector_and_stiffness_(asm, Uu, p)
        fec.set_profiling(asm, False)
        kernels = set(fec.profile(asm))
        assert names[design] in kernels, (design, kernels)
        R = fec.residual(asm).cpu().numpy().copy()
        K = asm.stiffness_storage.cpu().numpy().copy()
        sn = None
        if st is not None:
            sn = np.zeros(st.size, dtype=np.float32)
            asm.ctx.call("fecx_get_state_new", sn.ctypes.data_as(ctypes.c_void_p))
        out[design] = (R, K, sn)
        if design != ATOMIC:  # the Jacobian-only call agrees with the residual + Jacobian call
            fec.assemble_stiffness_(asm, fec.stiffness, Uu, p)
            assert np.array_equal(asm.stiffness_storage.cpu().numpy(), K)
    Rt, Kt, stt = out[TWO_PHASE]
    for design in designs:
        if design == TWO_PHASE:
            continue
        Rd, Kd, sd = out[design]
        dR = np.abs(Rd - Rt).max() / np.abs(Rt).max()
        dK = np.abs(Kd - Kt).max() / np.abs(Kt).max()
        print(f"[fused] {name}: {names[design]} vs two-phase R {dR:.1e} K {dK:.1e}, bit-equal K {np.mean(Kd == Kt):.4f}",
              flush=True)
        assert dR <= 2e-15 and dK <= 2e-15, (design, dR, dK)
        if sd is not None:
            assert np.array_equal(sd, stt)
    Rf, Kf, sf = out[FUSED]
    # the oracle
    k = {"HEX8": orc.HEX8}[m.element_type]
    props = wl.physics.props()
    sto = None if st is None else st.astype(np.float64)
    R_ref = orc.assemble_residual(k, m.coords, m.conn, wl.nf, wl.physics.kind, props, U, sto,
                                  None if sto is None else np.zeros_like(sto))[unk - 1]
    _, d2u = orc.update_dofs(wl.nf * m.num_nodes, asm.dof.dirichlet_dofs)
    storage = orc.assemble_stiffness_storage(k, m.coords, m.conn, wl.nf, wl.physics.kind, props, U, sto)
    rp, cv, vals = orc.sparse_csr_rows(wl.nf, m.conn, d2u, len(unk), storage)
    eR, eK = _errs(Rf, R_ref), _errs(Kf, vals)  # residual(asm): the unknowns already
    print(f"[fused] {name}: vs oracle R {eR:.2e} K {eK:.2e}", flush=True)
    assert eR < 1e-6 and eK < 1e-6
