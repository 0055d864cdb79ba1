"""Parity against the FP64 oracle at the configurations' full sizes, computed on the host of the GPU box.

  neohookean_hex8 96^3 (884,736 elements, 217,238,121 CSR values): residual, CSR values, matrix-free and assembled
      Jacobian actions, each within 1e-6 relative error of FP64 in the 2-norm and in the max-norm -- the bound the
      FP32 element kernels are held to on this configuration. The oracle's CSR is its element storage reduced row by
      row with sparse!'s sums (bit-identical to its COO sparse! path, pinned in test_oracle_pins.py); the pattern is
      compared bit for bit.
  neohookean_tet10 80^3 x 6 (3,072,000 jittered elements): residual and both actions. Here FP32 element arithmetic
      lands above 1e-6 (measured 2.6e-6 for the residual's 2-norm, 1.2-1.6e-6 in the max-norms): interior residuals
      are cancelling sums of O(1) element contributions, so their relative error grows like eps / h, and on the
      jittered tets it comes from every stage of the element chain (geometry, stress, contraction) alike. The test
      states that error: 4e-6 in the 2-norm, 3e-6 in the max-norm.
  neohookean_tet10 80^3 x 6 with the FP64 element geometry, field gradient and stress (fecx_set_fp64_geometry):
      residual, CSR values and both actions, each within 1e-6 in both norms.
  poisson_hex8 128^3 and j2_hex8 64^3 (last load step on a plastic state): every output within 1e-6 in both norms,
      and the J2 updated state."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _errs(a, b):
    return (float(np.linalg.norm(a - b) / np.linalg.norm(b)), float(np.abs(a - b).max() / np.abs(b).max()))


def _setup(fec, wl, fp64_geometry=False):
    V = fec.FunctionSpace(wl.mesh, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if wl.nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=0, fp64_geometry=fp64_geometry)
    p = fec.create_parameters(wl.mesh, asm, wl.physics)
    return asm, p


def _actions(fec, orc, wl, asm, p, Uu, kind, U=None, st=None):
    m = wl.mesh
    U = wl.U if U is None else U
    N = wl.nf * m.num_nodes
    Vf = np.cos(np.arange(N, dtype=np.float64))
    Vu = torch.from_numpy(Vf).to(Uu.device)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    mf = fec.hvp(asm, Vu).cpu().numpy().copy()
    fec.assemble_matrix_action_(asm, fec.stiffness, Uu, Vu, p)
    am = fec.hvp(asm, Vu).cpu().numpy().copy()
    props = wl.physics.props()
    mf_ref = orc.assemble_action(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, U, Vf, st, matrix_free=True)
    am_ref = orc.assemble_action(kind, m.coords, m.conn, wl.nf, wl.physics.kind, props, U, Vf, st)
    return _errs(mf, mf_ref), _errs(am, am_ref)



def test_neohookean_hex8_96_every_output_within_1e6(fec, orc, gpu):
    print("[fullsize parity] neohookean_hex8 96^3", flush=True)
    wl = fec.workloads.neohookean_hex8(96)
    m = wl.mesh
    asm, p = _setup(fec, wl)
    Uu = torch.from_numpy(wl.U).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = asm.stiffness_storage.cpu().numpy()
    props = wl.physics.props()
    orc.set_threads(16)
    R_ref = orc.assemble_residual(orc.HEX8, m.coords, m.conn, 3, orc.NEOHOOKEAN, props, wl.U)
    eR = _errs(R, R_ref)
    del R_ref
    N = 3 * m.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    st = orc.assemble_stiffness_storage(orc.HEX8, m.coords, m.conn, 3, orc.NEOHOOKEAN, props, wl.U)
    rp, cv, vals = orc.sparse_csr_rows(3, m.conn, d2u, N, st)
    del st
    rp_g, cv_g = (t.cpu().numpy() + 1 for t in asm.pattern())
    assert np.array_equal(rp_g, rp) and np.array_equal(cv_g, cv)
    eK = _errs(K, vals)
    del K, vals, rp, cv, rp_g, cv_g
    eMF, eAM = _actions(fec, orc, wl, asm, p, Uu, orc.HEX8)
    print(f"[fullsize parity] hex8 96^3 (l2/max): R {eR[0]:.2e}/{eR[1]:.2e} K {eK[0]:.2e}/{eK[1]:.2e} "
          f"MF {eMF[0]:.2e}/{eMF[1]:.2e} ASM {eAM[0]:.2e}/{eAM[1]:.2e}", flush=True)
    for e in (eR, eK, eMF, eAM):
        assert e[0] < 1e-6 and e[1] < 1e-6


def test_neohookean_tet10_80_residual_and_actions(fec, orc, gpu):
    print("[fullsize parity] neohookean_tet10 80^3 x 6", flush=True)
    wl = fec.workloads.neohookean_tet10(80)
    m = wl.mesh
    asm, p = _setup(fec, wl)
    Uu = torch.from_numpy(wl.U).to(gpu)
    fec.assemble_vector_(asm, fec.residual, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    orc.set_threads(16)
    R_ref = orc.assemble_residual(orc.TET10, m.coords, m.conn, 3, orc.NEOHOOKEAN, wl.physics.props(), wl.U)
    eR = _errs(R, R_ref)
    eMF, eAM = _actions(fec, orc, wl, asm, p, Uu, orc.TET10)
    print(f"[fullsize parity] tet10 80^3x6 (l2/max): R {eR[0]:.2e}/{eR[1]:.2e} MF {eMF[0]:.2e}/{eMF[1]:.2e} "
          f"ASM {eAM[0]:.2e}/{eAM[1]:.2e}", flush=True)
    for e in (eR, eMF, eAM):
        assert e[0] < 4e-6 and e[1] < 3e-6


def test_neohookean_tet10_80_fp64_geometry_every_output_within_1e6(fec, orc, gpu):
    print("[fullsize parity] neohookean_tet10 80^3 x 6, FP64 element geometry", flush=True)
    wl = fec.workloads.neohookean_tet10(80)
    m = wl.mesh
    asm, p = _setup(fec, wl, fp64_geometry=True)
    Uu = torch.from_numpy(wl.U).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = asm.stiffness_storage.cpu().numpy()
    props = wl.physics.props()
    orc.set_threads(16)
    R_ref = orc.assemble_residual(orc.TET10, m.coords, m.conn, 3, orc.NEOHOOKEAN, props, wl.U)
    eR = _errs(R, R_ref)
    del R_ref
    N = 3 * m.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    st = orc.assemble_stiffness_storage(orc.TET10, m.coords, m.conn, 3, orc.NEOHOOKEAN, props, wl.U)
    rp, cv, vals = orc.sparse_csr_rows(3, m.conn, d2u, N, st)
    del st
    rp_g, cv_g = (t.cpu().numpy() + 1 for t in asm.pattern())
    assert np.array_equal(rp_g, rp) and np.array_equal(cv_g, cv)
    eK = _errs(K, vals)
    del K, vals, rp, cv, rp_g, cv_g
    eMF, eAM = _actions(fec, orc, wl, asm, p, Uu, orc.TET10)
    print(f"[fullsize parity] tet10 80^3x6 fp64 geometry (l2/max): R {eR[0]:.2e}/{eR[1]:.2e} "
          f"K {eK[0]:.2e}/{eK[1]:.2e} MF {eMF[0]:.2e}/{eMF[1]:.2e} ASM {eAM[0]:.2e}/{eAM[1]:.2e}", flush=True)
    for e in (eR, eK, eMF, eAM):
        assert e[0] < 1e-6 and e[1] < 1e-6


def _csr_errs(fec, orc, wl, asm, K, kind, U, st=None):
    m = wl.mesh
    N = wl.nf * m.num_nodes
    _, d2u = orc.update_dofs(N, np.zeros(0, np.int64))
    storage = orc.assemble_stiffness_storage(kind, m.coords, m.conn, wl.nf, wl.physics.kind, wl.physics.props(), U, st)
    rp, cv, vals = orc.sparse_csr_rows(wl.nf, m.conn, d2u, N, storage)
    del storage
    rp_g, cv_g = (t.cpu().numpy() + 1 for t in asm.pattern())
    assert np.array_equal(rp_g, rp) and np.array_equal(cv_g, cv)
    return _errs(K, vals)


def test_poisson_hex8_128_every_output_within_1e6(fec, orc, gpu):
    print("[fullsize parity] poisson_hex8 128^3", flush=True)
    wl = fec.workloads.poisson_hex8(128)
    m = wl.mesh
    asm, p = _setup(fec, wl)
    Uu = torch.from_numpy(wl.U).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = asm.stiffness_storage.cpu().numpy()
    orc.set_threads(16)
    props = wl.physics.props()
    eR = _errs(R, orc.assemble_residual(orc.HEX8, m.coords, m.conn, 1, wl.physics.kind, props, wl.U))
    eK = _csr_errs(fec, orc, wl, asm, K, orc.HEX8, wl.U)
    del K
    eMF, eAM = _actions(fec, orc, wl, asm, p, Uu, orc.HEX8)
    print(f"[fullsize parity] poisson hex8 128^3 (l2/max): R {eR[0]:.2e}/{eR[1]:.2e} K {eK[0]:.2e}/{eK[1]:.2e} "
          f"MF {eMF[0]:.2e}/{eMF[1]:.2e} ASM {eAM[0]:.2e}/{eAM[1]:.2e}", flush=True)
    for e in (eR, eK, eMF, eAM):
        assert e[0] < 1e-6 and e[1] < 1e-6


def test_j2_hex8_64_plastic_every_output_within_1e6(fec, orc, gpu):
    """J2 at 64^3 with the last load step's displacement on a seeded plastic state (every point with an eqps > 0
    and a back-stress-free plastic strain of ~1e-3), so the return mapping takes its plastic branch: residual, updated
    state, CSR values and both actions against the oracle."""
    import ctypes
    print("[fullsize parity] j2_hex8 64^3, plastic state", flush=True)
    W = fec.workloads
    wl = W.j2_hex8(64)
    m = wl.mesh
    path = wl.extra["path"]
    U = W.j2_displacement(m.coords, path[-1], 1.0, wl.extra["seed"])
    g = fec.rng.normal(7, 13, m.num_elements * 8 * 7).reshape(m.num_elements, 8, 7)
    st = np.concatenate([1e-3 * g[..., :6], 1e-3 * np.abs(g[..., 6:])], axis=2).ravel().astype(np.float32)
    asm, p = _setup(fec, wl)
    asm.ctx.call("fecx_set_state_old", st.ctypes.data_as(ctypes.c_void_p))
    st64 = st.astype(np.float64)
    Uu = torch.from_numpy(U).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy().copy()
    K = asm.stiffness_storage.cpu().numpy()
    sn = np.zeros(st.size, dtype=np.float32)
    asm.ctx.call("fecx_get_state_new", sn.ctypes.data_as(ctypes.c_void_p))
    orc.set_threads(16)
    props = wl.physics.props()
    sn_ref = np.zeros_like(st64)
    R_ref = orc.assemble_residual(orc.HEX8, m.coords, m.conn, 3, orc.J2, props, U, st64, sn_ref)
    eR = _errs(R, R_ref)
    plastic = float((sn_ref[6::7] > st64[6::7]).mean())
    eS = float(np.abs(sn - sn_ref).max() / max(1.0, np.abs(sn_ref).max()))  # the states are O(1e-3): absolute
    eK = _csr_errs(fec, orc, wl, asm, K, orc.HEX8, U, st64)
    del K
    eMF, eAM = _actions(fec, orc, wl, asm, p, Uu, orc.HEX8, U=U, st=st64)
    print(f"[fullsize parity] j2 hex8 64^3 (l2/max; {100 * plastic:.0f} % of the points yield): R {eR[0]:.2e}/{eR[1]:.2e} "
          f"state {eS:.2e} K {eK[0]:.2e}/{eK[1]:.2e} MF {eMF[0]:.2e}/{eMF[1]:.2e} ASM {eAM[0]:.2e}/{eAM[1]:.2e}",
          flush=True)
    assert plastic > 0.5
    assert eS < 1e-6
    for e in (eR, eK, eMF, eAM):
        assert e[0] < 1e-6 and e[1] < 1e-6
