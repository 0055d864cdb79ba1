"""Exodus II meshes (the reference's FileMesh / UnstructuredMesh, src/meshes/Exodus.jl:18-141) and results files
(PostProcessor, src/meshes/Exodus.jl:148-290), on the reference's own mesh files tests/golden/cube.g and
tests/golden/mechanics_coarse.g.

CPU: what the files fix (counts, node and side sets on their planes, id maps), the product's host pattern builder on
the file meshes against the oracle (bit-exact), and a write -> read round trip. GPU: assembly on the file meshes
against the oracle, to the tolerances of test_gpu_parity.py."""
import os

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CUBE = os.path.join(ROOT, "tests", "golden", "cube.g")
QUADS = os.path.join(ROOT, "tests", "golden", "mechanics_coarse.g")


def test_cube_mesh(fec):
    m = fec.exodus.UnstructuredMesh(CUBE)
    assert (m.element_type, m.num_dims, m.num_nodes, m.num_elements) == ("HEX8", 3, 27, 8)
    assert m.conn.min() == 1 and m.conn.max() == 27 and len(np.unique(m.conn)) == 27
    assert np.allclose(m.coords.min(0), -0.5) and np.allclose(m.coords.max(0), 0.5)
    planes = {"x": 0, "y": 1, "z": 2}
    for name, nodes in m.nodeset_nodes.items():     # nsx-, nsx+, ... : the 9 nodes of each face, sorted
        ax, side = planes[name[2]], -0.5 if name[3] == "-" else 0.5
        assert len(nodes) == 9 and np.all(np.diff(nodes) > 0)
        assert np.allclose(m.coords[nodes - 1, ax], side)
    assert set(m.nodeset_nodes) == {"nsx-", "nsx+", "nsy-", "nsy+", "nsz-", "nsz+"}
    exo = m.exodus
    for name in ("ssx-", "ssx+", "ssy-", "ssz-", "ssz+"):
        # side nodes through the Exodus HEX8 face table equal the node set of the same face
        assert np.array_equal(exo.sideset_nodes[name], m.nodeset_nodes["ns" + name[2:]])
        assert len(exo.sideset_elems[name]) == 4 and np.all(np.diff(exo.sideset_elems[name]) > 0)
    # the file's ssy+ lists the same (element, side) pairs as ssy-
    assert np.array_equal(exo.sideset_elems["ssy+"], exo.sideset_elems["ssy-"])
    # global node ids: a permutation of 1..27, kept aside from the positional numbering the connectivity uses
    assert np.array_equal(np.sort(exo.node_id_map), np.arange(1, 28)) and exo.node_id_map[0] != 1


def test_quad_mesh(fec):
    m = fec.exodus.UnstructuredMesh(QUADS)
    assert (m.element_type, m.num_dims, m.num_nodes, m.num_elements) == ("QUAD4", 2, 121, 100)
    assert np.allclose(m.coords.min(0), 0.0) and np.allclose(m.coords.max(0), 1.0)
    edges = []
    for name, nodes in m.nodeset_nodes.items():
        assert len(nodes) == 11
        on = [ax for ax in (0, 1) for v in (0.0, 1.0) if np.allclose(m.coords[nodes - 1, ax], v)]
        assert len(on) == 1
        edges.append(on[0])
        # the side set of the same number covers the same edge
        assert np.array_equal(m.exodus.sideset_nodes["sset_" + name.split("_")[1]], nodes)
    assert sorted(edges) == [0, 0, 1, 1]
    # counter-clockwise elements: positive signed area
    x = m.coords[m.conn - 1]
    area = 0.5 * np.sum(x[:, :, 0] * np.roll(x[:, :, 1], -1, 1) - np.roll(x[:, :, 0], -1, 1) * x[:, :, 1], axis=1)
    assert np.all(area > 0) and np.isclose(area.sum(), 1.0)
    # the global ids are not the positional numbering
    assert m.exodus.node_id_map[0] != 1 and len(np.unique(m.exodus.node_id_map)) == 121


@pytest.mark.parametrize("path,nf", [(CUBE, 1), (CUBE, 3), (QUADS, 1)])
def test_pattern_on_file_mesh(fec, orc, path, nf):
    m = fec.exodus.UnstructuredMesh(path)
    nodes = next(iter(m.nodeset_nodes.values()))
    dd = np.sort(np.concatenate([nf * (nodes - 1) + d + 1 for d in range(nf)])).astype(np.int64)
    hp = fec.HostPattern(m, nf, dd)
    unknown, d2u = orc.update_dofs(nf * m.num_nodes, dd)
    assert np.array_equal(hp.unknown_dofs, unknown)
    Is, Js, _, _ = orc.pattern(nf, m.conn, d2u)
    rowptr, colval, _, _ = orc.sparse_csr(len(unknown), Is, Js, want_map=True)
    rp, cv = hp.pattern(base=1)
    assert np.array_equal(rp, rowptr) and np.array_equal(cv, colval)


def test_write_read_round_trip(fec, tmp_path):
    m = fec.exodus.UnstructuredMesh(CUBE)
    u = np.stack([np.sin(t + m.coords[:, 0]) for t in (0.0, 0.5, 1.0)])
    out = str(tmp_path / "out.e")
    fec.exodus.write_exodus(out, m, {"u": u, "v": 2 * u}, times=[0.0, 0.5, 1.0])
    r = fec.exodus.ExodusMesh(out)
    assert np.array_equal(r.coords, m.coords)
    (etype, conn, _), = r.blocks.values()
    assert etype == "HEX8" and np.array_equal(conn, m.conn)
    from scipy.io import netcdf_file
    with netcdf_file(out, "r", mmap=False) as f:
        assert np.array_equal(f.variables["time_whole"][:], [0.0, 0.5, 1.0])
        names = [b"".join(r).strip(b"\x00").decode() for r in f.variables["name_nod_var"][:]]
        assert names == ["u", "v"]
        assert np.array_equal(f.variables["vals_nod_var1"][:], u)
        assert np.array_equal(f.variables["vals_nod_var2"][:], 2 * u)


@pytest.mark.gpu
@pytest.mark.parametrize("path,physics", [(CUBE, "poisson"), (CUBE, "neohookean"), (QUADS, "poisson")])
def test_assembly_on_file_mesh(fec, orc, gpu, path, physics):
    print(f"[gpu exodus] {os.path.basename(path)} {physics}", flush=True)
    m = fec.exodus.UnstructuredMesh(path)
    nf = 1 if physics == "poisson" else 3
    V = fec.FunctionSpace(m, "H1Field", "Lagrange")
    var = fec.ScalarFunction(V, "u") if nf == 1 else fec.VectorFunction(V, "displ")
    asm = fec.SparseMatrixAssembler(var, sparse_matrix_type="csr", device=gpu.index or 0)
    nset = next(iter(m.nodeset_nodes))
    bcs = [fec.DirichletBC(name, lambda X, t: 0.0, nset_name=nset) for name in var.names]
    phys = fec.Poisson("sines", 1.0) if nf == 1 else fec.NeoHookean(1.0, 0.5)
    p = fec.create_parameters(m, asm, phys, dirichlet_bcs=bcs)
    unknown = asm.dof.unknown_dofs
    N = nf * m.num_nodes
    _, d2u = orc.update_dofs(N, asm.dof.dirichlet_dofs)
    U = 0.05 * np.sin(3.0 * np.arange(N))
    U[asm.dof.dirichlet_dofs - 1] = 0.0
    Vv = np.zeros(N)
    Vv[unknown - 1] = np.cos(np.arange(len(unknown)))
    k = {"QUAD4": orc.QUAD4, "HEX8": orc.HEX8}[m.element_type]
    props = phys.props()
    R_ref = orc.assemble_residual(k, m.coords, m.conn, nf, phys.kind, props, U, None, None)[unknown - 1]
    rp, cv, vals = orc.stiffness_csr(k, m.coords, m.conn, nf, phys.kind, props, U, d2u, len(unknown), None)
    KV_ref = orc.assemble_action(k, m.coords, m.conn, nf, phys.kind, props, U, Vv, None, matrix_free=True)[unknown - 1]

    Uu = torch.from_numpy(U[unknown - 1]).to(gpu)
    Vu = torch.from_numpy(Vv[unknown - 1]).to(gpu)
    fec.assemble_vector_and_stiffness_(asm, Uu, p)
    R = fec.residual(asm).cpu().numpy()
    K = fec.stiffness(asm)
    fec.assemble_matrix_free_action_(asm, fec.stiffness_action, Uu, Vu, p)
    KV = fec.hvp(asm, Vu).cpu().numpy()

    def rel(a, b):  # the larger of the relative 2-norm and max-norm errors
        return max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)),
                   float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))

    assert np.array_equal(K.rowptr.cpu().numpy() + 1, rp) and np.array_equal(K.colval.cpu().numpy() + 1, cv)
    errs = (rel(R, R_ref), rel(K.values.cpu().numpy(), vals), rel(KV, KV_ref))
    print(f"[gpu exodus] R={errs[0]:.2e} K={errs[1]:.2e} KV={errs[2]:.2e}", flush=True)
    assert max(errs) < 1e-6
