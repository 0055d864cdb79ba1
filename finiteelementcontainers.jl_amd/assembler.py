"""Host-side mirror of FiniteElementContainers.jl's assembler interface, backed by libfecx.so on one GPU.

Names follow the reference with Julia's ``!`` spelled as a trailing underscore:

    reference (Julia)                                   here (Python)
    SparseMatrixAssembler(u; sparse_matrix_type=:csr)   SparseMatrixAssembler(u, sparse_matrix_type="csr")
    create_parameters(mesh, asm, physics, props; ...)   create_parameters(mesh, asm, physics, props, ...)
    assemble_vector!(asm, residual, Uu, p)              assemble_vector_(asm, residual, Uu, p)
    assemble_stiffness!(asm, stiffness, Uu, p)          assemble_stiffness_(asm, stiffness, Uu, p)
    assemble_matrix_action!(asm, stiffness, Uu, Vu, p)  assemble_matrix_action_(asm, stiffness, Uu, Vu, p)
    assemble_matrix_free_action!(asm, stiffness_action, Uu, Vu, p)
    residual(asm), stiffness(asm), hvp(asm, v)          residual(asm), stiffness(asm), hvp(asm, v)

Call sequences, output shapes (reduced vectors of the unknowns, CSR of n_unknowns x n_unknowns) and the
Dirichlet write-then-assemble order (src/Parameters.jl:404-425) are the reference's. Fields live in HBM as torch
float64 tensors; the element kernels and reductions run in libfecx.so. There is no CPU compute path.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import FECXError

# ------------------------------------------------------------------ physics descriptors


class Physics:
    kind = None
    nf = None

    def props(self):
        raise NotImplementedError


class Poisson(Physics):
    """-div grad u = f (test/poisson/TestPoissonCommon.jl:75-127).

    source="constant": f = value; source="sines": f = value * nd pi^2 prod_d sin(pi x_d)
    (2 pi^2 sin(pi x) sin(pi y) in 2D, the Poisson test's forcing)."""
    kind = _lib.POISSON
    nf = 1

    def __init__(self, source="constant", value=1.0):
        self.source = source
        self.value = float(value)

    def props(self):
        return [0.0 if self.source == "constant" else 1.0, self.value]


class ThreeDimensional:
    """Formulation tag (src/Formulations.jl:456-611)."""
    num_dims = 3


class PlaneStrain:
    """Formulation tag (src/Formulations.jl:318-427): 2D displacement gradients embedded in 3D with a zero third row
    and column; stress and tangent restricted to the in-plane components. Mechanics only (the test linear-elastic
    model on quad4 meshes with two displacement fields)."""
    num_dims = 2


class Mechanics(Physics):
    """The reference's test Mechanics physics: psi = K/2 tr(eps)^2 + G dev eps : dev eps
    (test/mechanics/TestMechanicsCommon.jl:1-236); props = (rho, K, G). Mechanics(PlaneStrain()) is its 2D form
    (test/mechanics/TestMechanics.jl:24)."""
    kind = _lib.LINEAR_ELASTIC

    def __init__(self, formulation=None, rho=1e3, K=10.0e9, G=1.0e9):
        self.formulation = formulation or ThreeDimensional()
        self.nf = self.formulation.num_dims
        self.p = [rho, K, G]

    def props(self):
        return list(self.p)


class NeoHookean(Physics):
    """Compressible neo-Hookean, psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2."""
    kind = _lib.NEOHOOKEAN
    nf = 3

    def __init__(self, lam=1.0, mu=0.5, formulation=None):
        self.formulation = formulation or ThreeDimensional()
        self.p = [lam, mu]

    def props(self):
        return list(self.p)


class J2Plasticity(Physics):
    """Small-strain J2 plasticity with linear isotropic hardening; 7 state variables per quadrature point."""
    kind = _lib.J2
    nf = 3

    def __init__(self, E=200.0, nu=0.3, sigma_y=0.25, H=2.0, formulation=None):
        self.formulation = formulation or ThreeDimensional()
        self.p = [E, nu, sigma_y, H]

    def props(self):
        return list(self.p)


# callback tags: in the reference these are the physics callbacks passed to the assemble functions and,
# with an assembler argument, the accessors of the assembled results (Julia multiple dispatch).
def residual(obj, *args):
    if isinstance(obj, SparseMatrixAssembler):
        return obj._residual()
    raise TypeError("residual(physics, ...) is evaluated inside libfecx.so; pass `residual` as the callback")


def stiffness(obj, *args):
    if isinstance(obj, SparseMatrixAssembler):
        return obj._stiffness()
    raise TypeError("stiffness(physics, ...) is evaluated inside libfecx.so; pass `stiffness` as the callback")


def stiffness_action(obj, *args):
    raise TypeError("stiffness_action(physics, ...) is evaluated inside libfecx.so; pass it as the callback")


def mass(obj, *args):
    """Callback tag of the consistent mass; mass(asm) is the assembled CSR mass matrix (Assemblers.jl:329-336)."""
    if isinstance(obj, SparseMatrixAssembler):
        return obj._mass()
    raise TypeError("mass(physics, ...) is evaluated inside libfecx.so; pass `mass` as the callback")


def lumped_mass(obj, *args):
    """Callback tag of the row-sum lumped mass; lumped_mass(asm) is the assembled vector over the unknowns
    (src/assemblers/LumpedMass.jl:69-80). It shares storage with residual(asm)."""
    if isinstance(obj, SparseMatrixAssembler):
        return obj._full_or_unknowns()
    raise TypeError("lumped_mass(physics, ...) is evaluated inside libfecx.so; pass `lumped_mass` as the callback")


def diagonal(asm):
    """The assembled diagonal over the unknowns, or the full vector in condensed mode (src/assemblers/Diagonal.jl:70-83);
    shares storage with residual(asm)."""
    return asm._full_or_unknowns()


def density(physics):
    """The density the test physics' mass callbacks multiply by: props[1] of Mechanics
    (test/mechanics/TestMechanicsCommon.jl:52-124), 1 for Poisson (test/poisson/TestPoissonCommon.jl:18-41), and the
    `density` attribute (default 1) of the other materials. With per-block physics every block must have the same
    density (the mass assemblies take one)."""
    if isinstance(physics, dict):
        ds = {density(ph) for ph in physics.values()}
        if len(ds) != 1:
            raise FECXError(-4, "mass assemblies take one density; the blocks' physics differ in density")
        return ds.pop()
    if isinstance(physics, Mechanics):
        return float(physics.props()[0])
    return float(getattr(physics, "density", 1.0))


# ------------------------------------------------------------------ function spaces and DOFs
_ELEM_KIND = {"QUAD4": _lib.QUAD4, "HEX8": _lib.HEX8, "TET10": _lib.TET10}


class FunctionSpace:
    def __init__(self, mesh, field_type="H1Field", interpolation="Lagrange"):
        self.mesh = mesh
        self.elem_kind = _ELEM_KIND[mesh.element_type]


class ScalarFunction:
    def __init__(self, fspace, name):
        self.fspace, self.name, self.nf = fspace, name, 1
        self.names = [name]


class VectorFunction:
    def __init__(self, fspace, name):
        self.fspace, self.name, self.nf = fspace, name, fspace.mesh.num_dims
        self.names = [f"{name}_{c}" for c in "xyz"[: self.nf]]


class DofManager:
    """NF x NN DOF layout; unknown_dofs after update_dofs! (src/DofManagers.jl:21-73, 227-298)."""

    def __init__(self, var):
        self.var = var
        self.nf = var.nf
        self.nn = var.fspace.mesh.num_nodes
        self.dirichlet_dofs = np.zeros(0, dtype=np.int64)
        self.unknown_dofs = np.arange(1, self.nf * self.nn + 1, dtype=np.int64)

    def __len__(self):
        return self.nf * self.nn

    def size(self):
        return (self.nf, self.nn)


@dataclass
class DirichletBC:
    """DirichletBC(var_name, func; nset_name=...|sset_name=...|block_name=...) (src/bcs/DirichletBCs.jl:1-48).

    func(X, t) -> value, vectorised over nodes (X is NN x nd)."""
    var_name: str
    func: object
    nset_name: str = None
    sset_name: str = None
    block_name: str = None


class DirichletBCs:
    def __init__(self, mesh, dof, bcs):
        self.entries = []
        names = dof.var.names
        for bc in bcs:
            comp = names.index(bc.var_name)
            if bc.nset_name is not None:
                nodes = np.asarray(mesh.nodeset_nodes[bc.nset_name], dtype=np.int64)
            elif bc.sset_name is not None:
                nodes = np.asarray(mesh.sideset_nodes[bc.sset_name], dtype=np.int64)
            else:
                nodes = np.unique(mesh.conn)
            dofs = dof.nf * (nodes - 1) + comp + 1
            _, first = np.unique(dofs, return_index=True)  # _unique_sort_perm (src/bcs/BoundaryConditions.jl:52-56)
            self.entries.append((dofs[first], nodes[first], bc.func))
        all_dofs = np.concatenate([e[0] for e in self.entries]) if self.entries else np.zeros(0, dtype=np.int64)
        self.dirichlet_dofs = np.unique(all_dofs)  # unique(sort(...)) (DirichletBCs.jl:396-398)

    def values(self, coords, t):
        """Values at dirichlet_dofs (later BCs overwrite earlier ones on shared dofs)."""
        vals = np.zeros(len(self.dirichlet_dofs))
        for dofs, nodes, func in self.entries:
            v = np.broadcast_to(np.asarray(func(coords[nodes - 1], t), dtype=np.float64), dofs.shape)
            vals[np.searchsorted(self.dirichlet_dofs, dofs)] = v
        return vals


# ------------------------------------------------------------------ assembler
class CSRMatrix:
    """CSR matrix on the GPU: rowptr/colval are 0-based int64 tensors, values float64 (SparseMatrixCSR
    uses 1-based indices; ``rowptr + 1`` / ``colval + 1`` give that convention)."""

    def __init__(self, n, rowptr, colval, values):
        self.shape = (n, n)
        self.rowptr, self.colval, self.values = rowptr, colval, values

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values.cpu().numpy(), self.colval.cpu().numpy(), self.rowptr.cpu().numpy()),
                             shape=self.shape)

    def to_torch(self):
        return torch.sparse_csr_tensor(self.rowptr, self.colval, self.values, size=self.shape)


class CSCMatrix:
    """CSC matrix on the GPU (sparse_matrix_type = :csc, the reference's default): colptr/rowval 0-based int64
    tensors, nzval float64 (SparseMatrixCSC uses 1-based indices). The element pattern is structurally symmetric and
    the assembled tangents and masses are symmetric bit for bit, so colptr/rowval/nzval are the CSR rowptr/colval/values
    of the same matrix, shared without a copy."""

    def __init__(self, n, colptr, rowval, nzval):
        self.shape = (n, n)
        self.colptr, self.rowval, self.nzval = colptr, rowval, nzval

    @property
    def values(self):
        return self.nzval

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval.cpu().numpy(), self.rowval.cpu().numpy(), self.colptr.cpu().numpy()),
                             shape=self.shape)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _set_mesh(ctx, mesh, elem_kind):
    """fecx_set_mesh_blocks with the mesh's element blocks (one block for meshes without ``block_conns``)."""
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int64)
    sizes = np.asarray(getattr(mesh, "block_sizes", [conn.shape[0]]), dtype=np.int64)
    ctx.call("fecx_set_mesh_blocks", ctypes.c_int(elem_kind), ctypes.c_int(coords.shape[1]),
             ctypes.c_int64(coords.shape[0]), _host_ptr(coords), ctypes.c_int(len(sizes)), _host_ptr(sizes),
             _host_ptr(conn))


def _host_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class SparseMatrixAssembler:
    """SparseMatrixAssembler (src/assemblers/SparseMatrixAssembler.jl:7-133) with one libfecx context.

    sparse_matrix_type "csr" or "csc" (the reference's default, SparseMatrixAssembler.jl:66); use_condensed keeps every
    DOF in the pattern and adjusts the results for the constraints in the accessors (Utils.jl:53-167); fp64_geometry
    evaluates the tet10 element geometry, field gradient and stress in FP64 (fecx_set_fp64_geometry)."""

    def __init__(self, var, sparse_matrix_type="csr", device=0, matrix_free=False, use_condensed=False,
                 penalty_scale=1.0e6, fp64_geometry=False):
        if sparse_matrix_type not in ("csr", "csc"):
            raise FECXError(-4, f"Unsupported sparse matrix type {sparse_matrix_type}. :csr and :csc are assembled "
                                "on the GPU.")
        if not torch.cuda.is_available():
            raise FECXError(-3, "no GPU: libfecx.so assembles on a HIP device only")
        self.device = torch.device("cuda", device)
        self.sparse_matrix_type = sparse_matrix_type
        self.matrix_free = matrix_free
        self.use_condensed = bool(use_condensed)
        self.penalty_scale = float(penalty_scale)
        self.var = var
        self.dof = DofManager(var)
        mesh = var.fspace.mesh
        self.ctx = _lib.Context(device)
        self.ctx.call("fecx_set_stream", ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        _set_mesh(self.ctx, mesh, var.fspace.elem_kind)
        if self.use_condensed:
            self.ctx.call("fecx_set_condensed", ctypes.c_int(1))
        if fp64_geometry:
            self.ctx.call("fecx_set_fp64_geometry", ctypes.c_int(1))
        self._pattern = None
        self._k_adjusted = self._m_adjusted = True
        self.update_dofs(np.zeros(0, dtype=np.int64))

    @property
    def n_unknowns(self):
        """Length of the unknown vectors: the free dofs, or every dof in condensed mode (create_unknowns,
        src/DofManagers.jl:168-175)."""
        return len(self.dof) if self.use_condensed else len(self.dof.unknown_dofs)

    # update_dofs!(asm, dbcs) (src/assemblers/SparseMatrixAssembler.jl:228-274)
    def update_dofs(self, dirichlet_dofs):
        d = np.ascontiguousarray(dirichlet_dofs, dtype=np.int64)
        self.ctx.call("fecx_set_dofs", ctypes.c_int(self.dof.nf), _host_ptr(d) if len(d) else None,
                      ctypes.c_int64(len(d)))
        n = ctypes.c_int64(0)
        self.ctx.call("fecx_get_unknown_dofs", ctypes.byref(n), None)
        unk = np.zeros(n.value, dtype=np.int64)
        self.ctx.call("fecx_get_unknown_dofs", ctypes.byref(n), _host_ptr(unk) if len(unk) else None)
        self.dof.dirichlet_dofs, self.dof.unknown_dofs = d, unk
        nnz = ctypes.c_int64(0)
        self.ctx.call("fecx_build_pattern", ctypes.byref(nnz))
        self.nnz = nnz.value
        self._pattern = None
        self._k_adjusted = self._m_adjusted = True
        dev, N, nu = self.device, len(self.dof), len(unk)
        f64 = torch.float64
        self.residual_storage = torch.zeros(N, dtype=f64, device=dev)
        self.residual_unknowns = torch.zeros(nu, dtype=f64, device=dev)
        self.stiffness_action_storage = torch.zeros(N, dtype=f64, device=dev)
        self.stiffness_action_unknowns = torch.zeros(nu, dtype=f64, device=dev)
        self.stiffness_storage = torch.zeros(0 if self.matrix_free else self.nnz, dtype=f64, device=dev)
        self.mass_storage = torch.zeros(0 if self.matrix_free else self.nnz, dtype=f64, device=dev)

    def pattern(self):
        """(rowptr, colval) 0-based int64 tensors on the GPU, bit-exact with sparse! + SparseMatrixCSR (and the
        (colptr, rowval) of sparse!'s CSC: the pattern is structurally symmetric)."""
        if self._pattern is None:
            nu = self.n_unknowns
            rp = np.zeros(nu + 1, dtype=np.int64)
            cv = np.zeros(self.nnz, dtype=np.int64)
            self.ctx.call("fecx_get_pattern", _host_ptr(rp), _host_ptr(cv), ctypes.c_int(0))
            self._pattern = (torch.from_numpy(rp).to(self.device), torch.from_numpy(cv).to(self.device))
        return self._pattern

    def coo_map(self):
        ne, nnpe = self.var.fspace.mesh.conn.shape
        ndof = nnpe * self.dof.nf
        m = np.zeros(ne * ndof * ndof, dtype=np.int64)
        self.ctx.call("fecx_get_coo_map", _host_ptr(m))
        return m

    def _set_material(self, p):
        """The physics of every block: one Physics for all blocks, or {block name: Physics} (the reference's
        per-block physics NamedTuple, src/Parameters.jl:37-73, handed out by foreach_block, src/Utils.jl:171-189)."""
        if isinstance(p.physics, dict):
            names = list(getattr(p.mesh, "block_names", ["block_1"]))
            if sorted(p.physics) != sorted(names):
                raise FECXError(-1, f"physics given for blocks {sorted(p.physics)}, the mesh has {names}")
            for b, name in enumerate(names):
                ph = p.physics[name]
                props = np.ascontiguousarray(ph.props(), dtype=np.float64)
                self.ctx.call("fecx_set_block_material", ctypes.c_int(b), ctypes.c_int(ph.kind), _host_ptr(props),
                              ctypes.c_int(len(props)))
            return
        props = np.ascontiguousarray(p.physics.props(), dtype=np.float64)
        self.ctx.call("fecx_set_material", ctypes.c_int(p.physics.kind), _host_ptr(props), ctypes.c_int(len(props)))

    def _residual(self):
        """residual(asm) (src/assemblers/Assemblers.jl:347-371): the free-dof subset, or in condensed mode the full
        vector with its constrained entries zeroed."""
        if self.use_condensed:
            self.ctx.call("fecx_adjust_vector_for_constraints", _ptr(self.residual_storage))
            return self.residual_storage
        self.ctx.call("fecx_extract_field_unknowns", _ptr(self.residual_storage), _ptr(self.residual_unknowns))
        return self.residual_unknowns

    def _full_or_unknowns(self):
        """The assembled diagonal / lumped mass: full vector in condensed mode (Diagonal.jl:70-83, LumpedMass.jl:69-80)."""
        if self.use_condensed:
            return self.residual_storage
        self.ctx.call("fecx_extract_field_unknowns", _ptr(self.residual_storage), _ptr(self.residual_unknowns))
        return self.residual_unknowns

    def _matrix(self, values):
        rp, cv = self.pattern()
        n = self.n_unknowns
        if self.sparse_matrix_type == "csc":
            return CSCMatrix(n, rp, cv, values)
        return CSRMatrix(n, rp, cv, values)

    def _stiffness(self):
        """stiffness(asm) (Assemblers.jl:376-388); condensed mode adjusts it for the constraints once per assembly."""
        if self.matrix_free:
            raise FECXError(-2, "stiffness is not available on a matrix-free assembler")
        if self.use_condensed and not self._k_adjusted:
            self.ctx.call("fecx_adjust_matrix_for_constraints", _ptr(self.stiffness_storage),
                          ctypes.c_double(self.penalty_scale))
            self._k_adjusted = True
        return self._matrix(self.stiffness_storage)

    def _mass(self):
        """mass(asm) (Assemblers.jl:329-336), adjusted like the stiffness in condensed mode."""
        if self.matrix_free:
            raise FECXError(-2, "mass is not available on a matrix-free assembler")
        if self.use_condensed and not self._m_adjusted:
            self.ctx.call("fecx_adjust_matrix_for_constraints", _ptr(self.mass_storage),
                          ctypes.c_double(self.penalty_scale))
            self._m_adjusted = True
        return self._matrix(self.mass_storage)

    def hvp(self, v):
        """hvp(asm, v) (Assemblers.jl:310-324): the free-dof subset of the action, or in condensed mode the full
        action with Av = (1 - c) Av + c v."""
        if self.use_condensed:
            _check_vec(self, v, "v")
            self.ctx.call("fecx_adjust_action_for_constraints", _ptr(self.stiffness_action_storage), _ptr(v))
            return self.stiffness_action_storage
        self.ctx.call("fecx_extract_field_unknowns", _ptr(self.stiffness_action_storage),
                      _ptr(self.stiffness_action_unknowns))
        return self.stiffness_action_unknowns


def create_unknowns(asm):
    """create_unknowns (src/DofManagers.jl:168-175): the free dofs, or every dof in condensed mode."""
    return torch.zeros(asm.n_unknowns, dtype=torch.float64, device=asm.device)


def create_field(asm):
    return torch.zeros(len(asm.dof), dtype=torch.float64, device=asm.device)


class Parameters:
    """Assembly state of the reference's Parameters (src/Parameters.jl:37-73): full fields in HBM,
    Dirichlet BCs, physics and properties, time."""

    def __init__(self, mesh, asm, physics, dirichlet_bcs, times):
        self.mesh, self.physics = mesh, physics
        self.coords = np.ascontiguousarray(mesh.coords)
        self.field = create_field(asm)
        self.field_old = create_field(asm)
        self.hvp_scratch_field = create_field(asm)
        self.dirichlet_bcs = DirichletBCs(mesh, asm.dof, dirichlet_bcs)
        self.t, self.dt = times
        self._bc_vals = torch.zeros(len(self.dirichlet_bcs.dirichlet_dofs), dtype=torch.float64, device=asm.device)
        self.update_bc_values()

    def update_bc_values(self):
        vals = self.dirichlet_bcs.values(self.coords, self.t)
        self._bc_vals.copy_(torch.from_numpy(vals))


def create_parameters(mesh, asm, physics, props=None, dirichlet_bcs=(), times=(0.0, 0.0)):
    """create_parameters (src/Parameters.jl:288-302): builds BCs, updates the DOFs and the pattern."""
    p = Parameters(mesh, asm, physics, list(dirichlet_bcs), times)
    if len(p.dirichlet_bcs.dirichlet_dofs):
        asm.update_dofs(p.dirichlet_bcs.dirichlet_dofs)
    asm._set_material(p)
    p._asm = asm
    return p


def _update_for_assembly(p, asm, Uu, Vu=None):
    """_update_for_assembly! (src/Parameters.jl:404-425): BC values, then unknowns, into the full fields."""
    if len(asm.dof.dirichlet_dofs):
        asm.ctx.call("fecx_update_field_dirichlet", _ptr(p._bc_vals), _ptr(p.field))
    asm.ctx.call("fecx_update_field_unknowns", _ptr(Uu), _ptr(p.field))
    if Vu is not None:
        asm.ctx.call("fecx_update_field_unknowns", _ptr(Vu), _ptr(p.hvp_scratch_field))


def _check_vec(asm, x, name, n=None):
    """The reference's @assert on vector sizes and backends (src/DofManagers.jl:399-408)."""
    n = asm.n_unknowns if n is None else n
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or x.device != asm.device or not x.is_contiguous():
        raise FECXError(-1, f"{name} must be a contiguous float64 tensor on {asm.device}")
    if x.numel() != n:
        raise FECXError(-1, f"{name} has {x.numel()} entries, expected {n}")


def assemble_vector_(asm, func, Uu, p):
    """assemble_vector! (src/assemblers/Vector.jl:4-74)."""
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    asm.ctx.call("fecx_assemble_vector", _ptr(p.field), _ptr(asm.residual_storage))


def assemble_stiffness_(asm, func, Uu, p):
    """assemble_stiffness! -> assemble_matrix! (src/assemblers/Matrix.jl:12-75)."""
    if asm.matrix_free:
        raise FECXError(-2, "assemble_stiffness! is not supported on a matrix-free assembler")
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    asm.ctx.call("fecx_assemble_matrix", _ptr(p.field), _ptr(asm.stiffness_storage))
    asm._k_adjusted = False


assemble_matrix_ = assemble_stiffness_


def assemble_vector_and_stiffness_(asm, Uu, p):
    """One pass producing the residual and the Jacobian (the pair of calls of IterativeLinearSolver.solve!)."""
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    asm.ctx.call("fecx_assemble_vector_matrix", _ptr(p.field), _ptr(asm.residual_storage),
                 _ptr(asm.stiffness_storage))
    asm._k_adjusted = False


def assemble_matrix_action_(asm, func, Uu, Vu, p):
    """assemble_matrix_action! (src/assemblers/MatrixAction.jl:154-238); func = stiffness or mass."""
    _check_vec(asm, Uu, "Uu")
    _check_vec(asm, Vu, "Vu")
    _update_for_assembly(p, asm, Uu, Vu)
    if func is mass:
        asm.ctx.call("fecx_assemble_mass_action", ctypes.c_double(density(p.physics)), _ptr(p.hvp_scratch_field),
                     _ptr(asm.stiffness_action_storage))
        return
    asm.ctx.call("fecx_assemble_matrix_action", _ptr(p.field), _ptr(p.hvp_scratch_field),
                 _ptr(asm.stiffness_action_storage))


def assemble_mass_(asm, func, Uu, p):
    """assemble_mass! (src/assemblers/Matrix.jl:1-10): the consistent mass into mass_storage (CSR values)."""
    if asm.matrix_free:
        raise FECXError(-2, "assemble_mass! called on a matrix-free SparseMatrixAssembler. Re-create the assembler "
                            "with matrix_free=False to enable matrix assembly.")
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    asm.ctx.call("fecx_assemble_mass", ctypes.c_double(density(p.physics)), _ptr(asm.mass_storage))
    asm._m_adjusted = False


def assemble_lumped_mass_(asm, func, Uu, p):
    """assemble_lumped_mass! (src/assemblers/LumpedMass.jl:29-58): the row-sum lumped mass into residual_storage."""
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    asm.ctx.call("fecx_assemble_lumped_mass", ctypes.c_double(density(p.physics)), _ptr(asm.residual_storage))


def assemble_diagonal_(asm, func, Uu, p):
    """assemble_diagonal! (src/assemblers/Diagonal.jl:17-68): the diagonal of the stiffness (func = stiffness) or
    of the consistent mass (func = mass) into residual_storage."""
    _check_vec(asm, Uu, "Uu")
    _update_for_assembly(p, asm, Uu)
    if func is mass:
        asm.ctx.call("fecx_assemble_mass_diagonal", ctypes.c_double(density(p.physics)), _ptr(asm.residual_storage))
    else:
        asm.ctx.call("fecx_assemble_diagonal", _ptr(p.field), _ptr(asm.residual_storage))


def assemble_matrix_free_action_(asm, func_action, Uu, Vu, p):
    """assemble_matrix_free_action! (src/assemblers/MatrixAction.jl:9-77)."""
    _check_vec(asm, Uu, "Uu")
    _check_vec(asm, Vu, "Vu")
    _update_for_assembly(p, asm, Uu, Vu)
    asm.ctx.call("fecx_assemble_matrix_free_action", _ptr(p.field), _ptr(p.hvp_scratch_field),
                 _ptr(asm.stiffness_action_storage))


def assemble_matrix_free_action_full_(asm, func_action, U_full, v_full, p):
    """assemble_matrix_free_action_full! (src/assemblers/MatrixAction.jl:99-149): caller-built full vectors; the
    scratch field's Dirichlet slots are reset to zero afterwards. The full action stays in stiffness_action_storage
    (no extraction of the unknowns)."""
    N = len(asm.dof)
    _check_vec(asm, U_full, "U_full", N)
    _check_vec(asm, v_full, "v_full", N)
    p.field.copy_(U_full)
    p.hvp_scratch_field.copy_(v_full)
    asm.ctx.call("fecx_assemble_matrix_free_action", _ptr(p.field), _ptr(p.hvp_scratch_field),
                 _ptr(asm.stiffness_action_storage))
    if len(asm.dof.dirichlet_dofs):
        idx = torch.from_numpy(asm.dof.dirichlet_dofs - 1).to(asm.device)
        p.hvp_scratch_field.index_fill_(0, idx, 0.0)


def hvp(asm, v):
    return asm.hvp(v)


def set_chunks(asm, chunks=0):
    """Number of pipelined pieces for matrix assemblies (0 or 1: one piece); results are unchanged."""
    asm.ctx.call("fecx_set_chunks", ctypes.c_int(chunks))


def set_profiling(asm, on=True):
    asm.ctx.call("fecx_set_profiling", ctypes.c_int(1 if on else 0))


def profile(asm):
    """{kernel name: (launches, total ms)} accumulated with HIP events on the assembler's stream."""
    n = ctypes.c_int(0)
    asm.ctx.call("fecx_profile_count", ctypes.byref(n))
    out = {}
    for i in range(n.value):
        buf = ctypes.create_string_buffer(64)
        launches, ms = ctypes.c_int64(0), ctypes.c_double(0)
        asm.ctx.call("fecx_profile_entry", ctypes.c_int(i), buf, ctypes.c_int(64), ctypes.byref(launches),
                     ctypes.byref(ms))
        out[buf.value.decode()] = (launches.value, ms.value)
    return out


def reset_profile(asm):
    asm.ctx.call("fecx_profile_reset")


class HostPattern:
    """libfecx's host pattern builder on a host-only context (device -1): DOF numbering, the reduced CSR pattern
    and the COO->CSR index map, without a GPU. Same code path as SparseMatrixAssembler.update_dofs."""

    def __init__(self, mesh, nf, dirichlet_dofs=(), condensed=False):
        self.ctx = _lib.Context(-1)
        self.ne, self.nnpe = mesh.conn.shape
        self.nf = nf
        _set_mesh(self.ctx, mesh, _ELEM_KIND[mesh.element_type])
        if condensed:
            self.ctx.call("fecx_set_condensed", ctypes.c_int(1))
        d = np.ascontiguousarray(dirichlet_dofs, dtype=np.int64)
        self.ctx.call("fecx_set_dofs", ctypes.c_int(nf), _host_ptr(d) if len(d) else None, ctypes.c_int64(len(d)))
        n = ctypes.c_int64(0)
        self.ctx.call("fecx_get_unknown_dofs", ctypes.byref(n), None)
        self.unknown_dofs = np.zeros(n.value, dtype=np.int64)
        self.ctx.call("fecx_get_unknown_dofs", ctypes.byref(n),
                      _host_ptr(self.unknown_dofs) if n.value else None)
        nnz = ctypes.c_int64(0)
        self.ctx.call("fecx_build_pattern", ctypes.byref(nnz))
        self.nnz = nnz.value
        self._nrows = nf * mesh.num_nodes if condensed else len(self.unknown_dofs)

    @property
    def nrows(self):
        """Rows of the pattern: the free dofs, or every dof in condensed mode."""
        return self._nrows

    def pattern(self, base=1):
        rp = np.zeros(self.nrows + 1, dtype=np.int64)
        cv = np.zeros(self.nnz, dtype=np.int64)
        self.ctx.call("fecx_get_pattern", _host_ptr(rp), _host_ptr(cv), ctypes.c_int(base))
        return rp, cv

    def coo_map(self):
        ndof = self.nnpe * self.nf
        m = np.zeros(self.ne * ndof * ndof, dtype=np.int64)
        self.ctx.call("fecx_get_coo_map", _host_ptr(m))
        return m
