# FECXExt.jl -- the FiniteElementContainers.jl side of the drop-in boundary: a package extension that routes the
# reference's assembler entry points to libfecx.so (C ABI: include/fecx.h) through ccall.
#
# A maintainer adds it as ext/FECXExt.jl with, in Project.toml,
#   [weakdeps]   AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#   [extensions] FECXExt = "AMDGPU"
# next to the package's own AMDGPU extension (ext/AMDGPUExt.jl:1-33), and points ENV["FECX_LIB"] at the built
# library. Julia is not installed where this library is built, so this file is source for that maintainer; the
# Python mirror (assembler.py) drives the same C ABI in the tests.
#
# What it replaces (each method cites the reference function it overrides):
#   SparseMatrixAssembler construction            src/assemblers/SparseMatrixAssembler.jl:59-133
#   update_dofs!(asm, dbcs, pbcs)                 src/assemblers/SparseMatrixAssembler.jl:228-274
#   assemble_vector!(asm, f, Uu, p)               src/assemblers/Vector.jl:4-74
#   assemble_stiffness! / assemble_matrix!        src/assemblers/Matrix.jl:12-75
#   assemble_matrix_action!(asm, f, Uu, Vu, p)    src/assemblers/MatrixAction.jl:154-238
#   assemble_matrix_free_action!(...), _full!     src/assemblers/MatrixAction.jl:9-149
#   assemble_mass!, assemble_lumped_mass!, assemble_diagonal!   src/assemblers/{Matrix,LumpedMass,Diagonal}.jl
#   residual, stiffness, mass, hvp, diagonal, lumped_mass       src/assemblers/Assemblers.jl:253-388,
#                                                               Diagonal.jl:70-83, LumpedMass.jl:60-80
# The DOF plumbing stays the reference's own (_update_for_assembly!, extract_field_unknowns!), on the same device
# arrays: the library sees the full fields the reference has written.
module FECXExt

using AMDGPU
using FiniteElementContainers
using SparseArrays
using SparseMatricesCSR

const FEC = FiniteElementContainers
const libfecx = get(ENV, "FECX_LIB", "libfecx.so")

# ------------------------------------------------------------------ constants of include/fecx.h
const FECX_QUAD4, FECX_HEX8, FECX_TET10 = Cint(0), Cint(1), Cint(2)
const FECX_POISSON, FECX_LINEAR_ELASTIC, FECX_NEOHOOKEAN, FECX_J2 = Cint(0), Cint(1), Cint(2), Cint(3)
const FECX_ERR_NUMERIC = Cint(-5)

struct FECXError <: FEC.AbstractFECError
  code::Cint
  msg::String
end
Base.showerror(io::IO, e::FECXError) = print(io, "libfecx error ", e.code, ": ", e.msg)

mutable struct Context
  ptr::Ptr{Cvoid}
end

function check(ctx::Context, rc::Cint)
  rc == 0 && return nothing
  msg = unsafe_string(ccall((:fecx_last_error, libfecx), Cstring, (Ptr{Cvoid},), ctx.ptr))
  throw(FECXError(rc, msg))
end

function Context(device::Integer)
  out = Ref{Ptr{Cvoid}}(C_NULL)
  rc = ccall((:fecx_create, libfecx), Cint, (Ref{Ptr{Cvoid}}, Cint), out, device)
  rc == 0 || throw(FECXError(rc, "fecx_create failed on device $device"))
  ctx = Context(out[])
  finalizer(c -> ccall((:fecx_destroy, libfecx), Cint, (Ptr{Cvoid},), c.ptr), ctx)
  # launch on the task-local stream KernelAbstractions would have used (src/Utils.jl:157-158)
  check(ctx, ccall((:fecx_set_stream, libfecx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx.ptr, AMDGPU.stream().stream))
  return ctx
end

dptr(x::ROCArray{Float64}) = reinterpret(Ptr{Float64}, pointer(x))
dptr(f::FEC.AbstractField) = dptr(f.data)

# ------------------------------------------------------------------ physics -> registered material
# Physics callbacks cannot cross a C ABI. A physics type is registered by a method of fecx_material returning the
# library's kind and its property vector; the test physics of the reference map as below (their types live in the
# reference's test tree, so the methods are added where those types are in scope):
#   fecx_material(::Poisson, props)   = (FECX_POISSON, [0.0, 1.0])            # f = 1 (source kind 0)
#   fecx_material(::Mechanics, props) = (FECX_LINEAR_ELASTIC, collect(props))  # {rho, K, G}
# (Mechanics(PlaneStrain()) on a quad4 mesh with two displacement fields maps the same way: the library takes the
# plane-strain form from the two fields), and a neo-Hookean or J2 physics of an application to FECX_NEOHOOKEAN
# {lambda, mu} / FECX_J2 {E, nu, sigma_y, H}. Dirichlet BCs on single components need nothing here: the DofManager's
# dirichlet_dofs go to fecx_set_dofs as they are, and the pattern returned is the dof-level one.
function fecx_material end
fecx_material(physics, props) =
  throw(FECXError(Cint(-4), "no libfecx material registered for $(typeof(physics)); add a fecx_material method"))

# ------------------------------------------------------------------ the assembler
# The reference's SparseMatrixAssembler fields (SparseMatrixAssembler.jl:7-28) with the COO element storage replaced
# by the CSR values the library writes, and the library context next to them.
mutable struct FECXAssembler{Condensed, Var} <: FEC.AbstractAssembler{FEC.DofManager}
  dof::FEC.DofManager
  ctx::Context
  sparse_matrix_type::Symbol          # :csr or :csc
  rowptr::Vector{Int64}               # 1-based, n + 1 (the CSC colptr too: the pattern is symmetric)
  colval::Vector{Int64}               # 1-based, nnz (the CSC rowval too)
  constraint_storage::ROCVector{Float64}
  residual_storage::FEC.AbstractField
  residual_unknowns::ROCVector{Float64}
  stiffness_storage::ROCVector{Float64}
  mass_storage::ROCVector{Float64}
  stiffness_action_storage::FEC.AbstractField
  stiffness_action_unknowns::ROCVector{Float64}
  stiffness_adjusted::Bool
  mass_adjusted::Bool
end

FEC._is_condensed(::FECXAssembler{C}) where C = C

function elem_kind(fspace, b)
  nnpe = FEC.num_entities_per_element(fspace, b)
  nd = size(FEC.coordinates(fspace), 1)
  nnpe == 4 && nd == 2 && return FECX_QUAD4
  nnpe == 8 && nd == 3 && return FECX_HEX8
  nnpe == 10 && nd == 3 && return FECX_TET10
  throw(FECXError(Cint(-4), "element with $nnpe nodes in $nd dimensions has no libfecx kernel"))
end

"""
    FECXAssembler(var; sparse_matrix_type = :csr, use_condensed = false, device = 0, fp64_geometry = false)

Drop-in for `SparseMatrixAssembler(var; sparse_matrix_type, use_condensed)` (SparseMatrixAssembler.jl:59-133)
whose element loops and sparse reduction run in libfecx.so on an MI355X. `fp64_geometry = true` evaluates the tet10
element geometry, field gradient and stress in FP64 (fecx_set_fp64_geometry).
"""
function FECXAssembler(var::FEC.AbstractFunction; sparse_matrix_type::Symbol = :csr, use_condensed::Bool = false,
                       device::Integer = 0, fp64_geometry::Bool = false)
  sparse_matrix_type in (:csr, :csc) || throw(FEC.UnsupportedSparseMatrixType(sparse_matrix_type))
  dof = FEC.DofManager(var; use_condensed = use_condensed)
  fspace = FEC.function_space(dof)
  ctx = Context(device)
  X = Array(FEC.coordinates(fspace).data)               # nd x nn
  conns = fspace.elem_conns                             # Connectivity: data, offsets, nelems per block
  nb = FEC.num_blocks(fspace)
  kind = elem_kind(fspace, 1)
  all(b -> elem_kind(fspace, b) == kind, 1:nb) ||
    throw(FECXError(Cint(-4), "libfecx takes blocks of one element kind"))
  block_ne = Int64[FEC.num_elements(fspace, b) for b in 1:nb]
  conn = Vector{Int64}(Array(conns.data))               # blocks concatenated, 1-based
  check(ctx, ccall((:fecx_set_mesh_blocks, libfecx), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Int64, Ptr{Float64}, Cint, Ptr{Int64}, Ptr{Int64}),
                   ctx.ptr, kind, size(X, 1), size(X, 2), X, nb, block_ne, conn))
  use_condensed && check(ctx, ccall((:fecx_set_condensed, libfecx), Cint, (Ptr{Cvoid}, Cint), ctx.ptr, 1))
  fp64_geometry && check(ctx, ccall((:fecx_set_fp64_geometry, libfecx), Cint, (Ptr{Cvoid}, Cint), ctx.ptr, 1))
  ND, NN = size(dof)
  asm = FECXAssembler{use_condensed, typeof(var)}(
    dof, ctx, sparse_matrix_type, Int64[], Int64[], AMDGPU.zeros(Float64, ND * NN),
    FEC.create_field(dof) |> AMDGPU.roc, AMDGPU.zeros(Float64, 0), AMDGPU.zeros(Float64, 0),
    AMDGPU.zeros(Float64, 0), FEC.create_field(dof) |> AMDGPU.roc, AMDGPU.zeros(Float64, 0), true, true)
  set_dofs!(asm, Int64[])
  return asm
end

# fecx_set_dofs + fecx_build_pattern + fecx_get_pattern: update_dofs! for this assembler
function set_dofs!(asm::FECXAssembler, ddofs::Vector{Int64})
  ctx, (ND, NN) = asm.ctx, size(asm.dof)
  check(ctx, ccall((:fecx_set_dofs, libfecx), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}, Int64),
                   ctx.ptr, ND, ddofs, length(ddofs)))
  nnz = Ref{Int64}(0)
  check(ctx, ccall((:fecx_build_pattern, libfecx), Cint, (Ptr{Cvoid}, Ref{Int64}), ctx.ptr, nnz))
  nu = Ref{Int64}(0)
  check(ctx, ccall((:fecx_get_unknown_dofs, libfecx), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}), ctx.ptr, nu,
                   C_NULL))
  nrows = FEC._is_condensed(asm) ? ND * NN : nu[]
  asm.rowptr = Vector{Int64}(undef, nrows + 1)
  asm.colval = Vector{Int64}(undef, nnz[])
  check(ctx, ccall((:fecx_get_pattern, libfecx), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Cint),
                   ctx.ptr, asm.rowptr, asm.colval, 1))
  asm.stiffness_storage = AMDGPU.zeros(Float64, nnz[])
  asm.mass_storage = AMDGPU.zeros(Float64, nnz[])
  asm.residual_unknowns = AMDGPU.zeros(Float64, nu[])
  asm.stiffness_action_unknowns = AMDGPU.zeros(Float64, nu[])
  c = zeros(ND * NN); c[ddofs] .= 1.0
  copyto!(asm.constraint_storage, c)
  return nothing
end

# update_dofs!(asm, dirichlet_bcs, periodic_bcs) (SparseMatrixAssembler.jl:228-274)
function FEC.update_dofs!(asm::FECXAssembler, dirichlet_bcs::FEC.DirichletBCs, periodic_bcs::FEC.PeriodicBCs)
  length(periodic_bcs) == 0 || throw(FECXError(Cint(-4), "periodic BCs are not assembled by libfecx"))
  ddofs = length(dirichlet_bcs) > 0 ? Vector{Int64}(Array(FEC.dirichlet_dofs(dirichlet_bcs))) : Int64[]
  FEC.update_dofs!(asm.dof, ddofs, Int64[], Int64[])
  set_dofs!(asm, ddofs)
  return nothing
end

# the Parameters' physics (one, or one per block as in foreach_block, src/Utils.jl:171-189) to the library
function set_materials!(asm::FECXAssembler, p::FEC.AbstractParameters)
  phys, props = values(p.physics), values(p.properties)
  for b in 1:length(phys)
    kind, pv = fecx_material(phys[b], props[b])
    check(asm.ctx, ccall((:fecx_set_block_material, libfecx), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cint),
                         asm.ctx.ptr, b - 1, kind, pv, length(pv)))
  end
  return nothing
end

# ------------------------------------------------------------------ entry points
function FEC.assemble_vector!(asm::FECXAssembler, func::F, Uu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu)
  set_materials!(asm, p)
  check(asm.ctx, ccall((:fecx_assemble_vector, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                       asm.ctx.ptr, dptr(p.field), dptr(asm.residual_storage)))
  return nothing
end

function FEC.assemble_stiffness!(asm::FECXAssembler, func::F, Uu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu)
  set_materials!(asm, p)
  check(asm.ctx, ccall((:fecx_assemble_matrix, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                       asm.ctx.ptr, dptr(p.field), dptr(asm.stiffness_storage)))
  asm.stiffness_adjusted = false
  return nothing
end

FEC.assemble_matrix!(asm::FECXAssembler, func::F, Uu, p) where F <: Function =
  FEC.assemble_stiffness!(asm, func, Uu, p)

function FEC.assemble_matrix_action!(asm::FECXAssembler, func::F, Uu, Vu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu, Vu)
  if func === FEC.mass
    check(asm.ctx, ccall((:fecx_assemble_mass_action, libfecx), Cint,
                         (Ptr{Cvoid}, Cdouble, Ptr{Float64}, Ptr{Float64}),
                         asm.ctx.ptr, density(p), dptr(p.hvp_scratch_field), dptr(asm.stiffness_action_storage)))
  else
    set_materials!(asm, p)
    check(asm.ctx, ccall((:fecx_assemble_matrix_action, libfecx), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         asm.ctx.ptr, dptr(p.field), dptr(p.hvp_scratch_field), dptr(asm.stiffness_action_storage)))
  end
  return nothing
end

function FEC.assemble_matrix_free_action!(asm::FECXAssembler, func::F, Uu, Vu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu, Vu)
  set_materials!(asm, p)
  check(asm.ctx, ccall((:fecx_assemble_matrix_free_action, libfecx), Cint,
                       (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                       asm.ctx.ptr, dptr(p.field), dptr(p.hvp_scratch_field), dptr(asm.stiffness_action_storage)))
  return nothing
end

# assemble_matrix_free_action_full! (MatrixAction.jl:99-149): caller-built full vectors, scratch BC slots reset
function FEC.assemble_matrix_free_action_full!(asm::FECXAssembler, func::F, U_full, v_full, p) where F <: Function
  FEC._update_for_assembly_full!(p, U_full, v_full)
  set_materials!(asm, p)
  check(asm.ctx, ccall((:fecx_assemble_matrix_free_action, libfecx), Cint,
                       (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                       asm.ctx.ptr, dptr(p.field), dptr(p.hvp_scratch_field), dptr(asm.stiffness_action_storage)))
  bc = asm.dof.dirichlet_dofs
  isempty(bc) || (p.hvp_scratch_field.data[bc] .= 0.0)
  return nothing
end

# the test physics' mass callbacks multiply by the density: props[1] of Mechanics, 1 otherwise
density(p) = (k = fecx_material(first(values(p.physics)), first(values(p.properties))); k[1] == FECX_LINEAR_ELASTIC ?
              k[2][1] : 1.0)

function FEC.assemble_mass!(asm::FECXAssembler, func::F, Uu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu)
  check(asm.ctx, ccall((:fecx_assemble_mass, libfecx), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}),
                       asm.ctx.ptr, density(p), dptr(asm.mass_storage)))
  asm.mass_adjusted = false
  return nothing
end

function FEC.assemble_lumped_mass!(asm::FECXAssembler, func::F, Uu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu)
  check(asm.ctx, ccall((:fecx_assemble_lumped_mass, libfecx), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}),
                       asm.ctx.ptr, density(p), dptr(asm.residual_storage)))
  return nothing
end

function FEC.assemble_diagonal!(asm::FECXAssembler, func::F, Uu, p) where F <: Function
  FEC._update_for_assembly!(p, asm.dof, Uu)
  if func === FEC.mass
    check(asm.ctx, ccall((:fecx_assemble_mass_diagonal, libfecx), Cint, (Ptr{Cvoid}, Cdouble, Ptr{Float64}),
                         asm.ctx.ptr, density(p), dptr(asm.residual_storage)))
  else
    set_materials!(asm, p)
    check(asm.ctx, ccall((:fecx_assemble_diagonal, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                         asm.ctx.ptr, dptr(p.field), dptr(asm.residual_storage)))
  end
  return nothing
end

# ------------------------------------------------------------------ accessors (Assemblers.jl:253-388)
function matrix(asm::FECXAssembler, vals::ROCVector{Float64})
  n = length(asm.rowptr) - 1
  v = Array(vals)
  asm.sparse_matrix_type === :csc && return SparseMatrixCSC(n, n, asm.rowptr, asm.colval, v)
  return SparseMatrixCSR{1}(n, n, asm.rowptr, asm.colval, v)
end

function FEC.residual(asm::FECXAssembler)
  if FEC._is_condensed(asm)
    check(asm.ctx, ccall((:fecx_adjust_vector_for_constraints, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}),
                         asm.ctx.ptr, dptr(asm.residual_storage)))
    return asm.residual_storage.data
  end
  FEC.extract_field_unknowns!(asm.residual_unknowns, asm.dof, asm.residual_storage)
  return asm.residual_unknowns
end

function FEC.stiffness(asm::FECXAssembler)
  if FEC._is_condensed(asm) && !asm.stiffness_adjusted
    check(asm.ctx, ccall((:fecx_adjust_matrix_for_constraints, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble),
                         asm.ctx.ptr, dptr(asm.stiffness_storage), 1.0e6))
    asm.stiffness_adjusted = true
  end
  return matrix(asm, asm.stiffness_storage)
end

function FEC.mass(asm::FECXAssembler)
  if FEC._is_condensed(asm) && !asm.mass_adjusted
    check(asm.ctx, ccall((:fecx_adjust_matrix_for_constraints, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cdouble),
                         asm.ctx.ptr, dptr(asm.mass_storage), 1.0e6))
    asm.mass_adjusted = true
  end
  return matrix(asm, asm.mass_storage)
end

function FEC.hvp(asm::FECXAssembler, v)
  if FEC._is_condensed(asm)
    check(asm.ctx, ccall((:fecx_adjust_action_for_constraints, libfecx), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                         asm.ctx.ptr, dptr(asm.stiffness_action_storage), dptr(v)))
    return asm.stiffness_action_storage.data
  end
  FEC.extract_field_unknowns!(asm.stiffness_action_unknowns, asm.dof, asm.stiffness_action_storage)
  return asm.stiffness_action_unknowns
end

function FEC.diagonal(asm::FECXAssembler)
  FEC._is_condensed(asm) && return asm.residual_storage.data
  FEC.extract_field_unknowns!(asm.residual_unknowns, asm.dof, asm.residual_storage)
  return asm.residual_unknowns
end

FEC.lumped_mass(asm::FECXAssembler) = FEC.diagonal(asm)

# ------------------------------------------------------------------ quadrature state (StateVariableField, Fields.jl:401-455)
# The library keeps FP32 copies of state_old / state_new: load the old state before a stateful assembly, read the
# new one back after it, and commit (state_old <- state_new) at the end of a converged load step.
function set_state_old!(asm::FECXAssembler, p)
  s = Vector{Float32}(Array(p.state_old.data))
  check(asm.ctx, ccall((:fecx_set_state_old, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float32}), asm.ctx.ptr, s))
end
function get_state_new!(p, asm::FECXAssembler)
  s = Vector{Float32}(undef, length(p.state_new.data))
  check(asm.ctx, ccall((:fecx_get_state_new, libfecx), Cint, (Ptr{Cvoid}, Ptr{Float32}), asm.ctx.ptr, s))
  copyto!(p.state_new.data, Float64.(s))
end

export FECXAssembler, fecx_material

end # module
