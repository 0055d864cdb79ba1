"""MI355X element-assembly path of FiniteElementContainers.jl: residual, CSR Jacobian and Jacobian actions
assembled by hand-written gfx950 HIP kernels in libfecx.so (C ABI: include/fecx.h).

The package directory name contains a dot, so load it by path (``__graft_entry__.load_package()``).
"""
from ._lib import FECXError, exported_symbols, LIB_PATH  # noqa: F401
from .mesh import (StructuredMesh, TetMesh, BlockMesh, split_into_blocks, fan_tet10_mesh, kuhn_tet10_mesh,  # noqa: F401
                   boundary_nodes)
from .exodus import ExodusMesh, UnstructuredMesh, write_exodus  # noqa: F401
from .assembler import (  # noqa: F401
    Poisson, Mechanics, NeoHookean, J2Plasticity, ThreeDimensional, PlaneStrain,
    residual, stiffness, stiffness_action, mass, lumped_mass, diagonal, density,
    FunctionSpace, ScalarFunction, VectorFunction, DofManager, DirichletBC, DirichletBCs,
    CSRMatrix, CSCMatrix, SparseMatrixAssembler, HostPattern, Parameters, create_parameters, create_unknowns, create_field,
    assemble_vector_, assemble_stiffness_, assemble_matrix_, assemble_vector_and_stiffness_,
    assemble_matrix_action_, assemble_matrix_free_action_, assemble_matrix_free_action_full_, hvp,
    assemble_mass_, assemble_lumped_mass_, assemble_diagonal_,
    set_chunks, set_profiling, profile, reset_profile,
)
from . import rng, workloads, surrogate, solvers, exodus  # noqa: F401
from .solvers import (IterativeLinearSolver, IterativeLinearSolverSettings, NewtonSolver,  # noqa: F401
                      NewtonSolverSettings)
from .surrogate import J2Surrogate, NeoHookeanSurrogate, Surrogate  # noqa: F401
