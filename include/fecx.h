/*
 * fecx.h -- C ABI of the MI355X element-assembly library (libfecx.so).
 *
 * This is the drop-in boundary for FiniteElementContainers.jl's element-assembly hot path.
 * Every entry point below replaces one reference function (cited per declaration,
 * paths relative to the FiniteElementContainers.jl repository). The Julia side keeps the
 * reference's own entry points (assemble_vector!, assemble_stiffness!/assemble_matrix!,
 * assemble_matrix_action!, assemble_matrix_free_action!, residual, stiffness, hvp) and
 * forwards to these functions through ccall.
 *
 * Conventions
 *  - Return value: 0 on success, a negative FECX_ERR_* code otherwise; the message is
 *    available from fecx_last_error(ctx) until the next call on that context.
 *  - Setup calls (fecx_set_*, fecx_get_*) take HOST pointers, borrowed for the call.
 *  - Assembly calls take DEVICE pointers (hipDeviceptr_t cast to the element type) that
 *    live on the context's device, e.g. a Julia ROCArray or a torch tensor; no copies.
 *  - Indices crossing the boundary are 1-based int64 like the reference, unless stated.
 *  - Nodal fields are NF x NN column-major (dof = NF*(node-1)+d, src/Fields.jl:36-40).
 *  - Calls on one context are not re-entrant. Device work is enqueued on the context's
 *    stream (fecx_set_stream). Assembly calls return after the stream is synchronised,
 *    so results can be read at once (the reference's callers read them immediately,
 *    src/Solvers.jl:128-153) and a folded element is reported by that call. The DOF
 *    plumbing calls (fecx_update_field_*, fecx_extract_field_unknowns) only enqueue:
 *    work ordered on the same stream sees their results.
 */
#ifndef FECX_H
#define FECX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fecx_ctx fecx_ctx;

/* error codes */
#define FECX_OK 0
#define FECX_ERR_ARG -1         /* bad argument or size mismatch (the reference's @assert) */
#define FECX_ERR_STATE -2       /* call out of order (e.g. assembly before fecx_build_pattern) */
#define FECX_ERR_HIP -3         /* HIP runtime failure */
#define FECX_ERR_UNSUPPORTED -4 /* element/physics/BC combination not supported (UnsupportedSparseMatrixType analogue) */
#define FECX_ERR_NUMERIC -5     /* non-positive Jacobian determinant or NaN met in an element kernel */

/* element kinds (ReferenceFiniteElements Quad4 / Hex8 / Tet10) */
#define FECX_QUAD4 0
#define FECX_HEX8 1
#define FECX_TET10 2

/* physics kinds (the reference's physics callbacks cannot cross a C ABI; each kind names one)
 * POISSON         test/poisson/TestPoissonCommon.jl:75-127; props = {source_kind, coefficient}
 *                 source_kind 0: f = coefficient; 1: f = coefficient * nd*pi^2 * prod_d sin(pi x_d)
 * LINEAR_ELASTIC  test/mechanics/TestMechanicsCommon.jl:14-20,140-222; props = {rho, K, G}
 * NEOHOOKEAN      compressible neo-Hookean; props = {lambda, mu}
 * J2              small-strain J2 plasticity, linear isotropic hardening; props = {E, nu, sigma_y, H};
 *                 7 state variables per quadrature point
 * J2_SURROGATE    J2 whose material update is the FECX_MLP_J2 surrogate, checked and with plain-model
 *                 fallback (fecx_surrogate_*); props as J2
 * NH_SURROGATE    neo-Hookean whose stress and tangent come from the FECX_MLP_NH surrogate, checked and
 *                 with plain-model fallback; props as NEOHOOKEAN */
#define FECX_POISSON 0
#define FECX_LINEAR_ELASTIC 1
#define FECX_NEOHOOKEAN 2
#define FECX_J2 3
#define FECX_J2_SURROGATE 4
#define FECX_NH_SURROGATE 5

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `device`. Replaces the `to_backend(ROCBackend, ...)` adaption
 * of the assembler and parameters (ext/AMDGPUExt.jl:8-14). */
int fecx_create(fecx_ctx** ctx, int device);

/* Frees the context and every device buffer it owns. */
int fecx_destroy(fecx_ctx* ctx);

/* Message of the last failed call on ctx (never NULL). */
const char* fecx_last_error(const fecx_ctx* ctx);

/* Stream for every kernel this context launches (a hipStream_t; NULL = the null stream).
 * Replaces the task-local stream KernelAbstractions launches on (src/Utils.jl:157-158). */
int fecx_set_stream(fecx_ctx* ctx, void* hip_stream);

/* ---------------------------------------------------------------- setup (host pointers) */

/* One element block: coords is nd x nn column-major FP64 (the mesh's nodal_coords H1Field,
 * src/meshes/StructuredMesh.jl:91), conn is nnpe x ne column-major 1-based (Connectivity.data
 * of that block, src/Fields.jl:129-170). Replaces FunctionSpace's connectivity/coordinate
 * storage (src/FunctionSpaces.jl:158-237). */
int fecx_set_mesh(fecx_ctx* ctx, int elem_kind, int nd, int64_t nn, const double* coords, int64_t ne,
                  const int64_t* conn);

/* Up to FECX_MAX_BLOCKS element blocks of one element kind (the reference's FunctionSpace blocks, at most
 * MAX_BLOCKS = 16, src/FunctionSpaces.jl:63): conn is the concatenation of the blocks' nnpe x block_ne[b]
 * connectivities in block order (Connectivity.data with its per-block offsets, src/Fields.jl:129-188), so element
 * ids run block after block as in the reference's pattern (SparsityPatterns.jl:53-117). Each block may carry its own
 * physics (fecx_set_block_material), run block by block like foreach_block (src/Utils.jl:171-263).
 * fecx_set_mesh is this call with one block. */
#define FECX_MAX_BLOCKS 16
int fecx_set_mesh_blocks(fecx_ctx* ctx, int elem_kind, int nd, int64_t nn, const double* coords, int nblocks,
                         const int64_t* block_ne, const int64_t* conn);

/* Number of element blocks of the mesh. */
int fecx_get_num_blocks(fecx_ctx* ctx, int* nblocks);

/* nf fields per node and the Dirichlet DOF list, sorted and unique, 1-based (the reference's
 * dirichlet_dofs(bcs), src/bcs/DirichletBCs.jl:396-398). Performs update_dofs!
 * (src/DofManagers.jl:227-298) without periodic pairs. For nf > 1 the list may fix single components of a node (a
 * DirichletBC on one component, src/bcs/DirichletBCs.jl:49-98): the CSR of fecx_get_pattern is then the reference's
 * dof-level pattern over the unknown dofs; the kernels assemble the node blocks of every node with a free dof and one
 * gather forms it (the solver calls take that CSR too). */
int fecx_set_dofs(fecx_ctx* ctx, int nf, const int64_t* dirichlet_dofs, int64_t n_dirichlet);

/* Condensed mode, SparseMatrixAssembler(var; use_condensed = true) (src/assemblers/SparseMatrixAssembler.jl:126-133,
 * DofManager{true}, src/DofManagers.jl:41-42). on = 1 keeps every DOF in the pattern: fecx_build_pattern then makes
 * the N x N CSR over all dofs, Dirichlet rows and columns included, as a condensed assembler never reduces its
 * pattern (SparseMatrixAssembler.jl:246-251); the Dirichlet set may then fix single components of a node. The
 * unknown vectors of fecx_update_field_unknowns are full length (U[unknown_dofs] = Uu[unknown_dofs], the condensed
 * kernel of update_field_unknowns!, src/DofManagers.jl:328-364), and the accessors adjust the assembled results with
 * the fecx_adjust_*_for_constraints calls below. Takes effect at the next fecx_set_dofs (called again after this). */
int fecx_set_condensed(fecx_ctx* ctx, int on);

/* n_unknowns and, if unknown_dofs != NULL, the unknown DOF ids (1-based, ascending):
 * DofManager.unknown_dofs after update_dofs! (src/DofManagers.jl:262-263). */
int fecx_get_unknown_dofs(fecx_ctx* ctx, int64_t* n_unknowns, int64_t* unknown_dofs);

/* Builds the reduced CSR sparsity pattern and the COO->CSR index maps on the host and uploads
 * them. Replaces SparseMatrixPattern(dof) + _update_dofs!(pattern, ...) + the COO->CSC->CSR
 * conversion of stiffness(asm) (src/assemblers/SparsityPatterns.jl:53-117,160-231,301-329).
 * *nnz receives the number of CSR entries. A node may couple with at most 255 nodes (itself
 * included): FECX_ERR_UNSUPPORTED otherwise. */
int fecx_build_pattern(fecx_ctx* ctx, int64_t* nnz);

/* CSR pattern (n_unknowns+1 row pointers, N+1 in condensed mode; nnz column indices) with the given index base
 * (1 = SparseMatrixCSR(csc) convention, 0 = C). Bit-exact with SparseArrays.sparse! followed by
 * SparseMatrixCSR (src/assemblers/SparsityPatterns.jl:301-308,326-329). The element pattern is structurally
 * symmetric, so the same arrays are the CSC colptr/rowval of sparse! itself (sparse_matrix_type = :csc, the
 * reference's default, SparseMatrixAssembler.jl:66, _csc_matrix SparsityPatterns.jl:301-308); the tangents and masses
 * assembled here are symmetric bit for bit, so the CSR values are also the CSC nzval. */
int fecx_get_pattern(fecx_ctx* ctx, int64_t* rowptr, int64_t* colval, int index_base);

/* The COO->CSR index map: for every element-storage slot (ne * (nf*nnpe)^2 entries, element-major,
 * then row i, then column j, as SparsityPatterns.jl:72-84 enumerates them) the 0-based CSR
 * position the slot sums into, or -1 when the slot's row or column is a Dirichlet DOF
 * (dropped by _update_dofs!, SparsityPatterns.jl:195-213). */
int fecx_get_coo_map(fecx_ctx* ctx, int64_t* map);

/* Physics kind and properties for every block; also (re)allocates the quadrature state
 * (NS x NQ x NE, FP32) set to zero when the kind carries state (src/Parameters.jl:1-23). */
int fecx_set_material(fecx_ctx* ctx, int kind, const double* props, int nprops);

/* Physics kind and properties of element block `block` (0-based): the per-block entries of the Parameters'
 * physics and properties tuples (src/Parameters.jl:37-73) that foreach_block hands to each block
 * (src/Utils.jl:171-189). Assembly needs every block set. The state (NS x NQ x NE over all blocks, element-id
 * order) has NS = 7 when any block is J2, and is set to zero. Surrogate kinds need a one-block mesh. */
int fecx_set_block_material(fecx_ctx* ctx, int block, int kind, const double* props, int nprops);

/* Number of state variables per quadrature point and quadrature points per element. */
int fecx_get_state_size(fecx_ctx* ctx, int* ns, int* nq);

/* Copies state_old from / state_new to host FP32 arrays of NS*NQ*NE (StateVariableField layout,
 * src/Fields.jl:401-455). */
int fecx_set_state_old(fecx_ctx* ctx, const float* state);
int fecx_get_state_new(fecx_ctx* ctx, float* state);

/* state_old <- state_new. The reference has no state swap (the caller owns it); this is it. */
int fecx_state_commit(fecx_ctx* ctx);

/* ---------------------------------------------------------------- assembly (device pointers) */

/* Residual into the full NF x NN field R (zeroed first). U is the full field with Dirichlet
 * values and unknowns already in place. Replaces assemble_vector!(storage, pattern, dof, func,
 * Uu, p) after _update_for_assembly! (src/assemblers/Vector.jl:25-74, Assemblers.jl:402-432,72-87).
 * Writes state_new for stateful physics. */
int fecx_assemble_vector(fecx_ctx* ctx, const double* U, double* R);

/* CSR values (nnz, FP64) of the Jacobian at U. Replaces assemble_matrix! followed by
 * stiffness(asm) for sparse_matrix_type = :csr (src/assemblers/Matrix.jl:35-75,
 * Assemblers.jl:376-388, SparsityPatterns.jl:301-329). */
int fecx_assemble_matrix(fecx_ctx* ctx, const double* U, double* vals);

/* Residual and Jacobian in one pass over the elements (one residual + one stiffness call of
 * IterativeLinearSolver.solve!, src/Solvers.jl:133-141). */
int fecx_assemble_vector_matrix(fecx_ctx* ctx, const double* U, double* R, double* vals);

/* fecx_assemble_vector_matrix that allocates nothing: the element-private outputs of its two-phase assembly go to a
 * file-scope device array of FECX_NOALLOC_SCRATCH_BYTES (64 MiB: about 48,000 hex8 or 31,000 tet10 elements with
 * three fields) that each GPU holds until the process ends, instead of the context's buffers; where the context takes
 * the fused kernel (fecx_get_scatter_path: structured hex8 meshes), no element-private output exists and the mesh
 * size is not limited. The atomic and coloured designs are not taken here (their zero fill is a memset). No
 * counterpart in the reference; it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The context's own set-up (mesh, dofs, pattern, material) is
 * done beforehand as for every entry point. Two-phase meshes whose element storage exceeds the array get
 * FECX_ERR_ARG; surrogate materials get FECX_ERR_UNSUPPORTED. All calls on one GPU share the array, which keeps the
 * last call's contents (every call overwrites what it reads), so calls that overlap on one GPU -- from several threads
 * or streams -- race. */
#define FECX_NOALLOC_SCRATCH_BYTES (64 << 20)
int fecx_assemble_vector_matrix_noalloc(fecx_ctx* ctx, const double* U, double* R, double* vals);

/* Assembled Jacobian action into the full field KV: per element K_el is built and multiplied by
 * v_el. Replaces assemble_matrix_action! (src/assemblers/MatrixAction.jl:154-238). */
int fecx_assemble_matrix_action(fecx_ctx* ctx, const double* U, const double* V, double* KV);

/* Matrix-free Jacobian action into the full field KV (per quadrature point K_q v).
 * Replaces assemble_matrix_free_action! / _full! (src/assemblers/MatrixAction.jl:9-77,99-149). */
int fecx_assemble_matrix_free_action(fecx_ctx* ctx, const double* U, const double* V, double* KV);

/* Element-Jacobian diagonals at U into the full field D (zeroed first); diagonal(asm) is D at the unknowns.
 * Replaces assemble_diagonal!(asm, stiffness, Uu, p) (src/assemblers/Diagonal.jl:17-68, the AssembledDiagonal
 * accumulation Assemblers.jl:42-45). */
int fecx_assemble_diagonal(fecx_ctx* ctx, const double* U, double* D);

/* Mass assemblies with the test physics' mass callbacks M_el = density JxW kron(N N^T, I_NF) and
 * m_el = density JxW N_a (test/mechanics/TestMechanicsCommon.jl:52-124; density = props[1] for Mechanics, 1 for
 * Poisson, test/poisson/TestPoissonCommon.jl:18-41). They need no material and no field.
 * fecx_assemble_mass: CSR values (nnz) of the consistent mass on the Jacobian's pattern; replaces assemble_mass!
 *   followed by mass(asm) (src/assemblers/Matrix.jl:1-10, Assemblers.jl:253-336).
 * fecx_assemble_lumped_mass: row-sum lumped mass into the full field M; replaces assemble_lumped_mass! and
 *   lumped_mass(asm) (src/assemblers/LumpedMass.jl:29-80).
 * fecx_assemble_mass_diagonal: diagonal of the consistent mass into the full field D; replaces
 *   assemble_diagonal!(asm, mass, Uu, p) (src/assemblers/Diagonal.jl:17-68).
 * fecx_assemble_mass_action: M V into the full field MV; replaces assemble_matrix_action!(asm, mass, Uu, Vu, p)
 *   (src/assemblers/MatrixAction.jl:154-238). */
int fecx_assemble_mass(fecx_ctx* ctx, double density, double* vals);
int fecx_assemble_lumped_mass(fecx_ctx* ctx, double density, double* M);
int fecx_assemble_mass_diagonal(fecx_ctx* ctx, double density, double* D);
int fecx_assemble_mass_action(fecx_ctx* ctx, double density, const double* V, double* MV);

/* Condensed-mode adjustments with constraint_storage c (1 at the Dirichlet dofs; src/assemblers/Utils.jl:53-167), on
 * device arrays, enqueued and synchronised. stiffness(asm) / mass(asm): every line of a constrained dof of the N x N
 * matrix (a CSR row, or a CSC column: the same values, see fecx_get_pattern) is multiplied by 1 - c = 0 and its
 * diagonal entry gets + penalty, penalty = penalty_scale tr(A) / N (the reference uses 1e6) with the trace taken
 * before the adjustment (Utils.jl:88-145). residual(asm): b = (1 - c) b (Utils.jl:161-167). hvp(asm, v):
 * Av = (1 - c) Av + c v (Utils.jl:150-159). FECX_ERR_STATE unless the context is in condensed mode. */
int fecx_adjust_matrix_for_constraints(fecx_ctx* ctx, double* vals, double penalty_scale);
int fecx_adjust_vector_for_constraints(fecx_ctx* ctx, double* b);
int fecx_adjust_action_for_constraints(fecx_ctx* ctx, double* Av, const double* v);

/* U[unknown_dofs] = Uu: update_field_unknowns! (src/DofManagers.jl:328-411); in condensed mode Uu is full length and
 * U[unknown_dofs] = Uu[unknown_dofs]. */
int fecx_update_field_unknowns(fecx_ctx* ctx, const double* Uu, double* U);

/* U[dirichlet_dofs] = vals (n_dirichlet values, in the order given to fecx_set_dofs):
 * update_field_dirichlet_bcs! (src/bcs/DirichletBCs.jl:411-418). */
int fecx_update_field_dirichlet(fecx_ctx* ctx, const double* vals, double* U);

/* Uu = U[unknown_dofs]: extract_field_unknowns! (src/DofManagers.jl:203-213). */
int fecx_extract_field_unknowns(fecx_ctx* ctx, const double* U, double* Uu);

/* ---------------------------------------------------------------- linear solves on the assembled Jacobian
 * The reference solves every Newton step with Krylov.jl's CG on stiffness(asm) on the CPU
 * (IterativeLinearSolver, src/Solvers.jl:97-153). These run on the GPU on the CSR values the
 * assembly wrote (device pointers; vectors over the unknowns, FP64). Reductions are deterministic. */

/* y = K x. */
int fecx_csr_spmv(fecx_ctx* ctx, const double* vals, const double* x, double* y);
/* dinv = 1 / diag(K) (Jacobi preconditioner; AssembledDiagonal, src/assemblers/Diagonal.jl). */
int fecx_csr_diagonal_inverse(fecx_ctx* ctx, const double* vals, double* dinv);
/* Conjugate gradients for K x = b from the initial x: stops when ||r|| <= atol + rtol ||b|| (Krylov.jl's
 * cg rule) or after maxit iterations; precond 1 = Jacobi, 0 = none. *iters and *resnorm report the
 * iterations taken and the final ||r|| (of the recurrence). */
int fecx_cg_solve(fecx_ctx* ctx, const double* vals, const double* b, double* x, double rtol, double atol, int maxit,
                  int precond, int* iters, double* resnorm);

/* ||v||_2 (host *out; deterministic two-stage sum): the Newton loop's norms (src/Solvers.jl:200-203). */
int fecx_vector_norm(fecx_ctx* ctx, const double* v, int64_t n, double* out);
/* y += a x: the Newton update map!((x, y) -> x + y, Uu, Uu, ΔUu) (src/Solvers.jl:147-149). */
int fecx_vector_axpy(fecx_ctx* ctx, int64_t n, double a, const double* x, double* y);

/* ---------------------------------------------------------------- learned constitutive surrogate
 * No counterpart in the reference, whose materials are energy functions differentiated by AD inside
 * the element loop (test/mechanics/TestMechanicsCommon.jl:140-222). A small MLP with GELU hidden
 * layers runs batched over quadrature points on BF16 MFMA with FP32 accumulation. It is trained on
 * the GPU by this library's own forward/backward kernels, on samples the plain material model
 * generates. The plain model stays the fallback wherever the surrogate's output fails a residual
 * check of the material equations, and it is the reference for the surrogate's error.
 *   FECX_MLP_J2  13 -> 128 -> 128 -> 128 -> 13. Raw input per sample (16 floats): H (9,
 *                row-major displacement gradient), state (7: plastic strain in the order
 *                11,21,31,22,32,33, then eqps). Raw output (13): stress (6, same order), plastic
 *                strain increment (6), eqps increment. Features: symmetric strain (6), state (7).
 *   FECX_MLP_NH  9 -> 256 -> 256 -> 15. Raw input H (9). Raw output: P (9, row-major), C^-1 - I
 *                (6, same symmetric order).
 * Inputs and outputs are standardized by per-column shift/scale (fecx_surrogate_fit_scaling).
 * Array arguments of the batched calls are device pointers. */
#define FECX_MLP_J2 0
#define FECX_MLP_NH 1

/* (Re)creates the network with He-normal weights (SplitMix64 stream `seed`) and zero biases. */
int fecx_surrogate_init(fecx_ctx* ctx, int arch, uint64_t seed);
int fecx_surrogate_num_params(fecx_ctx* ctx, int64_t* nparams);
/* Canonical FP32 parameter vector (host): per layer W (out x in, row-major), then b. */
int fecx_surrogate_get_params(fecx_ctx* ctx, float* params);
int fecx_surrogate_set_params(fecx_ctx* ctx, const float* params);
/* Standardization (host arrays): features (13 or 9) and raw outputs (13 or 15). */
int fecx_surrogate_set_scaling(fecx_ctx* ctx, const float* in_shift, const float* in_scale, const float* out_shift,
                               const float* out_scale);
int fecx_surrogate_get_scaling(fecx_ctx* ctx, float* in_shift, float* in_scale, float* out_shift, float* out_scale);
/* Column means and standard deviations of the features of X and of T (n samples, device). */
int fecx_surrogate_fit_scaling(fecx_ctx* ctx, const float* X, const float* T, int64_t n);
/* Network output in raw units for n raw inputs. */
int fecx_surrogate_forward(fecx_ctx* ctx, const float* X, int64_t n, float* Y);
/* Mean-squared error of the standardized outputs over n samples (host *loss) and its gradient with
 * respect to the canonical parameters (device, nparams floats); no update. */
int fecx_surrogate_gradient(fecx_ctx* ctx, const float* X, const float* T, int64_t n, double* loss, float* grad);
/* Adam training: `epochs` passes over the n samples in shuffled minibatches of `batch` (SplitMix64
 * stream `seed`), learning rate lr (beta1 0.9, beta2 0.999, eps 1e-8). loss_per_epoch (host,
 * `epochs` doubles, or NULL) receives the mean minibatch loss of each epoch. */
int fecx_surrogate_train(fecx_ctx* ctx, const float* X, const float* T, int64_t n, int epochs, int batch, float lr,
                         uint64_t seed, double* loss_per_epoch);
/* Residual-check tolerance of the surrogate material update (relative; default 0.02). */
int fecx_surrogate_set_tolerance(fecx_ctx* ctx, float tol);
/* on = 1: the next surrogate passes reuse the last pass's per-point decisions (network output or plain
 * model) instead of checking again, so that the residual is one fixed function of the field during a
 * Newton solve; on = 0 (default): every pass checks. */
int fecx_surrogate_freeze(fecx_ctx* ctx, int on);
/* on = 0: assemblies with a surrogate physics use the plain model (state and pattern untouched);
 * on = 1 (default): the surrogate runs. */
int fecx_surrogate_enable(fecx_ctx* ctx, int on);
/* Number of quadrature points that fell back to the plain model in the last surrogate pass. */
int fecx_surrogate_fallbacks(fecx_ctx* ctx, int64_t* n_fallback, int64_t* n_points);

/* Batched material update over n samples with the properties of the context's material:
 * kind FECX_J2 or FECX_NEOHOOKEAN runs the plain model, FECX_J2_SURROGATE or FECX_NH_SURROGATE the
 * surrogate with residual check and plain fallback. Raw inputs and outputs as the network's. */
int fecx_material_update(fecx_ctx* ctx, int kind, const float* in, int64_t n, float* out);

/* Displacement gradients at every quadrature point of the mesh at the full field U (device):
 * Hq (ne x nq x 9, element-major in the input element order, row-major H). Training samples for
 * the neo-Hookean surrogate come from here. */
int fecx_qp_gradients(fecx_ctx* ctx, const double* U, float* Hq);

/* J2 training samples from the plain model: npaths random macroscopic strain paths of `steps`
 * steps (4 random waypoints, norms max_norm * U(0.2, 1.1)); X (npaths*steps x 16) raw inputs and
 * T (x 13) raw outputs, device. Uses the context's J2 properties. */
int fecx_j2_generate_samples(fecx_ctx* ctx, int64_t npaths, int steps, float max_norm, uint64_t seed, float* X,
                             float* T);

/* ---------------------------------------------------------------- instrumentation */

/* Pipelining of matrix assemblies (no counterpart in the reference, whose CPU and
 * KernelAbstractions paths run the element loop and sparse! one after the other,
 * src/assemblers/SparsityPatterns.jl:301-329): the element kernel and the reduction run in
 * `chunks` Morton-ordered pieces on two streams, element piece j+1 overlapping reduction piece j.
 * 0 (the default) and 1 run them back to back as one piece. Results do not depend on it (same
 * sums, same order). */
int fecx_set_chunks(fecx_ctx* ctx, int chunks);

/* Launch form of the two-phase reduction on tet10 meshes with 3 fields (no counterpart in the reference, whose
 * sparse! walks the COO entries in sorted order, src/assemblers/SparsityPatterns.jl:301-329): 1 (the default)
 * reduces the edge-midpoint and the vertex rows in one launch, their work-groups interleaved along the Morton curve
 * so that the rows of the same elements read their blocks at about the same time (28 % fewer bytes read from HBM);
 * meshes with a row of more than 65 blocks (32 for an edge-midpoint row) take the two launches whatever the setting.
 * 0 runs the two row buckets as two launches. Results do not depend on it (same sums, same order). */
int fecx_set_merged_rows(fecx_ctx* ctx, int on);

/* Precision of the tet10 element chain (the reference evaluates its element loop in FP64 throughout,
 * src/assemblers/Assemblers.jl:402-432): 0 (the default) evaluates every element quantity in FP32; 1 evaluates the
 * isoparametric map, shape gradients, field gradient and stress of tet10 elements in FP64, from FP64 coordinate and
 * field differences, and passes their FP32 roundings to the tangent, the residual contraction and the element
 * outputs, which stay FP32 (the reduction sums them in FP64 as always). On jittered tet10 meshes the residual's
 * interior rows are cancelling sums of O(1) element contributions, so the FP32 error grows like eps / h; this keeps
 * every output within 1e-6 of FP64 at 80^3 x 6 elements. Other element kinds are unaffected. */
int fecx_set_fp64_geometry(fecx_ctx* ctx, int on);

/* Scatter design of matrix assemblies (no counterpart in the reference, whose element loop stores element-private
 * matrices and sparse! sums them afterwards, src/assemblers/Matrix.jl:35-75, SparsityPatterns.jl:301-329):
 *   FECX_SCATTER_AUTO       (default) FUSED where it is the faster design on MI355X (vector-valued materials, nf = 3),
 *                           TWO_PHASE elsewhere
 *   FECX_SCATTER_FUSED      on a one-block hex8 mesh numbered as StructuredMesh numbers it
 *                           (src/meshes/StructuredMesh.jl:85-131; detected by fecx_build_pattern), residual + Jacobian
 *                           and Jacobian-only assemblies with a plain material run as one kernel that sums each CSR
 *                           row from the element matrices in on-chip memory; the two-phase path elsewhere
 *   FECX_SCATTER_TWO_PHASE  element kernel into element-private storage, then the FP64 owner-computes reduction
 *   FECX_SCATTER_ATOMIC     zero fill, then FP64 atomic adds of every element block (structured hex8, Poisson and
 *                           neo-Hookean; the two-phase path elsewhere); summation order not deterministic
 *   FECX_SCATTER_COLOURED   zero fill, then 8 element colours whose elements share no node, plain adds
 *                           (same meshes as atomic)
 * The designs sum the same FP32 element contributions in FP64 in different orders, so they agree to FP64 rounding.
 * fecx_get_scatter_path reports the design (FUSED .. COLOURED) a residual + Jacobian call of the context takes. */
#define FECX_SCATTER_FUSED 0
#define FECX_SCATTER_TWO_PHASE 1
#define FECX_SCATTER_ATOMIC 2
#define FECX_SCATTER_COLOURED 3
#define FECX_SCATTER_AUTO 4
int fecx_set_scatter(fecx_ctx* ctx, int design);
int fecx_get_scatter_path(fecx_ctx* ctx, int* path);

/* When on, every kernel launch is bracketed by HIP events on the stream it runs on and its
 * duration accumulated per kernel name (the reference's only instrumentation is TimerOutputs
 * sections, src/Solvers.jl:133-152). */
int fecx_set_profiling(fecx_ctx* ctx, int on);

/* Number of distinct kernels profiled so far. */
int fecx_profile_count(fecx_ctx* ctx, int* count);

/* Entry i: kernel name (copied into name, NUL-terminated, at most name_len bytes), number of
 * launches and total milliseconds. */
int fecx_profile_entry(fecx_ctx* ctx, int i, char* name, int name_len, int64_t* launches, double* total_ms);

/* Clears the accumulated profile. */
int fecx_profile_reset(fecx_ctx* ctx);

/* HBM bandwidth ceiling of the context's device, for the rooflines the assembly kernels are priced against (no
 * counterpart in the reference): the average milliseconds of one streaming pass over the device buffer `buf` of
 * `bytes` (a multiple of 32, at least 4096) by a hand-written kernel, over `reps` passes after one warm-up, timed with
 * HIP events on the context's stream. op 0: read every byte; 1: write every byte; 2: copy the first half of the
 * buffer to the second (bytes / 2 read and bytes / 2 written). */
int fecx_measure_bandwidth(fecx_ctx* ctx, int op, void* buf, int64_t bytes, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* FECX_H */
