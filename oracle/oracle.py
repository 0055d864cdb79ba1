"""ctypes wrapper of the FP64 oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates FiniteElementContainers.jl's assembly algorithm on the CPU in FP64
(oracle.cpp, file:line citations there). Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg may import this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

QUAD4, HEX8, TET10 = 0, 1, 2
POISSON, LINEAR_ELASTIC, NEOHOOKEAN, J2 = 0, 1, 2, 3

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_pattern_count.restype = ctypes.c_int64
        _lib.orc_sparse_csr.restype = ctypes.c_int64
        _lib.orc_sparse_csr_rows.restype = ctypes.c_int64
        _lib.orc_sparse_csc.restype = ctypes.c_int64
        _lib.orc_update_dofs.restype = ctypes.c_int64
        _lib.orc_cg.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _i64(x):
    return ctypes.c_int64(int(x))


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def structured_mesh(kind, mins, maxs, counts):
    """StructuredMesh(kind, mins, maxs, counts) -> (coords nd x NN col-major as (NN, nd), conn (NE, nnpe) 1-based)."""
    nd = len(counts)
    nn = int(np.prod(counts))
    ne = int(np.prod([c - 1 for c in counts]))
    nnpe = 4 if kind == QUAD4 else 8
    coords = np.zeros((nn, nd))
    conn = np.zeros((ne, nnpe), dtype=np.int64)
    rc = lib().orc_structured_mesh(ctypes.c_int(kind), _p(np.asarray(mins, dtype=np.float64)),
                                   _p(np.asarray(maxs, dtype=np.float64)),
                                   _p(np.asarray(counts, dtype=np.int64)), _p(coords), _p(conn))
    assert rc == 0
    return coords, conn


def update_dofs(ndofs, dirichlet):
    dirichlet = np.ascontiguousarray(dirichlet, dtype=np.int64)
    unknown = np.zeros(ndofs, dtype=np.int64)
    d2u = np.zeros(ndofs, dtype=np.int64)
    k = lib().orc_update_dofs(_i64(ndofs), _p(dirichlet), _i64(len(dirichlet)), _p(unknown), _p(d2u))
    return unknown[:k].copy(), d2u


def pattern(nf, conn, dof_to_unknown):
    """Reduced SparseMatrixPattern: Is, Js (unknown numbering), slots (1-based storage index), perm (1-based)."""
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    d2u = np.ascontiguousarray(dof_to_unknown, dtype=np.int64)
    ne, nnpe = conn.shape
    n = lib().orc_pattern_count(ctypes.c_int(nf), ctypes.c_int(nnpe), _i64(ne), _p(conn), _p(d2u))
    Is = np.zeros(n, dtype=np.int64)
    Js = np.zeros(n, dtype=np.int64)
    slots = np.zeros(n, dtype=np.int64)
    perm = np.zeros(n, dtype=np.int64)
    lib().orc_pattern(ctypes.c_int(nf), ctypes.c_int(nnpe), _i64(ne), _p(conn), _p(d2u), _p(Is), _p(Js), _p(slots),
                      _p(perm))
    return Is, Js, slots, perm


def sparse_csr(nrows, Is, Js, V=None, want_map=False):
    """sparse!(Is, Js, V, n, n, +) then SparseMatrixCSR: (rowptr, colval) 1-based, vals, COO->CSR map (0-based)."""
    Is = np.ascontiguousarray(Is, dtype=np.int64)
    Js = np.ascontiguousarray(Js, dtype=np.int64)
    Vv = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    n = len(Is)
    nnz = lib().orc_sparse_csr(_i64(nrows), _i64(n), _p(Is), _p(Js), None, None, None, None, None)
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    colval = np.zeros(nnz, dtype=np.int64)
    vals = np.zeros(nnz)
    cmap = np.zeros(n, dtype=np.int64) if want_map else None
    lib().orc_sparse_csr(_i64(nrows), _i64(n), _p(Is), _p(Js), _p(Vv), _p(rowptr), _p(colval), _p(vals), _p(cmap))
    return rowptr, colval, vals, cmap


def sparse_csc(nrows, Is, Js, V=None):
    """sparse!(Is, Js, V, n, n, +): the CSC of sparse_matrix_type = :csc, (colptr, rowval) 1-based, nzval."""
    Is = np.ascontiguousarray(Is, dtype=np.int64)
    Js = np.ascontiguousarray(Js, dtype=np.int64)
    Vv = None if V is None else np.ascontiguousarray(V, dtype=np.float64)
    n = len(Is)
    nnz = lib().orc_sparse_csc(_i64(nrows), _i64(n), _p(Is), _p(Js), None, None, None, None)
    colptr = np.zeros(nrows + 1, dtype=np.int64)
    rowval = np.zeros(nnz, dtype=np.int64)
    nzval = np.zeros(nnz)
    lib().orc_sparse_csc(_i64(nrows), _i64(n), _p(Is), _p(Js), _p(Vv), _p(colptr), _p(rowval), _p(nzval))
    return colptr, rowval, nzval


def constraint_storage(ndofs, dirichlet):
    """The condensed assembler's constraint_storage: 1.0 at the Dirichlet dofs (1-based), 0.0 elsewhere."""
    c = np.zeros(ndofs)
    c[np.asarray(dirichlet, dtype=np.int64) - 1] = 1.0
    return c


def adjust_matrix(ptr, idx, nz, c, penalty_scale=1e6):
    """Condensed-mode matrix adjustment in place on a compressed (CSR rows or CSC columns) 1-based matrix."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    assert nz.dtype == np.float64 and nz.flags.c_contiguous
    c = np.ascontiguousarray(c, dtype=np.float64)
    lib().orc_adjust_matrix(_i64(len(ptr) - 1), _p(ptr), _p(idx), _p(nz), _p(c), ctypes.c_double(penalty_scale))
    return nz


def adjust_vector(b, c):
    assert b.dtype == np.float64 and b.flags.c_contiguous
    lib().orc_adjust_vector(_i64(len(b)), _p(b), _p(np.ascontiguousarray(c, dtype=np.float64)))
    return b


def adjust_action(Av, c, v):
    assert Av.dtype == np.float64 and Av.flags.c_contiguous
    lib().orc_adjust_action(_i64(len(Av)), _p(Av), _p(np.ascontiguousarray(c, dtype=np.float64)),
                            _p(np.ascontiguousarray(v, dtype=np.float64)))
    return Av


def sparse_csr_rows(nf, conn, dof_to_unknown, n_unknowns, storage):
    """The CSR of the element storage, as pattern -> gather_slots -> sparse_csr gives it (same pattern, same bits),
    formed row by row in parallel without COO arrays: (rowptr, colval) 1-based, vals."""
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    d2u = np.ascontiguousarray(dof_to_unknown, dtype=np.int64)
    storage = np.ascontiguousarray(storage, dtype=np.float64)
    ne, nnpe = conn.shape
    nn = len(d2u) // nf
    rowptr = np.zeros(n_unknowns + 1, dtype=np.int64)
    args = (ctypes.c_int(nf), ctypes.c_int(nnpe), _i64(nn), _i64(ne), _p(conn), _p(d2u), _i64(n_unknowns),
            _p(storage), _p(rowptr))
    nnz = lib().orc_sparse_csr_rows(*args, None, None)
    colval = np.zeros(nnz, dtype=np.int64)
    vals = np.zeros(nnz)
    lib().orc_sparse_csr_rows(*args, _p(colval), _p(vals))
    return rowptr, colval, vals


def _common(elem_kind, X, conn, nf, phys, props):
    X = np.ascontiguousarray(X, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    props = np.ascontiguousarray(props, dtype=np.float64)
    nn, nd = X.shape
    ne = conn.shape[0]
    return X, conn, props, nn, nd, ne


def assemble_residual(elem_kind, X, conn, nf, phys, props, U, state_old=None, state_new=None):
    X, conn, props, nn, nd, ne = _common(elem_kind, X, conn, nf, phys, props)
    U = np.ascontiguousarray(U, dtype=np.float64)
    R = np.zeros(nn * nf)
    lib().orc_assemble_residual(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                                ctypes.c_int(nf), ctypes.c_int(phys), _p(props), ctypes.c_int(len(props)), _p(U),
                                _p(state_old), _p(state_new), _p(R))
    return R


def assemble_stiffness_storage(elem_kind, X, conn, nf, phys, props, U, state_old=None):
    X, conn, props, nn, nd, ne = _common(elem_kind, X, conn, nf, phys, props)
    U = np.ascontiguousarray(U, dtype=np.float64)
    ndof = conn.shape[1] * nf
    st = np.zeros(ne * ndof * ndof)
    lib().orc_assemble_stiffness_storage(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne),
                                         _p(conn), ctypes.c_int(nf), ctypes.c_int(phys), _p(props),
                                         ctypes.c_int(len(props)), _p(U), _p(state_old), _p(st))
    return st


def assemble_action(elem_kind, X, conn, nf, phys, props, U, V, state_old=None, matrix_free=False):
    X, conn, props, nn, nd, ne = _common(elem_kind, X, conn, nf, phys, props)
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    KV = np.zeros(nn * nf)
    lib().orc_assemble_action(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                              ctypes.c_int(nf), ctypes.c_int(phys), _p(props), ctypes.c_int(len(props)), _p(U), _p(V),
                              _p(state_old), ctypes.c_int(1 if matrix_free else 0), _p(KV))
    return KV


def assemble_mass_storage(elem_kind, X, conn, nf, rho):
    X, conn, _, nn, nd, ne = _common(elem_kind, X, conn, nf, 0, [])
    ndof = conn.shape[1] * nf
    st = np.zeros(ne * ndof * ndof)
    lib().orc_assemble_mass_storage(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                                    ctypes.c_int(nf), ctypes.c_double(rho), _p(st))
    return st


def assemble_lumped_mass(elem_kind, X, conn, nf, rho):
    X, conn, _, nn, nd, ne = _common(elem_kind, X, conn, nf, 0, [])
    M = np.zeros(nn * nf)
    lib().orc_assemble_lumped_mass(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                                   ctypes.c_int(nf), ctypes.c_double(rho), _p(M))
    return M


def assemble_diagonal(elem_kind, X, conn, nf, phys, props, U, state_old=None, which="stiffness", rho=1.0):
    """Full-field element diagonals of the stiffness (at U) or of the consistent mass (density rho)."""
    X, conn, props, nn, nd, ne = _common(elem_kind, X, conn, nf, phys, props)
    U = np.ascontiguousarray(U if U is not None else np.zeros(nn * nf), dtype=np.float64)
    Dg = np.zeros(nn * nf)
    lib().orc_assemble_diagonal(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                                ctypes.c_int(nf), ctypes.c_int(phys), _p(props), ctypes.c_int(len(props)), _p(U),
                                _p(state_old), ctypes.c_int(1 if which == "mass" else 0), ctypes.c_double(rho),
                                _p(Dg))
    return Dg


def assemble_mass_action(elem_kind, X, conn, nf, rho, V):
    X, conn, _, nn, nd, ne = _common(elem_kind, X, conn, nf, 0, [])
    V = np.ascontiguousarray(V, dtype=np.float64)
    MV = np.zeros(nn * nf)
    lib().orc_assemble_mass_action(ctypes.c_int(elem_kind), ctypes.c_int(nd), _i64(nn), _p(X), _i64(ne), _p(conn),
                                   ctypes.c_int(nf), ctypes.c_double(rho), _p(V), _p(MV))
    return MV


def mass_csr(elem_kind, X, conn, nf, rho, dof_to_unknown, n_unknowns):
    """The reference's mass(asm) for the CSR type: consistent mass storage -> sparse! -> CSR."""
    Is, Js, slots, perm = pattern(nf, conn, dof_to_unknown)
    st = assemble_mass_storage(elem_kind, X, conn, nf, rho)
    V = gather_slots(slots, st)
    rowptr, colval, vals, _ = sparse_csr(n_unknowns, Is, Js, V)
    return rowptr, colval, vals


def gather_slots(slots, storage):
    slots = np.ascontiguousarray(slots, dtype=np.int64)
    out = np.zeros(len(slots))
    lib().orc_gather_slots(_i64(len(slots)), _p(slots), _p(np.ascontiguousarray(storage)), _p(out))
    return out


def stiffness_csr(elem_kind, X, conn, nf, phys, props, U, dof_to_unknown, n_unknowns, state_old=None):
    """The reference's stiffness(asm) on CPU for the CSR type: element storage -> sparse! -> CSR."""
    Is, Js, slots, perm = pattern(nf, conn, dof_to_unknown)
    st = assemble_stiffness_storage(elem_kind, X, conn, nf, phys, props, U, state_old)
    V = gather_slots(slots, st)
    rowptr, colval, vals, _ = sparse_csr(n_unknowns, Is, Js, V)
    return rowptr, colval, vals


def material_point(phys, props, nd, nf, grad, Xq=None, state_old=None):
    props = np.ascontiguousarray(props, dtype=np.float64)
    grad = np.ascontiguousarray(grad, dtype=np.float64).ravel()
    Xq = np.zeros(3) if Xq is None else np.ascontiguousarray(Xq, dtype=np.float64)
    D = nf * nd
    P = np.zeros(D)
    src = np.zeros(nf)
    A = np.zeros(D * D)
    st_new = np.zeros(7) if phys == J2 else None
    so = None if state_old is None else np.ascontiguousarray(state_old, dtype=np.float64)
    if phys == J2 and so is None:
        so = np.zeros(7)
    lib().orc_material_point(ctypes.c_int(phys), _p(props), ctypes.c_int(len(props)), ctypes.c_int(nd),
                             ctypes.c_int(nf), _p(grad), _p(Xq), _p(so), _p(st_new), _p(P), _p(src), _p(A))
    return P, src, A.reshape(D, D), st_new


def ref_element(kind):
    nnpe, nd = {QUAD4: (4, 2), HEX8: (8, 3), TET10: (10, 3)}[kind]
    xi = np.zeros((8, 3))
    w = np.zeros(8)
    N = np.zeros((8, nnpe))
    dN = np.zeros((8, nnpe, nd))
    nq = lib().orc_ref_element(ctypes.c_int(kind), _p(xi), _p(w), _p(N), _p(dN))
    return xi[:nq], w[:nq], N[:nq], dN[:nq]


def cg(rowptr, colval, vals, b, x0=None, rtol=1e-12, maxit=10000):
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    rel = ctypes.c_double(0)
    it = lib().orc_cg(_i64(n), _p(np.ascontiguousarray(rowptr, dtype=np.int64)),
                      _p(np.ascontiguousarray(colval, dtype=np.int64)), _p(np.ascontiguousarray(vals)),
                      _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x), ctypes.c_double(rtol), ctypes.c_int(maxit),
                      ctypes.byref(rel))
    return x, it, rel.value
