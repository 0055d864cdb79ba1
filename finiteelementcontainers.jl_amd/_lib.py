"""ctypes binding of libfecx.so (the C ABI declared in include/fecx.h).

The library is the only compute path: if it is missing or fails to load, every call raises
immediately. There is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfecx.so")

QUAD4, HEX8, TET10 = 0, 1, 2
POISSON, LINEAR_ELASTIC, NEOHOOKEAN, J2, J2_SURROGATE, NH_SURROGATE = 0, 1, 2, 3, 4, 5
MLP_J2, MLP_NH = 0, 1

ERRORS = {
    -1: "FECX_ERR_ARG",
    -2: "FECX_ERR_STATE",
    -3: "FECX_ERR_HIP",
    -4: "FECX_ERR_UNSUPPORTED",
    -5: "FECX_ERR_NUMERIC",
}


class FECXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

_SIGS = {
    "fecx_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int],
    "fecx_destroy": [ctypes.c_void_p],
    "fecx_set_stream": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_set_mesh": [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                      ctypes.c_void_p],
    "fecx_set_dofs": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64],
    "fecx_get_unknown_dofs": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p],
    "fecx_build_pattern": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)],
    "fecx_get_pattern": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
    "fecx_get_coo_map": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_set_material": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int],
    "fecx_get_state_size": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "fecx_set_state_old": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_get_state_new": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_state_commit": [ctypes.c_void_p],
    "fecx_assemble_vector": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_matrix": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_vector_matrix": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_matrix_action": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_matrix_free_action": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_diagonal": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_assemble_mass": [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p],
    "fecx_assemble_lumped_mass": [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p],
    "fecx_assemble_mass_diagonal": [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p],
    "fecx_assemble_mass_action": [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_update_field_unknowns": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_update_field_dirichlet": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_extract_field_unknowns": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_csr_spmv": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_csr_diagonal_inverse": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_cg_solve": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                      ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double)],
    "fecx_vector_norm": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)],
    "fecx_vector_axpy": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_surrogate_init": [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64],
    "fecx_surrogate_num_params": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)],
    "fecx_surrogate_get_params": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_surrogate_set_params": [ctypes.c_void_p, ctypes.c_void_p],
    "fecx_surrogate_set_scaling": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_surrogate_get_scaling": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_surrogate_fit_scaling": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64],
    "fecx_surrogate_forward": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p],
    "fecx_surrogate_gradient": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_double), ctypes.c_void_p],
    "fecx_surrogate_train": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                             ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p],
    "fecx_surrogate_set_tolerance": [ctypes.c_void_p, ctypes.c_float],
    "fecx_surrogate_freeze": [ctypes.c_void_p, ctypes.c_int],
    "fecx_surrogate_enable": [ctypes.c_void_p, ctypes.c_int],
    "fecx_surrogate_fallbacks": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "fecx_material_update": [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p],
    "fecx_qp_gradients": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "fecx_j2_generate_samples": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                 ctypes.c_void_p, ctypes.c_void_p],
    "fecx_set_chunks": [ctypes.c_void_p, ctypes.c_int],
    "fecx_set_merged_rows": [ctypes.c_void_p, ctypes.c_int],
    "fecx_set_fp64_geometry": [ctypes.c_void_p, ctypes.c_int],
    "fecx_set_profiling": [ctypes.c_void_p, ctypes.c_int],
    "fecx_profile_count": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "fecx_profile_entry": [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)],
    "fecx_profile_reset": [ctypes.c_void_p],
}


def exported_symbols():
    """Names the binding expects libfecx.so to export (all entry points of include/fecx.h)."""
    return list(_SIGS) + ["fecx_last_error"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FECXError(-3, f"libfecx.so not built at {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.fecx_last_error.argtypes = [ctypes.c_void_p]
        L.fecx_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


class Context:
    """Owns one fecx_ctx (one element block on one device)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        rc = lib().fecx_create(ctypes.byref(self._h), int(device))
        if rc != 0:
            raise FECXError(rc, f"fecx_create failed on device {device}")

    def call(self, name, *args):
        rc = getattr(lib(), name)(self._h, *args)
        if rc != 0:
            raise FECXError(rc, f"{name}: {lib().fecx_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            lib().fecx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
