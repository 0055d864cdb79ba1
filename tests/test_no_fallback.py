"""The HIP library is the only compute path: without libfecx.so, or without a HIP device, the assembler refuses to run
instead of falling back to a CPU implementation."""
import numpy as np
import pytest
import torch


def test_missing_library_raises(fec, monkeypatch):
    import sys
    L = sys.modules[fec.__name__ + "._lib"]
    monkeypatch.setattr(L, "LIB_PATH", L.LIB_PATH + ".missing")
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(fec.FECXError) as ei:
        L.Context(-1)
    assert ei.value.code == -3 and "not built" in str(ei.value)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the refusal on a host without a HIP device")
def test_assembler_without_gpu_raises(fec):
    mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (3, 3, 3))
    var = fec.ScalarFunction(fec.FunctionSpace(mesh, "H1Field", "Lagrange"), "u")
    with pytest.raises(fec.FECXError) as ei:
        fec.SparseMatrixAssembler(var, sparse_matrix_type="csr")
    assert ei.value.code == -3


def test_host_context_refuses_kernels(fec):
    """A host-only context (device -1) builds patterns but refuses every assembly call."""
    mesh = fec.StructuredMesh("hex", (0, 0, 0), (1, 1, 1), (3, 3, 3))
    hp = fec.HostPattern(mesh, 1, np.zeros(0, np.int64))
    U, R = np.zeros(27), np.zeros(27)  # host arrays: the call is refused before any pointer is used
    with pytest.raises(fec.FECXError) as ei:
        hp.ctx.call("fecx_assemble_vector", U.ctypes.data_as(__import__("ctypes").c_void_p),
                    R.ctypes.data_as(__import__("ctypes").c_void_p))
    assert "host-only" in str(ei.value)
