"""Complex GEMM kernels (matrix-core and generic) against numpy, for every
operand form the engine uses and ragged sizes; FP64 to 1e-12 relative,
FP32 to 1e-5 relative."""
import ctypes as ct

import numpy as np
import pytest

from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu

c_dp = ct.POINTER(ct.c_double)


def _run(prec, opA, opB, M, N, K, mfma, rng):
    lib = _abi.load_library()
    lib.yaqs_debug_gemm.argtypes = [ct.c_int] * 6 + [c_dp, c_dp, c_dp, ct.c_int]
    a_shape = (M, K) if opA == 0 else (K, M)
    b_shape = (K, N) if opB == 0 else (N, K)
    A = rng.standard_normal(a_shape) + 1j * rng.standard_normal(a_shape)
    B = rng.standard_normal(b_shape) + 1j * rng.standard_normal(b_shape)
    out = np.zeros(2 * M * N)
    a = np.ascontiguousarray(A).view(np.float64).ravel()
    b = np.ascontiguousarray(B).view(np.float64).ravel()
    rc = lib.yaqs_debug_gemm(prec, M, N, K, opA, opB, a.ctypes.data_as(c_dp), b.ctypes.data_as(c_dp),
                             out.ctypes.data_as(c_dp), mfma)
    assert rc == 0
    op = {0: lambda X: X, 1: lambda X: X.T, 2: lambda X: X.conj().T}
    ref = op[opA](A) @ op[opB](B)
    got = out.view(np.complex128).reshape(M, N)
    return np.abs(got - ref).max() / max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
@pytest.mark.parametrize("mfma", [1, 0])
def test_gemm_forms(prec, tol, mfma):
    print(f"case: gemm prec={prec} mfma={mfma}", flush=True)
    rng = np.random.default_rng(7)
    for opA, opB in [(a, b) for a in range(3) for b in range(3)]:
        for M, N, K in [(64, 64, 16), (70, 33, 21), (5, 130, 67), (320, 256, 64), (1, 1, 1)]:
            err = _run(prec, opA, opB, M, N, K, mfma, rng)
            assert err < tol, (opA, opB, M, N, K, err)
