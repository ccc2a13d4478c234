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


def _tile_run(lib, prec, opA, M, N, K, tile, a, b, opB=0):
    out = np.zeros(2 * M * N)
    rc = lib.yaqs_debug_gemm_tile(prec, M, N, K, opA, opB, a.ctypes.data_as(c_dp), b.ctypes.data_as(c_dp),
                                  out.ctypes.data_as(c_dp), tile)
    assert rc == 0
    return out


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_tile_shapes_bit_identical(prec, tol):
    """Every C tile shape the tuner may pick gives the same bits as the
    default 64 x 64 tile (the same MFMA sequence per output element), and
    matches numpy."""
    print(f"case: gemm tile shapes prec={prec}", flush=True)
    lib = _abi.load_library()
    lib.yaqs_debug_gemm_tile.argtypes = [ct.c_int] * 6 + [c_dp, c_dp, c_dp, ct.c_int]
    rng = np.random.default_rng(11)
    op = {0: lambda X: X, 1: lambda X: X.T, 2: lambda X: X.conj().T}
    # tiles 4 and 5 (k tiles loaded straight into LDS) apply to complex-FP32,
    # op(A) = N and whole 64 x 64 tiles: (320, 256, 64), (64, 128, 48),
    # (128, 64, 320); elsewhere it runs as tile 0
    for opA in range(3):
        for M, N, K in [(80, 64, 16), (320, 256, 64), (33, 47, 70), (128, 20, 320), (1, 3, 5), (64, 128, 48),
                        (128, 64, 320)]:
            a_shape = (M, K) if opA == 0 else (K, M)
            A = rng.standard_normal(a_shape) + 1j * rng.standard_normal(a_shape)
            B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
            a = np.ascontiguousarray(A).view(np.float64).ravel()
            b = np.ascontiguousarray(B).view(np.float64).ravel()
            base = _tile_run(lib, prec, opA, M, N, K, 0, a, b)
            ref = op[opA](A) @ B
            got = base.view(np.complex128).reshape(M, N)
            assert np.abs(got - ref).max() / max(1.0, np.abs(ref).max()) < tol
            for tile in range(1, 6):
                other = _tile_run(lib, prec, opA, M, N, K, tile, a, b)
                assert np.array_equal(other, base), (opA, M, N, K, tile)
            # a conjugated right operand (the LQ side of the gauge steps'
            # projections): the 32 x 64, 64 x 32 and 32 x 32 tiles too
            Bc = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
            bc = np.ascontiguousarray(Bc).view(np.float64).ravel()
            base_c = _tile_run(lib, prec, opA, M, N, K, 0, a, bc, opB=2)
            ref_c = op[opA](A) @ Bc.conj().T
            got_c = base_c.view(np.complex128).reshape(M, N)
            assert np.abs(got_c - ref_c).max() / max(1.0, np.abs(ref_c).max()) < tol
            for tile in range(1, 4):
                other = _tile_run(lib, prec, opA, M, N, K, tile, a, bc, opB=2)
                assert np.array_equal(other, base_c), (opA, M, N, K, tile, "opB=C")


def _csr(Wd):
    rows, cols = np.nonzero(np.abs(Wd) > 0)
    rowptr = np.zeros(Wd.shape[0] + 1, dtype=np.int32)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    vals = np.ascontiguousarray(Wd[rows, cols]).view(np.float64).ravel()
    return rowptr, cols.astype(np.int32), vals


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_fused_mpo_step_matches_unfused(prec, tol):
    """GEMM with the MPO step in its epilogue = GEMM, then wstep, bit for bit
    (same FP64 accumulation order), with register-staged and with
    direct-to-LDS k tiles, and = numpy."""
    print(f"case: fused MPO step prec={prec}", flush=True)
    lib = _abi.load_library()
    ip = ct.POINTER(ct.c_int)
    lib.yaqs_debug_gemm_w.argtypes = [ct.c_int] * 7 + [c_dp, c_dp, ip, ip, c_dp, c_dp, c_dp, c_dp]
    rng = np.random.default_rng(5)
    n_dl = 0
    for P, Dl, dph, Dr, Q, K in [(64, 5, 2, 5, 64, 64), (13, 5, 4, 5, 37, 29), (64, 3, 2, 3, 128, 64),
                                 (7, 1, 2, 5, 3, 9), (30, 5, 4, 5, 64, 30), (20, 2, 2, 3, 50, 17),
                                 (33, 4, 4, 4, 40, 64), (32, 5, 4, 5, 32, 48), (16, 1, 4, 2, 16, 16),
                                 (64, 5, 4, 5, 64, 64)]:
        M, N, I, J = P * Dl, dph * Q, Dl * dph, dph * Dr
        A = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
        B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
        Wd = (rng.standard_normal((J, I)) + 1j * rng.standard_normal((J, I))) * (rng.random((J, I)) < 0.3)
        rowptr, col, val = _csr(Wd)
        a = np.ascontiguousarray(A).view(np.float64).ravel()
        b = np.ascontiguousarray(B).view(np.float64).ravel()
        yf = np.zeros(2 * P * J * Q)
        yu = np.zeros(2 * P * J * Q)
        yd = np.zeros(2 * P * J * Q)
        rc = lib.yaqs_debug_gemm_w(prec, P, Dl, dph, Dr, Q, K, a.ctypes.data_as(c_dp), b.ctypes.data_as(c_dp),
                                   rowptr.ctypes.data_as(ip), col.ctypes.data_as(ip), val.ctypes.data_as(c_dp),
                                   yf.ctypes.data_as(c_dp), yu.ctypes.data_as(c_dp), yd.ctypes.data_as(c_dp))
        assert rc & 1 == 0, (P, Dl, dph, Dr, Q, K)
        assert np.array_equal(yf, yu), (P, Dl, dph, Dr, Q, K)
        # k tiles straight into LDS: complex-FP32 with P % 16, Q % (64 / dph), K % 16 == 0
        fits = prec == 1 and P % 16 == 0 and Q % (64 // dph) == 0 and K % 16 == 0
        assert (rc & 4 == 0) == fits, (P, Dl, dph, Dr, Q, K, rc)
        if fits:
            n_dl += 1
            assert np.array_equal(yd, yf), (P, Dl, dph, Dr, Q, K)
        X = (A @ B).reshape(P, Dl, dph, Q).reshape(P, I, Q)
        ref = np.einsum("ji,piq->pjq", Wd, X)
        got = yf.view(np.complex128).reshape(P, J, Q)
        assert np.abs(got - ref).max() / max(1.0, np.abs(ref).max()) < tol, (P, Dl, dph, Dr, Q, K)
    assert n_dl == (5 if prec == 1 else 0)


@pytest.mark.parametrize("form", [0, 1])
def test_gram_fp64_matrix_cores(form):
    """FP64 Gram matrices of complex-FP32 operands on the FP64 matrix cores
    (the bond split's G = Y Y^H or Y^H Y): against numpy on the FP32-rounded
    operands (products of FP32 values are exact in FP64, so only the sum
    order differs: 1e-13 of |Y|^2) and against the FP64 generic kernel;
    with `lower` the tiles above the diagonal stay untouched (zero) and the
    lower triangle is the same."""
    print(f"case: gram form={form}", flush=True)
    lib = _abi.load_library()
    f = lib.yaqs_debug_gram
    f.argtypes = [ct.c_int] * 5 + [c_dp, c_dp, ct.c_int]
    rng = np.random.default_rng(11 + form)
    for M, K, cnt in [(256, 256, 3), (128, 128, 2), (100, 33, 2), (64, 200, 1), (17, 5, 2)]:
        shape = (cnt, M, K) if form == 0 else (cnt, K, M)
        Y = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64).astype(np.complex128)
        y = np.ascontiguousarray(Y).view(np.float64).ravel()
        ref = Y @ np.conj(np.transpose(Y, (0, 2, 1))) if form == 0 else np.conj(np.transpose(Y, (0, 2, 1))) @ Y
        scale = np.abs(Y).max() ** 2 * K
        out = {}
        for lower in (0, 1):
            for mf in (0, 1):
                g = np.zeros(2 * cnt * M * M)
                rc = f(M, K, form, lower, cnt, y.ctypes.data_as(c_dp), g.ctypes.data_as(c_dp), mf)
                assert rc == 0
                out[lower, mf] = g.view(np.complex128).reshape(cnt, M, M)
        full = out[0, 1]
        assert np.abs(full - ref).max() < 1e-13 * scale, (M, K)
        assert np.abs(full - out[0, 0]).max() < 1e-13 * scale, (M, K)
        low = out[1, 1]
        tril = np.tril(np.ones((M, M), bool))
        assert np.abs(low[:, tril] - full[:, tril]).max() == 0.0, (M, K)
        # tiles strictly above the diagonal (64 x 64) are not written
        r, c = np.indices((M, M))
        above = (c // 64) > (r // 64)
        assert np.abs(low[:, above]).max(initial=0.0) == 0.0, (M, K)
