"""Blocked CGS2 orthonormalisation of the gauge steps (QR of (2 chi, chi)
column blocks, LQ of (chi, 2 chi) row blocks) against its definition:
A = Q Cf^T (column k of A is sum_j Cf[k][j] q_j), Cf lower triangular with a
real non-negative diagonal, live vectors orthonormal, dead directions
(remaining norm <= 1e-12 ||A||_F) exactly zero.  Shapes cross the 32-vector
panels; FP64 to 1e-12, FP32 to 2e-5 (relative to ||A||)."""
import ctypes as ct

import numpy as np
import pytest

from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu
c_dp = ct.POINTER(ct.c_double)


def _cgs2(mats, lq, prec):
    lib = _abi.load_library()
    lib.yaqs_debug_cgs2.argtypes = [ct.c_int] * 5 + [c_dp, c_dp]
    cnt = len(mats)
    if lq:
        nvec, ln = mats[0].shape
    else:
        ln, nvec = mats[0].shape
    a = np.ascontiguousarray(np.array(mats).astype(np.complex128)).view(np.float64).ravel().copy()
    c = np.zeros(2 * cnt * nvec * nvec)
    rc = lib.yaqs_debug_cgs2(prec, nvec, ln, lq, cnt, a.ctypes.data_as(c_dp), c.ctypes.data_as(c_dp))
    assert rc == 0
    Q = a.view(np.complex128).reshape((cnt,) + mats[0].shape)
    C = c.view(np.complex128).reshape(cnt, nvec, nvec)
    return Q, C


def _mats(nvec, ln, lq, rng):
    shape = (nvec, ln) if lq else (ln, nvec)
    full = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    # rank-deficient: a few vectors repeated; zero-padded: trailing vectors zero
    dup = full.copy()
    if lq:
        dup[nvec // 2:] = dup[: nvec - nvec // 2]
    else:
        dup[:, nvec // 2:] = dup[:, : nvec - nvec // 2]
    pad = full.copy()
    if lq:
        pad[max(1, nvec - 5):] = 0
    else:
        pad[:, max(1, nvec - 5):] = 0
    f = np.exp(-np.arange(nvec) / 12.0)
    decay = full * (f[:, None] if lq else f[None, :])
    return [full, dup, pad, decay]


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
@pytest.mark.parametrize("lq", [0, 1])
def test_cgs2_blocked(prec, tol, lq):
    print(f"case: cgs2 prec={prec} lq={lq}", flush=True)
    rng = np.random.default_rng(3 + lq)
    for nvec, ln in [(64, 128), (128, 256), (33, 70), (5, 300), (96, 96), (1, 4)]:
        mats = _mats(nvec, ln, lq, rng)
        if prec == 1:
            # an exactly repeated vector leaves a residual at FP32 rounding
            # level, above the dead threshold: only FP64 takes it as dead
            mats = [mats[0], mats[2], mats[3]]
        Q, C = _cgs2(mats, lq, prec)
        for A, q, c in zip(mats, Q, C):
            V = q if lq else q.T            # rows = vectors
            Av = A if lq else A.T
            scale = max(1.0, np.linalg.norm(A))
            assert np.abs(np.triu(c, 1)).max() == 0.0
            assert np.abs(np.diag(c).imag).max() == 0.0 and np.diag(c).real.min() >= 0.0
            # a_k = sum_j Cf[k][j] q_j
            assert np.abs(Av - c @ V).max() < tol * scale * 10, (nvec, ln)
            live = np.diag(c).real > 0
            G = V[live].conj() @ V[live].T
            assert np.abs(G - np.eye(live.sum())).max() < tol * 10 * max(1, nvec / 32), (nvec, ln)
            assert np.abs(V[~live]).max(initial=0.0) == 0.0
