"""Batched Hermitian eigensolver of the bond splits (Householder + implicit
QL, FP64) against numpy.linalg.eigh, on Gram matrices like those of the
2-site splits: full rank, rank-deficient, degenerate and zero-padded."""
import ctypes as ct

import numpy as np
import pytest

from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu
c_dp = ct.POINTER(ct.c_double)


def _heev(G, rule=0, thr=1e-9, max_bond=10 ** 6):
    lib = _abi.load_library()
    lib.yaqs_debug_heev.argtypes = [ct.c_int, c_dp, c_dp, c_dp, ct.POINTER(ct.c_int), ct.c_int, ct.c_double, ct.c_int]
    n = G.shape[0]
    g = np.ascontiguousarray(G.astype(np.complex128)).view(np.float64).ravel()
    ev = np.zeros(n)
    U = np.zeros(2 * n * n)
    keep = ct.c_int(0)
    rc = lib.yaqs_debug_heev(n, g.ctypes.data_as(c_dp), ev.ctypes.data_as(c_dp), U.ctypes.data_as(c_dp),
                             ct.byref(keep), rule, thr, max_bond)
    assert rc == 0
    return ev, U.view(np.complex128).reshape(n, n), keep.value


def _check(G, tol=1e-12):
    n = G.shape[0]
    ev, U, _ = _heev(G)
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(ev - ref).max() < tol * scale * n
    assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-11 * n
    assert np.abs(G @ U - U * ev).max() < tol * scale * n


@pytest.mark.parametrize("n", [1, 2, 3, 8, 33, 64, 100, 128, 160, 256])
def test_random_gram(n):
    print(f"case: heev n={n}", flush=True)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    _check(X @ X.conj().T / n)


def test_rank_deficient_degenerate_and_padded():
    print("case: heev special spectra", flush=True)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((96, 20)) + 1j * rng.standard_normal((96, 20))
    _check(X @ X.conj().T)                               # rank 20
    Q, _ = np.linalg.qr(rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64)))
    lam = np.repeat([1.0, 0.5, 0.25, 1e-6], 16)
    _check((Q * lam) @ Q.conj().T)                       # 16-fold degenerate eigenvalues
    P = np.zeros((128, 128), complex)
    Y = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    P[:40, :40] = Y @ Y.conj().T
    _check(P)                                            # zero padding
    _check(np.diag(np.linspace(2, 0, 50)).astype(complex))  # already diagonal


def test_truncation_rules():
    """T-A keeps max(k, 2) with absolute discarded weight >= thr; T-C is
    relative; T-D drops while the cumulative tail stays <= thr."""
    print("case: heev truncation rules", flush=True)
    s2 = np.array([0.6, 0.3, 0.09, 0.009, 1e-10, 4e-11, 0.0, 0.0])
    G = np.diag(s2).astype(complex)
    _, _, ka = _heev(G, rule=0, thr=1e-9)
    _, _, kc = _heev(G, rule=1, thr=1e-9)
    _, _, kd = _heev(G, rule=2, thr=1e-9)
    assert ka == 4 and kd == 4 and kc == 4
    _, _, ka2 = _heev(np.diag([1.0, 0, 0, 0]).astype(complex), rule=0, thr=1e-9)
    _, _, kd2 = _heev(np.diag([1.0, 0, 0, 0]).astype(complex), rule=2, thr=1e-9)
    assert ka2 == 2 and kd2 == 1
    _, _, km = _heev(G, rule=0, thr=1e-12, max_bond=3)
    assert km == 3
