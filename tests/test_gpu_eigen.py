"""Batched Hermitian eigensolver of the bond splits (Householder + implicit
QL, FP64) against numpy.linalg.eigh, on Gram matrices like those of the
2-site splits: full rank, rank-deficient, degenerate and zero-padded.  The
multi-kernel path (n <= 128 with at most 64 kept columns), the wide path
(n <= 512, at most 256 kept columns) and the single-workgroup solver
(everything else, inverse-iteration breakdowns) are covered; the residual
bound is ||G u - lambda u|| <= tol * ||G||."""
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


def _heev_batch(Gs, rule=0, thr=1e-9, max_bond=64, path=0):
    lib = _abi.load_library()
    f = lib.yaqs_debug_heev_batch
    f.argtypes = [ct.c_int, ct.c_int, c_dp, c_dp, c_dp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.c_int,
                  ct.c_double, ct.c_int, ct.c_int]
    cnt, n, _ = Gs.shape
    g = np.ascontiguousarray(Gs.astype(np.complex128)).view(np.float64).ravel()
    ev = np.zeros(cnt * n)
    U = np.zeros(2 * cnt * n * n)
    keep = np.zeros(cnt, np.int32)
    fb = np.zeros(cnt, np.int32)
    ip = ct.POINTER(ct.c_int)
    rc = f(n, cnt, g.ctypes.data_as(c_dp), ev.ctypes.data_as(c_dp), U.ctypes.data_as(c_dp), keep.ctypes.data_as(ip),
           fb.ctypes.data_as(ip), rule, thr, max_bond, path)
    assert rc == 0
    return ev.reshape(cnt, n), U.view(np.complex128).reshape(cnt, n, n), keep, fb


def _gram_batch(n, cnt, seed):
    """Gram matrices of 2-site-like tensors: decaying, rank-deficient,
    degenerate and zero-padded spectra, trace 1."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(cnt):
        kind = b % 5
        if kind == 0:
            X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
            X *= np.exp(-np.arange(n) / rng.uniform(3, 30))
        elif kind == 1:
            r = int(rng.integers(1, n + 1))
            X = rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))
        elif kind == 2:
            Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
            lam = np.repeat(np.array([0.5, 0.2, 0.05, 1e-7, 0.0]), n // 5 + 1)[:n]
            X = Q * np.sqrt(lam)
        elif kind == 3:
            X = np.zeros((n, n), complex)
            h = max(1, n // 2)
            X[:h, :h] = rng.standard_normal((h, h)) + 1j * rng.standard_normal((h, h))
        else:
            X = np.diag(np.linspace(1, 0, n)).astype(complex)
        G = X @ X.conj().T
        out.append(G / max(np.trace(G).real, 1e-300))
    return np.array(out)


@pytest.mark.parametrize("n,max_bond", [(2, 64), (7, 4), (33, 64), (64, 64), (100, 50), (128, 64), (128, 20)])
def test_batched_fast_path(n, max_bond):
    """Eigenvalues, kept columns and kept ranks of the multi-kernel path on a
    batch of 70 matrices (two 64-lane QL groups) against numpy and against
    the single-workgroup solver."""
    print(f"case: heev batch n={n} max_bond={max_bond}", flush=True)
    Gs = _gram_batch(n, 70, n)
    ev, U, keep, fb = _heev_batch(Gs, max_bond=max_bond)
    ev1, U1, keep1, _ = _heev_batch(Gs, max_bond=max_bond, path=1)
    assert fb.sum() == 0
    assert np.array_equal(keep, keep1)
    kout = min(n, max_bond)
    for b in range(len(Gs)):
        G = Gs[b]
        ref = np.sort(np.linalg.eigvalsh(G))[::-1]
        assert np.abs(ev[b] - ref).max() < 1e-13 * n
        assert np.abs(ev[b] - ev1[b]).max() < 1e-13 * n
        k = keep[b]  # the columns the split uses
        Uk = U[b][:, :k]
        assert np.abs(Uk.conj().T @ Uk - np.eye(k)).max() < 1e-12 * n
        assert np.abs(G @ Uk - Uk * ev[b][:k]).max() < 1e-13 * n
        # same kept subspace as the single-workgroup solver where the cut is in a gap
        if k < n and ev[b][k - 1] - ev[b][k] > 1e-6:
            Pk = U[b][:, :k] @ U[b][:, :k].conj().T
            P1 = U1[b][:, :k] @ U1[b][:, :k].conj().T
            assert np.abs(Pk - P1).max() < 1e-9


@pytest.mark.parametrize("n,max_bond", [(130, 64), (200, 100), (256, 128), (256, 256), (384, 128), (512, 256),
                                        (128, 128), (100, 90), (70, 65)])
def test_batched_wide_path(n, max_bond):
    """The global-memory path (128 < n <= 512, up to 256 kept columns, and
    n <= 128 with more than 64 kept) on a batch of 10 matrices of every
    spectrum kind, against numpy and against the single-workgroup solver."""
    print(f"case: heev wide batch n={n} max_bond={max_bond}", flush=True)
    Gs = _gram_batch(n, 10, n + 1)
    ev, U, keep, fb = _heev_batch(Gs, max_bond=max_bond)
    ev1, U1, keep1, _ = _heev_batch(Gs, max_bond=max_bond, path=1)
    assert fb.sum() == 0
    assert np.array_equal(keep, keep1)
    for b in range(len(Gs)):
        G = Gs[b]
        ref = np.sort(np.linalg.eigvalsh(G))[::-1]
        assert np.abs(ev[b] - ref).max() < 1e-13 * n
        k = keep[b]
        Uk = U[b][:, :k]
        assert np.abs(Uk.conj().T @ Uk - np.eye(k)).max() < 1e-12 * n
        assert np.abs(G @ Uk - Uk * ev[b][:k]).max() < 1e-13 * n
        if k < n and ev[b][k - 1] - ev[b][k] > 1e-6:
            Pk = U[b][:, :k] @ U[b][:, :k].conj().T
            P1 = U1[b][:, :k] @ U1[b][:, :k].conj().T
            assert np.abs(Pk - P1).max() < 1e-9


def _check(G, tol=1e-12):
    n = G.shape[0]
    ev, U, k = _heev(G)
    ref = np.sort(np.linalg.eigvalsh(G))[::-1]
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(ev - ref).max() < tol * scale * n
    U = U[:, :k]  # the kept columns
    assert np.abs(U.conj().T @ U - np.eye(k)).max() < 1e-11 * n
    assert np.abs(G @ U - U * ev[:k]).max() < tol * scale * n


@pytest.mark.parametrize("n", [1, 2, 3, 8, 33, 64, 100, 128, 160, 256, 512, 600, 1024])
def test_random_gram(n):
    # 512: the widest of the wide path (chi = 256 splits); 600 and 1024 (the
    # chi = 512 splits of the large Ising chain): the single-workgroup solver
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
