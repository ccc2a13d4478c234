"""Pins the oracle's dense kernels against numpy/scipy, closed forms and the
Philox known-answer vectors (CPU only)."""
import ctypes as ct

import numpy as np
import pytest
import scipy.linalg as sla

rng = np.random.default_rng(1234)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m,n", [(8, 8), (12, 5), (5, 12), (32, 32), (2, 2), (1, 4)])
def test_svd_matches_numpy(oracle_lib, m, n):
    A = crand(m, n)
    U, S, Vt = oracle_lib.svd(A)
    r = min(m, n)
    assert np.allclose(U @ np.diag(S) @ Vt, A, atol=1e-12)
    assert np.allclose(S, np.linalg.svd(A, compute_uv=False), atol=1e-12)
    assert np.allclose(U.conj().T @ U, np.eye(r), atol=1e-12)
    assert np.allclose(Vt @ Vt.conj().T, np.eye(r), atol=1e-12)
    assert np.all(np.diff(S) <= 1e-15)


def test_svd_rank_deficient_has_orthonormal_u(oracle_lib):
    a, b = crand(6, 1), crand(1, 6)
    U, S, Vt = oracle_lib.svd(a @ b)
    assert S[1] < 1e-12 and abs(S[0] - np.linalg.norm(a) * np.linalg.norm(b)) < 1e-12
    assert np.allclose(U.conj().T @ U, np.eye(6), atol=1e-10)


@pytest.mark.parametrize("m,n", [(8, 4), (4, 8), (6, 6), (128, 64)])
def test_qr_lq(oracle_lib, m, n):
    A = crand(m, n)
    Q, R = oracle_lib.qr(A)
    assert np.allclose(Q @ R, A, atol=1e-12)
    assert np.allclose(Q.conj().T @ Q, np.eye(Q.shape[1]), atol=1e-12)
    assert np.abs(np.tril(R, -1)).max(initial=0) == 0
    Lm, Q2 = oracle_lib.lq(A)
    assert np.allclose(Lm @ Q2, A, atol=1e-12)
    assert np.allclose(Q2 @ Q2.conj().T, np.eye(Q2.shape[0]), atol=1e-12)
    assert np.abs(np.triu(Lm, 1)).max(initial=0) == 0


def test_qr_lq_against_lapack(oracle_lib):
    """Full rank: the thin factors equal numpy's LAPACK ones up to the phase
    of each column of Q (rows of the LQ factor), the freedom of the QR."""
    for m, n in ((12, 5), (5, 12), (9, 9)):
        A = crand(m, n)
        Q, R = oracle_lib.qr(A)
        Qn, Rn = np.linalg.qr(A)
        ph = np.diag(Qn.conj().T @ Q)
        assert np.allclose(np.abs(ph), 1, atol=1e-12)
        assert np.allclose(Q, Qn * ph, atol=1e-12) and np.allclose(R, ph.conj()[:, None] * Rn, atol=1e-12)


def test_qr_lq_dead_directions(oracle_lib):
    """Exactly dependent columns (rows): LAPACK completes Q from a reflector
    of rounding noise (numpy's Q column differs under a 1e-15 change of A),
    the oracle drops the direction: the next live column takes its slot, the
    unused last slot is a zero column of Q and a zero row of R, the live
    columns orthonormal, Q R = A (a residual below 1e-12 of ||A||_F is
    dead).  With more columns than rows a live direction after a dead column
    still finds a slot (Q R = A for a rank-4 4 x 6 matrix whose second column
    repeats the first)."""
    a = crand(10, 3)
    A = np.c_[a[:, :2], a[:, 0] * (0.3 - 0.2j) + a[:, 1] * 0.7, a[:, 2:]]  # column 2 dependent
    Q, R = oracle_lib.qr(A)
    assert np.abs(Q[:, 3]).max() == 0 and np.abs(R[3]).max() == 0
    assert np.allclose(Q[:, :3].conj().T @ Q[:, :3], np.eye(3), atol=1e-12)
    assert np.allclose(Q @ R, A, atol=1e-12)
    # LAPACK's completion of the dead column follows the rounding noise
    Qn0, _ = np.linalg.qr(A)
    Qn1, _ = np.linalg.qr(A + 1e-15 * crand(10, 4))
    assert np.abs(np.abs(np.vdot(Qn0[:, 2], Qn1[:, 2])) - 1) > 1e-6
    Lm, Q2 = oracle_lib.lq(A.conj().T)
    assert np.abs(Q2[3]).max() == 0 and np.abs(Lm[:, 3]).max() == 0
    assert np.allclose(Lm @ Q2, A.conj().T, atol=1e-12)
    b = crand(4, 5)
    B = np.c_[b[:, :1], 2 * b[:, :1], b[:, 1:]]  # rank 4, dead column 1
    Q, R = oracle_lib.qr(B)
    assert np.allclose(Q @ R, B, atol=1e-12)
    assert np.allclose(Q.conj().T @ Q, np.eye(4), atol=1e-12)


def test_tridiagonal_exponential(oracle_lib):
    for k in (1, 2, 7, 25):
        a = rng.standard_normal(k)
        b = np.r_[0.0, rng.standard_normal(k - 1)]
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        for tau in (0.05, -0.3, 2.0):
            c = oracle_lib.tridiag_exp(a, b, tau)
            assert np.allclose(c, sla.expm(-1j * tau * T)[:, 0], atol=1e-13)


def test_krylov_closed_forms(oracle_lib):
    """test/test_Algorithms.jl:17-35: exp(-iZdt)|0>, exp(-iXdt)|0>."""
    Zm = np.diag([1.0, -1.0]).astype(complex)
    Xm = np.array([[0, 1], [1, 0]], complex)
    v0 = np.array([1, 0], complex)
    out, _ = oracle_lib.expm_krylov_dense(Zm, v0, 0.5, 5)
    assert np.allclose(out, v0 * np.exp(-0.5j), atol=1e-10)
    out, _ = oracle_lib.expm_krylov_dense(Xm, v0, 0.5, 5)
    assert np.allclose(out, [np.cos(0.5), -1j * np.sin(0.5)], atol=1e-10)
    out, it = oracle_lib.expm_krylov_dense(Zm, np.zeros(2, complex), 0.5, 5)
    assert it == 0 and np.all(out == 0)  # ‖v‖ = 0 returns v (src/Algorithms.jl:145-148)


@pytest.mark.parametrize("n,dt", [(16, 0.05), (64, 0.025), (256, -0.025), (200, 0.1)])
def test_krylov_vs_scipy(oracle_lib, n, dt):
    H = crand(n, n)
    H = (H + H.conj().T) / 2
    v = crand(n)
    v /= np.linalg.norm(v)
    out, it = oracle_lib.expm_krylov_dense(H, v, dt, 25, 1e-10)
    assert it <= 25
    assert np.abs(out - sla.expm(-1j * dt * H) @ v).max() < 1e-10


def test_effective_hamiltonian_and_envs(oracle_lib):
    """project_site / update_{left,right}_environment index contractions
    (src/Algorithms.jl:199-220) against einsum restatements."""
    kl, p, kr, ml, mr = 5, 4, 3, 3, 5
    x, Le, W, Re = crand(kl, p, kr), crand(kl, ml, kl), crand(ml, p, p, mr), crand(kr, mr, kr)
    y = oracle_lib.project_site(x, Le, W, Re)
    assert np.allclose(y, np.einsum("amk,kpr,mqpn,bnr->aqb", Le, x, W, Re), atol=1e-11)
    A = crand(kl, 2, kr)
    W1 = crand(ml, 2, 2, mr)
    El = oracle_lib.env_left(A, W1, Le)
    assert np.allclose(El, np.einsum("amk,kpr,mqpn,aqb->bnr", Le, A, W1, A.conj()), atol=1e-11)
    Er = oracle_lib.env_right(A, W1, Re)
    assert np.allclose(Er, np.einsum("kpr,bnr,mqpn,aqb->amk", A, Re, W1, A.conj()), atol=1e-11)


def _philox(lib, c, k):
    c = (ct.c_uint32 * 4)(*c)
    k = (ct.c_uint32 * 2)(*k)
    o = (ct.c_uint32 * 4)()
    lib.oracle_philox_raw(c, k, o)
    return list(o)


def test_philox_known_answers(oracle_lib):
    """Random123 known-answer vectors for philox4x32-10."""
    L = oracle_lib.lib()
    assert _philox(L, [0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _philox(L, [0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _philox(L, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_rng_uniform_contract(oracle_lib):
    u = np.array([oracle_lib.rng_uniform(7, t, k) for t in range(50) for k in range(40)])
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.03
    # pure function of (seed, traj, k)
    assert oracle_lib.rng_uniform(7, 3, 5) == oracle_lib.rng_uniform(7, 3, 5)
    assert oracle_lib.rng_uniform(7, 3, 5) != oracle_lib.rng_uniform(8, 3, 5)
    # top 53 bits of the first two words of the block with counter (k, 0, t, 0)
    w = _philox(oracle_lib.lib(), [5, 0, 3, 0], [7, 0])
    bits = ((w[0] >> 5) << 26) | (w[1] >> 6)
    assert oracle_lib.rng_uniform(7, 3, 5) == bits / 2.0 ** 53
