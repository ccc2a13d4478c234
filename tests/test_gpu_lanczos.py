"""The device tridiagonal exponential of the Lanczos stopping rule,
c = exp(-i tau T) e0 for the Krylov matrix T (symmetric tridiagonal, k <= 64
rows): a Chebyshev-Bessel expansion over the Gershgorin interval of T, with
Taylor sub-steps when the expansion would run past 60 terms.  Checked
against the dense T's eigen-decomposition (scipy) and the oracle's Jacobi
route, on both branches, for the shapes the Lanczos
loop produces (k from 1 to 64) and special cases (T = c I, tau = 0,
negative tau).  Absolute tolerance 1e-12 (|c| <= 1) against the dense
route, 1e-11 against the oracle."""
import ctypes as ct

import numpy as np
import pytest
import scipy.linalg as sla

from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu
c_dp = ct.POINTER(ct.c_double)


def _gpu(alpha, beta, tau):
    lib = _abi.load_library()
    lib.yaqs_debug_tridiag_exp.argtypes = [ct.c_int, c_dp, c_dp, ct.c_double, c_dp]
    k = len(alpha)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    out = np.zeros(2 * k)
    assert lib.yaqs_debug_tridiag_exp(k, a.ctypes.data_as(c_dp), b.ctypes.data_as(c_dp), tau,
                                      out.ctypes.data_as(c_dp)) == 0
    return out.view(np.complex128)


def _dense(alpha, beta, tau):
    # V diag(e^{-i tau lambda}) V^T e0 of the dense symmetric T
    T = np.diag(alpha) + np.diag(beta[1:], 1) + np.diag(beta[1:], -1)
    lam, V = sla.eigh(T)
    return V @ (np.exp(-1j * tau * lam) * V[0, :])


@pytest.mark.parametrize("k", [1, 2, 5, 12, 25, 40, 64])
def test_tridiag_exp_random(k, oracle_lib):
    print(f"case: tridiag exp k={k}", flush=True)
    rng = np.random.default_rng(100 + k)
    # Krylov-like scales: diagonal around an energy offset, couplings of a
    # few units; tau of the TDVP half steps up to steps that need Taylor
    for shift, scale, tau in [(-30.0, 8.0, 0.05), (-30.0, 8.0, 0.025), (5.0, 2.0, -0.3), (0.0, 1.0, 2.0),
                              (-50.0, 20.0, 0.1), (0.0, 40.0, 1.5), (100.0, 30.0, 1.0)]:
        alpha = shift + scale * rng.standard_normal(k)
        beta = np.r_[0.0, np.abs(scale * rng.standard_normal(k - 1))]
        got = _gpu(alpha, beta, tau)
        ref = _dense(alpha, beta, tau)
        assert np.abs(got - ref).max() < 1e-12, (k, shift, scale, tau, np.abs(got - ref).max())
        # the oracle's Jacobi route carries phase errors of order eps |tau| ||T|| k
        orc = oracle_lib.tridiag_exp(alpha, beta, tau)
        assert np.abs(got - orc).max() < 1e-11


def test_tridiag_exp_special_cases():
    print("case: tridiag exp special cases", flush=True)
    # T = c I (zero Gershgorin radius), tau = 0, a single row
    for alpha, beta, tau in [(np.full(6, -3.5), np.zeros(6), 0.7), (np.array([1.0, 2.0, 3.0]), np.array([0.0, 0.5, 0.5]),
                                                                     0.0), (np.array([-7.25]), np.zeros(1), 0.05)]:
        got = _gpu(alpha, beta, tau)
        ref = _dense(alpha, beta, tau)
        assert np.abs(got - ref).max() < 1e-13, (alpha, beta, tau)
    # decoupled tail: rows past a zero coupling stay exactly zero
    alpha = np.array([1.0, -2.0, 0.5, 4.0, 3.0])
    beta = np.array([0.0, 0.8, 0.3, 0.0, 1.2])
    got = _gpu(alpha, beta, 0.4)
    assert np.all(got[3:] == 0)
    assert np.abs(got - _dense(alpha, beta, 0.4)).max() < 1e-13
