"""Exact dense solvers for small chains (TEST INFRASTRUCTURE ONLY).

Dense Schrödinger evolution and the Lindblad master equation by a vectorised
Liouvillian, for L <= 8.  Site 1 is the most significant qubit of the dense
index; |0> has Z = +1 (src/GateLibrary.jl:56-58).  These play the role of the
reference's exact checks (test/benchmark_tdvp.jl:94-216 and the QuTiP
comparison of 00_safety_checks_full_algorithms/analogTJM_check/
benchmark_analog_tjm.jl:20-227).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)
LOWER = np.array([[0, 1], [0, 0]], dtype=complex)   # |0><1|
RAISE = np.array([[0, 0], [1, 0]], dtype=complex)   # |1><0|


def site_op(op, i, L):
    """op on site i (1-based), identity elsewhere; site 1 most significant."""
    return sp.kron(sp.kron(sp.identity(2 ** (i - 1), format="csr"), sp.csr_matrix(op)),
                   sp.identity(2 ** (L - i), format="csr"), format="csr")


def ising(L, J, g):
    """H = -Σ J_i Z_i Z_{i+1} - Σ g_i X_i (the Hamiltonian of init_ising)."""
    Jv = np.broadcast_to(np.asarray(J, float), (L - 1,)) if np.ndim(J) == 0 else np.asarray(J, float)
    gv = np.broadcast_to(np.asarray(g, float), (L,)) if np.ndim(g) == 0 else np.asarray(g, float)
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=complex)
    for i in range(1, L):
        H = H - Jv[i - 1] * (site_op(Z, i, L) @ site_op(Z, i + 1, L))
    for i in range(1, L + 1):
        H = H - gv[i - 1] * site_op(X, i, L)
    return H


def general(L, Jxx, Jyy, Jzz, hx, hy, hz):
    H = sp.csr_matrix((2 ** L, 2 ** L), dtype=complex)
    for i in range(1, L):
        H = H + Jxx * site_op(X, i, L) @ site_op(X, i + 1, L) + Jyy * site_op(Y, i, L) @ site_op(Y, i + 1, L) \
            + Jzz * site_op(Z, i, L) @ site_op(Z, i + 1, L)
    for i in range(1, L + 1):
        H = H + hx * site_op(X, i, L) + hy * site_op(Y, i, L) + hz * site_op(Z, i, L)
    return H


def product_state(L, bits):
    v = np.zeros(2 ** L, dtype=complex)
    idx = 0
    for b in bits:
        idx = 2 * idx + int(b)
    v[idx] = 1.0
    return v


def schrodinger(H, psi0, dt, steps, ops):
    """<op>(t_k) for k = 0..steps under exp(-i H dt)."""
    U = sla.expm(-1j * dt * H.toarray())
    psi = psi0.copy()
    out = np.zeros((len(ops), steps + 1))
    for k in range(steps + 1):
        for q, o in enumerate(ops):
            out[q, k] = np.real(np.vdot(psi, o @ psi))
        psi = U @ psi
    return out


def lindblad(H, jumps, rho0, dt, steps, ops):
    """Tr(op ρ(t_k)) for dρ/dt = -i[H,ρ] + Σ γ (LρL† - ½{L†L,ρ}); jumps is a
    list of (γ, L).  Column-stacking vec: vec(AρB) = (Bᵀ ⊗ A) vec(ρ)."""
    n = H.shape[0]
    Id = sp.identity(n, format="csr")
    Hs = sp.csr_matrix(H)
    Lv = -1j * (sp.kron(Id, Hs) - sp.kron(Hs.T, Id))
    for g, Lk in jumps:
        Lk = sp.csr_matrix(Lk)
        LdL = (Lk.conj().T @ Lk)
        Lv = Lv + g * (sp.kron(Lk.conj(), Lk) - 0.5 * sp.kron(Id, LdL) - 0.5 * sp.kron(LdL.T, Id))
    Lv = sp.csc_matrix(Lv)
    r = rho0.reshape(-1, order="F")
    traj = spla.expm_multiply(Lv, r, start=0.0, stop=dt * steps, num=steps + 1, endpoint=True)
    out = np.zeros((len(ops), steps + 1))
    for k in range(steps + 1):
        rho = traj[k].reshape(n, n, order="F")
        for q, o in enumerate(ops):
            out[q, k] = np.real(np.trace(o @ rho))
    return out
