"""Unravelings of local noise processes (src/Noise.jl:134-295, 299-420).

Pinned three ways: the reference's own NoiseModel unit tests
(test/test_Noise.jl:11-71), the Lindblad dissipator of every expansion
against that of the standard process it unravels (the identity that makes
them unravelings: sum_k g_k (A_k rho A_k^+ - {A_k^+ A_k, rho}/2) must equal
g (P rho P^+ - {P^+ P, rho}/2), which fails for a wrong strength, sign or
weight), and the oracle's trajectory ensemble with unraveled dephasing
against the exact Lindblad solution of the standard model (CPU only)."""
import warnings

import numpy as np
import pytest
from scipy.linalg import expm

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi
from yaqs_julia_amd.noise import gauss_angles
from oracle import exact as E

X = np.array([[0, 1], [1, 0]], complex)
Yp = np.array([[0, -1j], [1j, 0]], complex)
Z = np.diag([1.0, -1.0]).astype(complex)


def test_reference_projector_case():
    """test/test_Noise.jl:24-53."""
    m = Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 0.2, "unraveling": "projector"}], 2)
    assert len(m.processes) == 2
    p1, p2 = m.processes
    assert p1.name.startswith("projector_plus") and p2.name.startswith("projector_minus")
    assert p1.strength == 0.1 and p2.strength == 0.1
    assert np.allclose(p1.matrix, np.eye(2) + Z) and np.allclose(p2.matrix, np.eye(2) - Z)
    assert not p1.pauli_name and not p2.pauli_name


def test_reference_unitary_2pt_case():
    """test/test_Noise.jl:55-71."""
    m = Y.NoiseModel([{"name": "pauli_x", "sites": [1], "strength": 0.2, "unraveling": "unitary_2pt",
                       "theta0": 0.1}], 2)
    assert len(m.processes) == 2
    p1, p2 = m.processes
    assert abs(p1.strength - (0.2 / np.sin(0.1) ** 2) / 2) < 1e-15
    assert np.allclose(p1.matrix, expm(1j * 0.1 * X)) and np.allclose(p2.matrix, expm(-1j * 0.1 * X))
    assert p1.name == "unitary2pt_plus_pauli_x" and p2.name == "unitary2pt_minus_pauli_x"


def test_standard_and_long_range():
    """test/test_Noise.jl:11-22; a long-range process becomes an MPO process
    (tests/test_oracle_longrange.py covers them)."""
    m = Y.NoiseModel([{"name": "pauli_x", "sites": [1], "strength": 0.1}], 2)
    assert len(m.processes) == 1 and m.processes[0].name == "pauli_x" and m.processes[0].pauli_name
    assert np.allclose(m.processes[0].matrix, X)
    lr = Y.NoiseModel([{"name": "crosstalk_xy", "sites": [1, 3], "strength": 0.05, "unraveling": "projector"}], 4)
    assert all(hasattr(p, "mpo") for p in lr.processes)
    with pytest.raises(ValueError):
        Y.NoiseModel([{"name": "pauli_x", "sites": [], "strength": 0.1}], 2)


def test_keywords_warnings_and_fallbacks():
    with pytest.warns(UserWarning, match="Stability"):
        Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 1.0, "unraveling": "unitary_2pt",
                       "theta0": 0.1}], 1, dt=0.1)
    with pytest.warns(UserWarning, match="deprecated"):
        m = Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 0.3, "unraveling": "analog_auto"}], 1)
    assert len(m.processes) == 1 and m.processes[0].name == "pauli_z"
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        # the default theta0 = pi/2 gives U = +-iP at rate gamma/2 each
        m = Y.NoiseModel([{"name": "pauli_y", "sites": [2], "strength": 0.4, "unraveling": "unitary_2pt"}], 2,
                         dt=0.1)
    assert np.allclose(m.processes[0].matrix, 1j * Yp) and np.allclose(m.processes[1].matrix, -1j * Yp)
    assert np.allclose([p.strength for p in m.processes], 0.2)
    with pytest.raises(ValueError):
        Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 0.1, "unraveling": "unitary_2pt",
                       "theta0": 0.0}], 1)
    with pytest.raises(ValueError):  # a single angle (0) cannot carry any rate
        Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 0.1, "unraveling": "unitary_gauss", "M": 1}], 1)


def test_gauss_grid():
    """src/Noise.jl:226-237: M odd gives M angles symmetric about 0 on
    [-theta_max, theta_max]; M even drops to 2*((M+1)//2) - 1 angles."""
    th, w = gauss_angles(1.0, 11, 4.0)
    assert len(th) == 11 and np.allclose(th, np.linspace(-4, 4, 11))
    assert abs(w.sum() - 1) < 1e-15 and np.array_equal(w, w[::-1])
    assert np.allclose(w, np.exp(-0.5 * th ** 2) / np.exp(-0.5 * th ** 2).sum())
    th10, _ = gauss_angles(0.5, 10, 2.0)
    assert len(th10) == 9 and th10[0] == -2.0 and th10[-1] == 2.0
    m = Y.NoiseModel([{"name": "pauli_x", "sites": [3], "strength": 0.1, "unraveling": "unitary_gauss"}], 3,
                     sigma=0.3, gauss_M=7)
    assert [p.name for p in m.processes] == [f"unitary_gauss_{k}_pauli_x" for k in range(7)]
    th7, w7 = gauss_angles(0.3, 7, 4.0 * 0.3)
    lam = 0.1 / np.sum(w7 * np.sin(th7) ** 2)
    for p, t, wk in zip(m.processes, th7, w7):
        assert abs(p.strength - lam * wk) < 1e-14 and np.allclose(p.matrix, expm(1j * t * X))
    m2 = Y.NoiseModel([{"name": "pauli_x", "sites": [3], "strength": 0.1, "unraveling": "unitary_gauss",
                        "M": 5, "theta_max": 1.0}], 3, sigma=2.0)
    assert len(m2.processes) == 5 and np.allclose(m2.processes[0].matrix, expm(-1j * 1.0 * X))


def _dissipator(procs, d):
    """Lindblad dissipator as a d^2 x d^2 superoperator on column-stacked rho."""
    I = np.eye(d)
    D = np.zeros((d * d, d * d), complex)
    for g, A in procs:
        AdA = A.conj().T @ A
        D += g * (np.kron(A.conj(), A) - 0.5 * np.kron(I, AdA) - 0.5 * np.kron(AdA.T, I))
    return D


CASES = [({"name": "pauli_z", "sites": [1]}, {}),
         ({"name": "pauli_x", "sites": [1]}, {}),
         ({"name": "pauli_y", "sites": [1], "theta0": 0.37}, {}),
         ({"name": "crosstalk_xz", "sites": [2, 1]}, {}),
         ({"name": "crosstalk_yy", "sites": [1, 2], "theta0": 1.1, "M": 9}, {"sigma": 0.4})]


@pytest.mark.parametrize("unravel", ["projector", "unitary_2pt", "unitary_gauss"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_unravelings_give_the_standard_dissipator(unravel, case):
    base, kw = CASES[case]
    g = 0.23
    std = Y.NoiseModel([dict(base, strength=g)], 2)
    ext = Y.NoiseModel([dict(base, strength=g, unraveling=unravel)], 2, **kw)
    P = std.processes[0].matrix
    d = P.shape[0]
    ref = _dissipator([(g, P)], d)
    got = _dissipator([(p.strength, p.matrix) for p in ext.processes], d)
    assert len(ext.processes) >= 2 and all(p.sites == std.processes[0].sites for p in ext.processes)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, max(p.strength for p in ext.processes))
    if unravel != "projector":  # unitary jumps: A^+ A = I, so the weight is rate x norm
        for p in ext.processes:
            assert np.allclose(p.matrix.conj().T @ p.matrix, np.eye(d))


@pytest.mark.parametrize("unravel,kw", [("projector", {}), ("unitary_2pt", {"theta0": 0.6}),
                                        ("unitary_gauss", {"sigma": 0.5})])
def test_oracle_ensemble_matches_lindblad(oracle_lib, unravel, kw):
    """Ising L=4 (J=1, g=0.5) from |++++>, dephasing gamma=0.15 unraveled,
    relaxation gamma=0.1 standard, dt=0.05, T=1, 300 trajectories: the
    oracle's mean <Z_i>(t) and <X_i>(t) against the exact Lindblad solution of the standard
    model within sampling error (measured: max 0.045, mean 0.009; against a
    Lindblad solution with twice the dephasing rate the same ensembles give
    mean deviations of 0.023-0.027, so the bound discriminates)."""
    L, ntraj, dt, steps = 4, 300, 0.05, 20
    procs = [{"name": "lowering", "sites": [i], "strength": 0.1} for i in range(1, L + 1)]
    procs += [{"name": "pauli_z", "sites": [i], "strength": 0.15, "unraveling": unravel} for i in range(1, L + 1)]
    noise = Y.NoiseModel(procs, L, **kw)
    obs = [Y.Observable(f"z{s}", Y.gates.ZGate(), s) for s in range(1, L + 1)] + \
          [Y.Observable(f"x{s}", Y.gates.XGate(), s) for s in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, dt * steps, dt=dt, num_traj=ntraj, max_bond_dim=16)
    h = _abi.pack_analog(Y.MPS(L, state="x+"), Y.init_ising(L, 1.0, 0.5), cfg, noise, seed=3)
    traj, mean, st = oracle_lib.run_analog(h, nthreads=8)
    assert st["jumps"] > 0
    H = E.ising(L, 1.0, 0.5)
    jumps = [(0.1, E.site_op(E.LOWER, s, L)) for s in range(1, L + 1)] + \
            [(0.15, E.site_op(E.Z, s, L)) for s in range(1, L + 1)]
    plus = np.ones(2) / np.sqrt(2)
    psi = plus
    for _ in range(L - 1):
        psi = np.kron(psi, plus)
    ops = [E.site_op(E.Z, s, L) for s in range(1, L + 1)] + [E.site_op(E.X, s, L) for s in range(1, L + 1)]
    ex = E.lindblad(H, jumps, np.outer(psi, psi.conj()), dt, steps, ops)
    dev = np.abs(mean - ex)
    assert dev.max() < 0.07 and dev.mean() < 0.015, (unravel, dev.max(), dev.mean())


def _pair_op(M, s, L):
    """A 4x4 operator in the reference's pair index (p1 + 2 p2, p1 on site s)
    as a dense operator on sites (s, s+1) of the exact solver (site 1 most
    significant)."""
    import scipy.sparse as sp
    P = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
    Mstd = P @ np.asarray(M, complex) @ P
    return sp.kron(sp.kron(sp.identity(2 ** (s - 1), format="csr"), sp.csr_matrix(Mstd)),
                   sp.identity(2 ** (L - s - 1), format="csr"), format="csr")


@pytest.mark.parametrize("name", ["custom", "raising_two", "crosstalk_xy"])
def test_oracle_two_site_ensemble_matches_lindblad(oracle_lib, name):
    """Adjacent 2-site processes (merged-pair dissipation with the T-C split,
    merged-pair weights, local_2 jumps): Ising L=4 (J=1, g=0.5) from |++++>,
    lowering gamma=0.1 on every site and the 2-site process gamma=0.2 on
    every pair, dt=0.05, T=1, 300 trajectories: the oracle's mean <Z_i>,
    <X_i> against the exact Lindblad solution within sampling error
    (measured mean deviations 0.009-0.013; against Lindblad solutions with
    twice the pair rate, the pair's sites swapped or the operator transposed
    the same ensembles give 0.024-0.11 wherever those change the model)."""
    L, ntraj, dt, steps = 4, 300, 0.05, 20
    rng = np.random.default_rng(11)
    M = (rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))) / 2 if name == "custom" else None
    procs = [{"name": "lowering", "sites": [i], "strength": 0.1} for i in range(1, L + 1)]
    for i in range(1, L):
        p = {"name": name, "sites": [i, i + 1], "strength": 0.2}
        if M is not None:
            p["matrix"] = M
        procs.append(p)
    noise = Y.NoiseModel(procs, L)
    obs = [Y.Observable(f"z{s}", Y.gates.ZGate(), s) for s in range(1, L + 1)] + \
          [Y.Observable(f"x{s}", Y.gates.XGate(), s) for s in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, dt * steps, dt=dt, num_traj=ntraj, max_bond_dim=16)
    h = _abi.pack_analog(Y.MPS(L, state="x+"), Y.init_ising(L, 1.0, 0.5), cfg, noise, seed=5)
    traj, mean, st = oracle_lib.run_analog(h, nthreads=8)
    assert st["jumps"] > 0
    H = E.ising(L, 1.0, 0.5)
    op2 = noise.processes[-1].matrix
    jumps = [(0.1, E.site_op(E.LOWER, s, L)) for s in range(1, L + 1)] + \
            [(0.2, _pair_op(op2, s, L)) for s in range(1, L)]
    plus = np.ones(2) / np.sqrt(2)
    psi = plus
    for _ in range(L - 1):
        psi = np.kron(psi, plus)
    ops = [E.site_op(E.Z, s, L) for s in range(1, L + 1)] + [E.site_op(E.X, s, L) for s in range(1, L + 1)]
    ex = E.lindblad(H, jumps, np.outer(psi, psi.conj()), dt, steps, ops)
    dev = np.abs(mean - ex)
    print(name, dev.max(), dev.mean())
    assert dev.max() < 0.07 and dev.mean() < 0.015, (name, dev.max(), dev.mean())
