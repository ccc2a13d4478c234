"""GPU parity of the batched analog TJM engine against the FP64 oracle.

FP64 mode runs the same algorithm on the same per-trajectory seeds, so every
trajectory (jump sequence included) must match the oracle: tolerance 1e-10 on
each <O>(t) (measured 1e-14 .. 3e-11; each test prints its deviation),
including after jumps that leave exactly rank-deficient site tensors
(raising on |0>, projector unravelings): there the reference's LAPACK
completes the null space from rounding noise, while the oracle and the
engine both drop the dead directions (the bond at its numerical rank).
Mixed mode (complex-FP32 contractions) is compared on ensemble means within
1e-4 where the jump sequences coincide, and within sampling error otherwise.
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def _gpu(ctx, mps, H, cfg, noise, precision="fp64", seed=0):
    h = _abi.pack_analog(mps, H, cfg, noise, precision=_abi.PREC_MIXED if precision == "mixed" else _abi.PREC_FP64,
                         seed=seed)
    ncol = len(cfg.times) if cfg.sample_timesteps else 1
    jid = ctx.submit_analog(h)
    traj, mean, _, stats = ctx.wait(jid, len(cfg.observables), ncol, cfg.num_traj)
    return traj, mean, stats


def _oracle(oracle_lib, mps, H, cfg, noise, seed=0):
    h = _abi.pack_analog(mps, H, cfg, noise, seed=seed)
    return oracle_lib.run_analog(h, nthreads=8)


def _err(tag, g, r):
    """Largest |gpu - oracle| over every observable, time and trajectory,
    printed so each run records the measured deviation."""
    e = float(np.abs(np.asarray(g) - np.asarray(r)).max())
    print(f"  {tag}: max|gpu-oracle| = {e:.3e}", flush=True)
    return e


def _zx(L):
    return [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
           [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]


def test_noise_free_tdvp_matches_oracle(ctx, oracle_lib):
    print("case: noise-free Ising L=6 chi=16", flush=True)
    L = 6
    cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.05, num_traj=1, max_bond_dim=16, truncation_threshold=1e-10)
    H = Y.init_ising(L, 1.0, 0.5)
    g, gm, st = _gpu(ctx, Y.MPS(L), H, cfg, None)
    r, rm, _ = _oracle(oracle_lib, Y.MPS(L), H, cfg, None)
    assert _err("g", g, r) < TOL_FP64
    assert st["matvecs"] > 0 and st["max_bond"] >= 4


def test_c1_per_trajectory_parity(ctx, oracle_lib):
    """C1 (reduced to 24 trajectories, T=2): Ising L=6 g=0.5, lowering +
    pauli_z gamma=0.1, dt=0.1, chi=16, threshold 1e-10."""
    print("case: C1 L=6 24 traj fp64", flush=True)
    L = 6
    cfg = Y.TimeEvolutionConfig(_zx(L), 2.0, dt=0.1, num_traj=24, max_bond_dim=16, truncation_threshold=1e-10)
    H = Y.init_ising(L, 1.0, 0.5)
    noise = Y.relaxation_dephasing(L, 0.1)
    g, gm, st = _gpu(ctx, Y.MPS(L), H, cfg, noise, seed=11)
    r, rm, rst = _oracle(oracle_lib, Y.MPS(L), H, cfg, noise, seed=11)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert _err("g", g, r) < TOL_FP64
    assert _err("gm", gm, rm) < TOL_FP64


def test_heisenberg_neel_parity(ctx, oracle_lib):
    """C2-like at small size: Heisenberg XXX from the Neel state, relaxation +
    dephasing, chi_max=8 so the bond cap binds."""
    print("case: Heisenberg L=8 chi=8 16 traj fp64", flush=True)
    L = 8
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.05, num_traj=16, max_bond_dim=8, truncation_threshold=1e-9)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    noise = Y.relaxation_dephasing(L, 0.05)
    g, gm, st = _gpu(ctx, Y.MPS(L, state="Neel"), H, cfg, noise, seed=5)
    r, rm, rst = _oracle(oracle_lib, Y.MPS(L, state="Neel"), H, cfg, noise, seed=5)
    assert st["jumps"] == rst["jumps"]
    assert _err("g", g, r) < TOL_FP64


def test_first_order_and_final_only(ctx, oracle_lib):
    print("case: order 1, sample_timesteps=False", flush=True)
    L = 5
    for order, sample in ((1, True), (2, False), (1, False)):
        cfg = Y.TimeEvolutionConfig(_zx(L), 0.6, dt=0.1, num_traj=6, max_bond_dim=8, order=order,
                                    sample_timesteps=sample)
        H = Y.init_ising(L, 0.8, 0.9)
        noise = Y.relaxation_dephasing(L, 0.2)
        g, _, _ = _gpu(ctx, Y.MPS(L), H, cfg, noise, seed=2)
        r, _, _ = _oracle(oracle_lib, Y.MPS(L), H, cfg, noise, seed=2)
        assert g.shape == r.shape
        assert _err("g", g, r) < TOL_FP64, (order, sample)


def test_two_site_observables_and_raising(ctx, oracle_lib):
    print("case: 2-site observables, raising noise", flush=True)
    L = 5
    zz = Y.gates.MatrixGate(np.kron(Y.gates.ZGate().matrix, Y.gates.ZGate().matrix))
    xy = Y.gates.MatrixGate(np.kron(Y.gates.XGate().matrix, Y.gates.YGate().matrix))
    obs = [Y.Observable("zz12", zz, [1, 2]), Y.Observable("xy34", xy, [3, 4]), Y.Observable("y5", Y.gates.YGate(), 5)]
    cfg = Y.TimeEvolutionConfig(obs, 0.5, dt=0.05, num_traj=8, max_bond_dim=16)
    H = Y.init_general_hamiltonian(L, 0.7, 0.4, 1.0, 0.3, 0.0, 0.2)
    noise = Y.NoiseModel([{"name": "raising", "sites": [i], "strength": 0.3} for i in range(1, L + 1)] +
                         [{"name": "pauli_x", "sites": [2], "strength": 0.2}], L)
    g, _, st = _gpu(ctx, Y.MPS(L), H, cfg, noise, seed=9)
    r, _, rst = _oracle(oracle_lib, Y.MPS(L), H, cfg, noise, seed=9)
    assert st["jumps"] == rst["jumps"]
    # raising jumps on |0> sites leave exactly rank-deficient bonds
    assert _err("g", g, r) < TOL_FP64


def test_mixed_precision_ensemble(ctx, oracle_lib):
    """Complex-FP32 contractions with FP64 Lanczos/Gram/norms: ensemble <Z_i>
    within 1e-4 of the FP64 oracle on the same seeds."""
    print("case: mixed precision L=8 32 traj", flush=True)
    L = 8
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.05, num_traj=32, max_bond_dim=16)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    noise = Y.relaxation_dephasing(L, 0.01)
    g, gm, st = _gpu(ctx, Y.MPS(L, state="Neel"), H, cfg, noise, precision="mixed", seed=4)
    r, rm, rst = _oracle(oracle_lib, Y.MPS(L, state="Neel"), H, cfg, noise, seed=4)
    same = np.all(np.abs(g - r).max(axis=(0, 1)) < 1e-2)
    assert same or st["jumps"] != rst["jumps"]
    if st["jumps"] == rst["jumps"]:
        assert _err("<Z_i> mean", gm[:L], rm[:L]) < 1e-4
        assert _err("<X_i> mean", gm[L:], rm[L:]) < 1e-4
    else:
        assert np.abs(gm[:L] - rm[:L]).max() < 0.05


def test_batch_invariance(ctx, oracle_lib):
    """Results do not depend on how trajectories are grouped into waves:
    a 10-trajectory job equals trajectories 0..9 of a 20-trajectory job."""
    print("case: batch invariance", flush=True)
    L = 6
    H = Y.init_ising(L, 1.0, 0.7)
    noise = Y.relaxation_dephasing(L, 0.15)
    cfg10 = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.1, num_traj=10, max_bond_dim=16)
    cfg20 = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.1, num_traj=20, max_bond_dim=16)
    a, _, _ = _gpu(ctx, Y.MPS(L), H, cfg10, noise, seed=8)
    b, _, _ = _gpu(ctx, Y.MPS(L), H, cfg20, noise, seed=8)
    assert np.abs(a - b[:, :, :10]).max() < 1e-10


@pytest.mark.parametrize("unravel,kw", [("projector", {}), ("unitary_2pt", {"theta0": 0.6}),
                                        ("unitary_gauss", {"sigma": 0.5, "gauss_M": 7})])
def test_unraveled_dephasing_parity(ctx, oracle_lib, unravel, kw):
    """Projector, two-point and Gaussian unitary unravelings of pauli_z
    dephasing (src/Noise.jl:134-252): 1-site processes with explicit
    matrices, non-Pauli names and (for the unitary ones) norm-preserving
    jumps.  Per-trajectory FP64 parity with the oracle, Ising L=6 from |++..>,
    lowering gamma=0.05 standard, dephasing gamma=0.2 unraveled."""
    print(f"case: unraveled dephasing {unravel} L=6 16 traj fp64", flush=True)
    L = 6
    procs = [{"name": "lowering", "sites": [i], "strength": 0.05} for i in range(1, L + 1)]
    procs += [{"name": "pauli_z", "sites": [i], "strength": 0.2, "unraveling": unravel} for i in range(1, L + 1)]
    noise = Y.NoiseModel(procs, L, **kw)
    cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.05, num_traj=16, max_bond_dim=16, truncation_threshold=1e-10)
    H = Y.init_ising(L, 1.0, 0.5)
    g, gm, st = _gpu(ctx, Y.MPS(L, state="x+"), H, cfg, noise, seed=4)
    r, rm, rst = _oracle(oracle_lib, Y.MPS(L, state="x+"), H, cfg, noise, seed=4)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert _err("g", g, r) < TOL_FP64
