"""single_site_tdvp! (src/Algorithms.jl:471-532) on the GPU, as the sweep
of an analog job (`tdvp_sites = 1`), against the FP64 oracle.

- Noise-free, full-bond random MPS (L=6, bonds 2,4,8,4,2), general XYZ
  Hamiltonian with fields: every <Z_i>, <X_i> at every step within 1e-10 of
  the oracle, which is exact there (tests/test_oracle_tdvp1.py).
- Noisy, bonds capped at 4 (L=8), lowering + dephasing: per trajectory,
  jump sequences included, within 1e-10 (FP64), and the mixed-precision
  ensemble within 1e-4 of the oracle's.
- From a product state (bond dimension 1, mean-field dynamics) with noise.
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def _zx(L):
    return [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
           [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]


def _mps(L, cap, seed):
    rng = np.random.default_rng(seed)
    dims = [min(cap, 2 ** i, 2 ** (L - i)) for i in range(L + 1)]
    ts = [rng.standard_normal((dims[i], 2, dims[i + 1])) + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))
          for i in range(L)]
    m = Y.MPS(L, tensors=ts)
    m.normalize()
    return m


def _run(ctx, oracle_lib, m, H, cfg, noise, precision, seed=0):
    L = m.length
    h = _abi.pack_analog(m, H, cfg, noise, precision=precision, seed=seed, tdvp_sites=1)
    jid = ctx.submit_analog(h)
    g, gm, _, st = ctx.wait(jid, len(cfg.observables), len(cfg.times), cfg.num_traj)
    r, rm, rst = oracle_lib.run_analog(_abi.pack_analog(m, H, cfg, noise, seed=seed, tdvp_sites=1), nthreads=8)
    return g, gm, st, r, rm, rst


def test_tdvp1_noise_free_full_bond(ctx, oracle_lib):
    print("case: 1-site TDVP noise-free full-bond L=6 fp64", flush=True)
    L = 6
    m = _mps(L, 64, seed=11)
    H = Y.init_general_hamiltonian(L, 1.0, 0.7, 0.4, 0.3, 0.0, 0.5)
    cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.1, num_traj=1, max_bond_dim=64, truncation_threshold=0.0)
    g, gm, st, r, rm, rst = _run(ctx, oracle_lib, m, H, cfg, None, _abi.PREC_FP64)
    e = float(np.abs(g - r).max())
    print(f"  max|gpu-oracle| = {e:.3e}", flush=True)
    assert e < 1e-10


@pytest.mark.parametrize("precision,tol", [(_abi.PREC_FP64, 1e-10), (_abi.PREC_MIXED, 1e-4)])
def test_tdvp1_noisy_capped_bonds(ctx, oracle_lib, precision, tol):
    print(f"case: 1-site TDVP noisy L=8 bonds<=4 16 traj precision={precision}", flush=True)
    L = 8
    m = _mps(L, 4, seed=5)
    H = Y.init_ising(L, 1.0, 0.6)
    cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.05, num_traj=16, max_bond_dim=16)
    noise = Y.relaxation_dephasing(L, 0.1)
    g, gm, st, r, rm, rst = _run(ctx, oracle_lib, m, H, cfg, noise, precision, seed=7)
    if precision == _abi.PREC_FP64:
        e = float(np.abs(g - r).max())
        print(f"  max|gpu-oracle| = {e:.3e}  jumps {st['jumps']} / {int(rst['jumps'])}", flush=True)
        assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
        assert e < tol
    else:
        same = np.all(np.abs(g - r).max(axis=(0, 1)) < 1e-2)
        em = float(np.abs(gm - rm).max())
        print(f"  mixed ensemble max|gpu-oracle| = {em:.3e}  jumps {st['jumps']} / {int(rst['jumps'])}", flush=True)
        assert st["jumps"] == rst["jumps"] and same
        assert em < tol


def test_tdvp1_product_state_with_noise(ctx, oracle_lib):
    print("case: 1-site TDVP from a product state, noisy, L=6 fp64", flush=True)
    L = 6
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0.2)
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.05, num_traj=12, max_bond_dim=8)
    g, gm, st, r, rm, rst = _run(ctx, oracle_lib, Y.MPS(L, state="x+"), H, cfg, Y.relaxation_dephasing(L, 0.2),
                                 _abi.PREC_FP64, seed=3)
    e = float(np.abs(g - r).max())
    print(f"  max|gpu-oracle| = {e:.3e}  jumps {st['jumps']} / {int(rst['jumps'])}", flush=True)
    assert st["jumps"] == rst["jumps"] and e < 1e-10


def test_tdvp1_padded_state_follows_exact_evolution(ctx):
    """The reference's way to use the 1-site sweep: pad a product state with
    pad_bond_dimension! (default noise 1e-8, padded Schmidt values near
    1e-16) and evolve.  The engine's measurements and sweeps never move the
    centre through the padded directions before evolving them, so its
    <Z_i>(t) follows the exact dense evolution of the padded state."""
    print("case: 1-site TDVP from a padded Neel state (noise 1e-8) vs exact, L=6 fp64", flush=True)
    from oracle import exact as E
    L = 6
    m = Y.MPS(L, state="Neel")
    m.pad_bond_dimension(8, rng=np.random.default_rng(0))
    psi0 = np.asarray(m.tensors[0])[0]
    for t in m.tensors[1:]:
        psi0 = np.tensordot(psi0, np.asarray(t), axes=([psi0.ndim - 1], [0]))
    psi0 = psi0.reshape(-1)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, 1.0, dt=0.1, num_traj=1, max_bond_dim=8, truncation_threshold=0.0)
    jid = ctx.submit_analog(_abi.pack_analog(m, H, cfg, None, precision=_abi.PREC_FP64, tdvp_sites=1))
    g, _, _, _ = ctx.wait(jid, len(obs), len(cfg.times), 1)
    ops = [E.site_op(E.Z, i, L) for i in range(1, L + 1)]
    ez = E.schrodinger(E.general(L, 1, 1, 1, 0, 0, 0), psi0, 0.1, 10, ops)
    e = float(np.abs(g[:, :, 0] - ez).max())
    print(f"  max|gpu-exact| = {e:.3e}", flush=True)
    assert e < 1e-8
