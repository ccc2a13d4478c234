"""single_site_tdvp! (_tdvp_sweep_hamiltonian_1site!, src/Algorithms.jl:
471-532, with project_bond :222-226) in the oracle.

- The reference's closed form (test/test_Algorithms.jl:37-63): the Ising
  eigenstate |0000> (J=1, g=0) after five dt=0.1 sweeps has
  <ψ0|ψ(t)> = exp(1.5i) (scalar_product(ψ(t), ψ0), src/MPS.jl:528-585).
- On an MPS whose bonds all have their maximal dimension the 1-site tangent
  space is the whole space and the projector-splitting sweep is exact: the
  overlap <ψ0|ψ(t)> and every <Z_i> equal the dense exp(-iHt) values for
  any dt.  A bond step evolved forward instead of backward, a dropped bond
  step or a left environment built from the site tensor before its QR
  breaks this (each was planted in the oracle while writing this test and
  failed it).
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi
from oracle import exact as E
from oracle import oracle as O


def _dense(m):
    v = np.asarray(m.tensors[0])[0]  # (d, dr)
    for t in m.tensors[1:]:
        t = np.asarray(t)
        v = np.tensordot(v, t, axes=([v.ndim - 1], [0]))
    return v.reshape(-1)


def _full_bond_mps(L, seed):
    rng = np.random.default_rng(seed)
    dims = [min(2 ** i, 2 ** (L - i)) for i in range(L + 1)]
    ts = [rng.standard_normal((dims[i], 2, dims[i + 1])) + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))
          for i in range(L)]
    m = Y.MPS(L, tensors=ts)
    m.normalize()
    return m


def test_single_site_tdvp_ising_eigenstate():
    L = 4
    cfg = Y.TimeEvolutionConfig([], 0.5, dt=0.1, num_traj=1)
    h = _abi.pack_analog(Y.MPS(L, state="zeros"), Y.init_ising(L, 1.0, 0.0), cfg, None)
    ov = O.tdvp_overlap(h, 5, 1)
    assert abs(abs(ov) - 1.0) < 1e-12
    assert abs(ov - np.exp(1.5j)) < 1e-12


@pytest.mark.parametrize("L,dt", [(4, 0.1), (5, 0.1), (5, 0.05), (6, 0.2)])
def test_single_site_tdvp_exact_at_full_bond(L, dt):
    m = _full_bond_mps(L, seed=L)
    H = Y.init_general_hamiltonian(L, 1.0, 0.7, 0.4, 0.3, 0.0, 0.5)
    T = 0.4
    steps = int(round(T / dt))
    cfg = Y.TimeEvolutionConfig([], T, dt=dt, num_traj=1, truncation_threshold=0.0)
    ov = O.tdvp_overlap(_abi.pack_analog(m, H, cfg, None), steps, 1)
    psi0 = _dense(m)
    Hd = E.general(L, 1.0, 0.7, 0.4, 0.3, 0.0, 0.5).toarray()
    import scipy.linalg as sla
    psit = sla.expm(-1j * T * Hd) @ psi0
    ex = np.vdot(psi0, psit)
    assert abs(ov - ex) < 1e-10, (ov, ex)
    # per-step <Z_i> through the analog driver without noise (one trajectory)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, T, dt=dt, num_traj=1, truncation_threshold=0.0)
    r, _, _ = O.run_analog(_abi.pack_analog(m, H, cfg, None, tdvp_sites=1), nthreads=1)
    ops = [E.site_op(E.Z, i, L) for i in range(1, L + 1)]
    ez = E.schrodinger(E.general(L, 1.0, 0.7, 0.4, 0.3, 0.0, 0.5), psi0, dt, steps, ops)
    assert np.abs(r[:, :, 0] - ez).max() < 1e-10


@pytest.mark.parametrize("noise_scale", [1e-8, 1e-4])
def test_pad_bond_dimension_then_single_site_tdvp(noise_scale):
    """MPS(L; state, pad) (src/MPS.jl:139-141, 453-485): the padded Néel
    state has full bonds, equals the product state to O(noise_scale^2), and
    the 1-site sweep from it follows the exact dense evolution of the padded
    state (entanglement grows from the padded directions; from bond
    dimension 1 it cannot)."""
    L = 6
    m = Y.MPS(L, state="Neel")
    m.pad_bond_dimension(8, noise_scale=noise_scale, rng=np.random.default_rng(0))
    assert [t.shape[2] for t in m.tensors] == [2, 4, 8, 4, 2, 1]
    psi0 = _dense(m)
    psi_neel = E.product_state(L, [0, 1, 0, 1, 0, 1])
    assert abs(abs(np.vdot(psi_neel, psi0)) - 1.0) < 10 * noise_scale ** 2
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    cfg = Y.TimeEvolutionConfig([], 1.0, dt=0.1, num_traj=1, truncation_threshold=0.0)
    ov = O.tdvp_overlap(_abi.pack_analog(m, H, cfg, None), 10, 1)
    import scipy.linalg as sla
    ex = np.vdot(psi0, sla.expm(-1j * E.general(L, 1, 1, 1, 0, 0, 0).toarray()) @ psi0)
    assert abs(ov - ex) < 1e-10
    assert abs(ex - 1.0) > 0.1  # the Néel state is not stationary
    flat = Y.MPS(L, state="Neel")
    ovf = O.tdvp_overlap(_abi.pack_analog(flat, H, cfg, None), 10, 1)
    assert abs(ovf - ex) > 0.1  # bond dimension 1: mean-field dynamics
