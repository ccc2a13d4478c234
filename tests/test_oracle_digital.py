"""The FP64 oracle's digital TJM (TEBD local gates, src/DigitalTJM.jl:238-305,
363-450) against closed forms and a dense statevector.

- Closed forms of the reference's own tests (test/test_DigitalTJM.jl:50-115):
  H then Rzz(pi/2) from |00>, <Z1> = 0 and <Z2> = 1; H H Rzz(pi/2),
  <Z1 Z2> = 0; X then pauli_z noise (strength 0.2), <Z1> = -1 on every
  trajectory whatever the jumps.
- Noise-free Haar brickwork without truncation against the dense state
  vector: the 4x4 gate acts on the pair index s_a + 2 s_{a+1} (the gate
  tensor is Julia's column-major reshape of the matrix, first listed site =
  low bit, src/DigitalTJM.jl:268-271), <Z_i> to 1e-10.
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

Z = np.diag([1.0, -1.0]).astype(complex)


def _oracle(oracle_lib, L, circ, obs, noise=None, ntraj=1, thr=1e-12, chi=64, seed=0):
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=chi, truncation_threshold=thr)
    h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, seed=seed)
    return oracle_lib.run_digital(h, nthreads=4)


def test_bell_closed_form(oracle_lib):
    c = Y.DigitalCircuit(2)
    c.add_gate(Y.gates.HGate(), [1])
    c.add_gate(Y.gates.RzzGate(np.pi / 2), [1, 2])
    obs = [Y.Observable("z1", Y.gates.ZGate(), 1), Y.Observable("z2", Y.gates.ZGate(), 2)]
    r, m, _, _ = _oracle(oracle_lib, 2, c, obs)
    assert abs(m[0, -1]) < 1e-10 and abs(m[1, -1] - 1.0) < 1e-10


def test_entanglement_closed_form(oracle_lib):
    c = Y.DigitalCircuit(2)
    c.add_gate(Y.gates.HGate(), [1])
    c.add_gate(Y.gates.HGate(), [2])
    c.add_gate(Y.gates.RzzGate(np.pi / 2), [1, 2])
    obs = [Y.Observable("zz", Y.gates.MatrixGate(np.kron(Z, Z)), [1, 2])]
    r, m, _, _ = _oracle(oracle_lib, 2, c, obs)
    assert abs(m[0, -1]) < 1e-10


def test_pauli_z_noise_closed_form(oracle_lib):
    c = Y.DigitalCircuit(2)
    c.add_gate(Y.gates.XGate(), [1])
    noise = Y.NoiseModel([{"name": "pauli_z", "sites": [1], "strength": 0.2}], 2)
    obs = [Y.Observable("z1", Y.gates.ZGate(), 1)]
    r, m, _, st = _oracle(oracle_lib, 2, c, obs, noise=noise, ntraj=64, seed=5)
    assert np.abs(r[0, -1, :] + 1.0).max() < 1e-10


def _dense_brickwork(L, circ):
    psi = np.zeros(2 ** L, complex)
    psi[0] = 1.0
    for g in circ.gates:
        if not g.sites:
            continue
        U = Y.gates.matrix(g.op)
        if len(g.sites) == 1:
            a = g.sites[0] - 1
            t = psi.reshape(2 ** a, 2, 2 ** (L - a - 1))
            psi = np.einsum("pq,lqr->lpr", U, t).reshape(-1)
        else:
            # pair index s_a + 2 s_{a+1}: the gate tensor is Julia's
            # column-major reshape (2,2,2,2) of the 4x4 matrix
            # (src/DigitalTJM.jl:268-271), first site = low bit
            a = min(g.sites) - 1
            assert max(g.sites) == a + 2
            t = psi.reshape(2 ** a, 2, 2, 2 ** (L - a - 2)).transpose(0, 2, 1, 3).reshape(2 ** a, 4, -1)
            t = np.einsum("pq,lqr->lpr", U, t)
            psi = t.reshape(2 ** a, 2, 2, -1).transpose(0, 2, 1, 3).reshape(-1)
    return psi


@pytest.mark.parametrize("L,depth,seed", [(4, 4, 1), (6, 6, 2), (7, 5, 3)])
def test_brickwork_matches_statevector(oracle_lib, L, depth, seed):
    circ = Y.haar_brickwork(L, depth, seed=seed)
    circ.gates.insert(0, Y.circuit.DigitalGate(Y.gates.HGate(), [2]))
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    r, m, _, _ = _oracle(oracle_lib, L, circ, obs, thr=0.0, chi=2 ** (L // 2))
    psi = _dense_brickwork(L, circ)
    p = np.abs(psi.reshape([2] * L)) ** 2
    for i in range(L):
        zi = p.sum(axis=tuple(k for k in range(L) if k != i)) @ np.array([1.0, -1.0])
        assert abs(m[i, -1] - zi) < 1e-10, (i, m[i, -1], zi)
