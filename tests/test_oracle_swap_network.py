"""Long-range 2-qubit gates with long_range_method :TEBD: the SWAP network of
apply_window! (src/DigitalTJM.jl:314-330) moves site b next to a with
nearest-neighbour SWAPs, applies the gate on (a, a+1) and unwinds.

Pinned by the dense state vector: without truncation the network is the
gate acting on the pair index s_a + 2 s_b of the original sites (closed
forms for CX, Haar circuits mixing long-range and nearest-neighbour gates
to 1e-10); TJMOptions picks the method per gate.  CPU only."""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

G = Y.gates
TEBD = Y.TJMOptions(local_method="TEBD", long_range_method="TEBD")


def _oracle(oracle_lib, L, circ, obs, opts=TEBD, noise=None, ntraj=1, thr=1e-12, chi=64, seed=0):
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=chi, truncation_threshold=thr)
    h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, seed=seed, alg_options=opts)
    return oracle_lib.run_digital(h, nthreads=4)


@pytest.mark.parametrize("flip,z", [(1, (-1.0, 1.0, 1.0)), (3, (-1.0, 1.0, -1.0))])
def test_long_range_cx_closed_form(oracle_lib, flip, z):
    """X on one site, then the reference's CX matrix on (1, 3): it flips
    site 1 when site 3 is 1 (pair index s_1 + 2 s_3)."""
    c = Y.DigitalCircuit(3)
    c.add_gate(G.XGate(), [flip])
    c.add_gate(G.CXGate(), [1, 3])
    obs = [Y.Observable(f"z{i}", G.ZGate(), i) for i in (1, 2, 3)]
    _, m, _, _ = _oracle(oracle_lib, 3, c, obs)
    assert np.allclose(m[:, -1], z, atol=1e-10)


def _dense(L, circ):
    psi = np.zeros(2 ** L, complex)
    psi[0] = 1.0
    for g in circ.gates:
        if not g.sites:
            continue
        U = g.op.matrix
        if len(g.sites) == 1:
            a = g.sites[0] - 1
            t = psi.reshape(2 ** a, 2, 2 ** (L - a - 1))
            psi = np.einsum("pq,lqr->lpr", U, t).reshape(-1)
        else:
            a, b = min(g.sites) - 1, max(g.sites) - 1
            t = np.moveaxis(psi.reshape([2] * L), (a, b), (0, 1))
            sh = t.shape
            # pair index s_a + 2 s_b: (s_b, s_a) row-major
            t = t.transpose(1, 0, *range(2, L)).reshape(4, -1)
            t = (U @ t).reshape(2, 2, *sh[2:]).transpose(1, 0, *range(2, L))
            psi = np.moveaxis(t, (0, 1), (a, b)).reshape(-1)
    return psi


def _long_range_circuit(L, depth, seed):
    rng = np.random.default_rng(seed)
    c = Y.DigitalCircuit(L)
    for q in range(1, L + 1):
        c.add_gate(G.RyGate(rng.uniform(0, np.pi)), [q])
    for _ in range(depth):
        a, b = sorted(rng.choice(np.arange(1, L + 1), 2, replace=False))
        U = G.haar_su4(rng)
        c.add_gate(G.MatrixGate(U), [int(a), int(b)] if rng.random() < 0.5 else [int(b), int(a)])
        q = int(rng.integers(1, L))
        c.add_gate(G.MatrixGate(G.haar_su4(rng)), [q, q + 1])
    return c


@pytest.mark.parametrize("L,depth,seed", [(4, 6, 1), (6, 8, 2), (7, 6, 3)])
def test_swap_network_matches_statevector(oracle_lib, L, depth, seed):
    circ = _long_range_circuit(L, depth, seed)
    obs = [Y.Observable(f"z{i}", G.ZGate(), i) for i in range(1, L + 1)]
    _, m, _, _ = _oracle(oracle_lib, L, circ, obs, thr=0.0, chi=2 ** (L // 2))
    psi = _dense(L, circ)
    p = np.abs(psi.reshape([2] * L)) ** 2
    for i in range(L):
        zi = p.sum(axis=tuple(k for k in range(L) if k != i)) @ np.array([1.0, -1.0])
        assert abs(m[i, -1] - zi) < 1e-10, (i, m[i, -1], zi)


def test_long_range_method_selects_the_gate_method():
    c = Y.DigitalCircuit(4)
    c.add_gate(G.RzzGate(0.3), [1, 3])
    c.add_gate(G.RzzGate(0.3), [3, 4])
    cfg = Y.StrongMeasurementConfig([Y.Observable("z1", G.ZGate(), 1)], num_traj=1)
    h = _abi.pack_digital(Y.MPS(4), c, cfg, None, alg_options=Y.TJMOptions(local_method="TEBD"))
    assert h.job.gates[0].method == _abi.GATE_WINDOW_TDVP and h.job.gates[1].method == _abi.GATE_EXACT
    h = _abi.pack_digital(Y.MPS(4), c, cfg, None, alg_options=Y.TJMOptions(long_range_method="TEBD"))
    assert h.job.gates[0].method == _abi.GATE_EXACT and h.job.gates[1].method == _abi.GATE_WINDOW_TDVP
