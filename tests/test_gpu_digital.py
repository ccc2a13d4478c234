"""GPU parity of the batched DigitalTJM (TEBD) path against the FP64 oracle.

The GPU keeps the orthogonality centre next to the last gate and applies
dissipation and jumps locally, the oracle follows the reference's full-chain
sweeps (src/DigitalTJM.jl:405-418); both are the same state up to gauge, so
per-trajectory <Z_i> and jump sequences must agree in FP64.
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


def _err(tag, g, r):
    """Largest |gpu - oracle|, printed so each run records the deviation."""
    e = float(np.abs(np.asarray(g) - np.asarray(r)).max())
    print(f"  {tag}: max|gpu-oracle| = {e:.3e}", flush=True)
    return e


def _depolarizing(L, p):
    procs = []
    for i in range(1, L + 1):
        for name in ("pauli_x", "pauli_y", "pauli_z"):
            procs.append({"name": name, "sites": [i], "strength": p / 3})
    return Y.NoiseModel(procs, L)


def _run(ctx, oracle_lib, L, circ, cfg, noise, precision=_abi.PREC_FP64, seed=0):
    hg = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, precision=precision, seed=seed)
    ncol = int(ctx.lib.yaqs_digital_num_samples(hg.job))
    jid = ctx.submit_digital(hg)
    g, gm, gb, st = ctx.wait(jid, len(cfg.observables), ncol, cfg.num_traj, want_bonds=True)
    hr = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, seed=seed)
    r, rm, rb, rst = oracle_lib.run_digital(hr, nthreads=8)
    return g, gm, gb, st, r, rm, rb, rst


def test_brickwork_noise_free(ctx, oracle_lib):
    print("case: digital noise-free L=8 depth 6", flush=True)
    L = 8
    circ = Y.haar_brickwork(L, 6, seed=3, sample_every_layer=True, sample_at_start=True)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=1, max_bond_dim=64, truncation_threshold=1e-12)
    g, gm, gb, st, r, rm, rb, rst = _run(ctx, oracle_lib, L, circ, cfg, None)
    assert g.shape == r.shape == (L, 7, 1)
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)


def test_brickwork_depolarizing_parity(ctx, oracle_lib):
    print("case: digital depolarizing L=8 depth 8, 16 traj", flush=True)
    L = 8
    circ = Y.haar_brickwork(L, 8, seed=5)
    circ.gates.insert(0, Y.circuit.DigitalGate(Y.gates.HGate(), [3]))
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=16, max_bond_dim=16, truncation_threshold=1e-9)
    g, gm, gb, st, r, rm, rb, rst = _run(ctx, oracle_lib, L, circ, cfg, _depolarizing(L, 0.05), seed=7)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)


def test_digital_mixed_precision(ctx, oracle_lib):
    print("case: digital mixed L=10 depth 10", flush=True)
    L = 10
    circ = Y.haar_brickwork(L, 10, seed=9)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=8, max_bond_dim=32, truncation_threshold=1e-9)
    g, gm, gb, st, r, rm, rb, rst = _run(ctx, oracle_lib, L, circ, cfg, _depolarizing(L, 0.01),
                                         precision=_abi.PREC_MIXED, seed=1)
    if st["jumps"] == rst["jumps"]:
        assert _err("gm", gm, rm) < 1e-4


def test_brickwork_unraveled_noise_parity(ctx, oracle_lib):
    """Two-point unitary unraveling of dephasing and the projector
    unraveling of bit flips (src/Noise.jl:134-193) in the digital path."""
    print("case: digital unraveled noise L=8 depth 8, 16 traj", flush=True)
    L = 8
    circ = Y.haar_brickwork(L, 8, seed=6)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=16, max_bond_dim=16, truncation_threshold=1e-9)
    procs = [{"name": "pauli_z", "sites": [i], "strength": 0.03, "unraveling": "unitary_2pt", "theta0": 0.7}
             for i in range(1, L + 1)]
    procs += [{"name": "pauli_x", "sites": [i], "strength": 0.02, "unraveling": "projector"} for i in range(1, L + 1)]
    noise = Y.NoiseModel(procs, L)
    g, gm, gb, st, r, rm, rb, rst = _run(ctx, oracle_lib, L, circ, cfg, noise, seed=8)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)


def _generator_circuit(L, depth, seed):
    """Layers of Rxx/Ryy/Rzz/CZ/CPhase/CX on alternating bonds (random site
    order) between single-qubit rotations."""
    G = Y.gates
    rng = np.random.default_rng(seed)
    c = Y.DigitalCircuit(L)
    for q in range(1, L + 1):
        c.add_gate(G.RyGate(rng.uniform(0, np.pi)), [q])
    for layer in range(depth):
        for a in range(1 + layer % 2, L, 2):
            th = rng.uniform(-np.pi, np.pi)
            op = [G.RxxGate(th), G.RyyGate(th), G.RzzGate(th), G.CZGate(), G.CPhaseGate(th),
                  G.CXGate()][rng.integers(6)]
            c.add_gate(op, [a, a + 1] if rng.random() < 0.5 else [a + 1, a])
        for q in range(1, L + 1):
            c.add_gate(G.RzGate(rng.uniform(-1, 1)), [q])
    return c


def _run_opts(ctx, oracle_lib, L, circ, cfg, noise, opts, seed=0, mps=None):
    mps = mps or Y.MPS(L)
    hg = _abi.pack_digital(mps, circ, cfg, noise, seed=seed, alg_options=opts)
    ncol = int(ctx.lib.yaqs_digital_num_samples(hg.job))
    jid = ctx.submit_digital(hg)
    g, gm, gb, st = ctx.wait(jid, len(cfg.observables), ncol, cfg.num_traj, want_bonds=True)
    hr = _abi.pack_digital(mps, circ, cfg, noise, seed=seed, alg_options=opts)
    r, rm, rb, rst = oracle_lib.run_digital(hr, nthreads=8)
    return g, gm, gb, st, r, rm, rb, rst


@pytest.mark.parametrize("noisy", [False, True])
def test_window_tdvp_gates_parity(ctx, oracle_lib, noisy):
    """local_method :TDVP (the reference default): 2-qubit gates as
    exp(-i h) of their window Hamiltonian with the TDVP split (keep max(k, 2)),
    against the oracle's Krylov window step, per trajectory."""
    print(f"case: digital TDVP windows L=8 depth 8 noisy={noisy}", flush=True)
    L = 8
    circ = _generator_circuit(L, 8, seed=12)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=16 if noisy else 1, max_bond_dim=16, truncation_threshold=1e-9)
    noise = _depolarizing(L, 0.05) if noisy else None
    g, gm, gb, st, r, rm, rb, rst = _run_opts(ctx, oracle_lib, L, circ, cfg, noise, Y.TJMOptions(), seed=3)
    assert st["jumps"] == rst["jumps"]
    if noisy:
        assert st["jumps"] > 0
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)


def _long_range_circuit(L, depth, seed, pairs):
    """Haar gates: long-range ones on the given pairs (random site order),
    nearest-neighbour ones on random bonds, between Ry rotations."""
    G = Y.gates
    rng = np.random.default_rng(seed)
    c = Y.DigitalCircuit(L)
    for q in range(1, L + 1):
        c.add_gate(G.RyGate(rng.uniform(0, np.pi)), [q])
    for _ in range(depth):
        a, b = pairs[int(rng.integers(len(pairs)))]
        c.add_gate(G.MatrixGate(G.haar_su4(rng)), [a, b] if rng.random() < 0.5 else [b, a])
        q = int(rng.integers(1, L))
        c.add_gate(G.MatrixGate(G.haar_su4(rng)), [q, q + 1])
    return c


@pytest.mark.parametrize("noise_sites", [None, (2, 3, 6, 7), (1, 2, 3, 4, 5, 6, 7, 8)])
def test_swap_network_parity(ctx, oracle_lib, noise_sites):
    """Long-range gates as TEBD SWAP networks (src/DigitalTJM.jl:314-330),
    per trajectory against the oracle: noise-free, noise only on the sites
    the long-range gates leave alone, and noise on every site (dissipation
    and jumps on the two sites of the long-range gate after the network)."""
    print(f"case: digital SWAP networks L=8 noise on {noise_sites}", flush=True)
    L = 8
    circ = _long_range_circuit(L, 10, seed=21, pairs=[(1, 4), (5, 8), (1, 8), (4, 5)])
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=16 if noise_sites else 1, max_bond_dim=16,
                                    truncation_threshold=1e-9)
    noise = None
    if noise_sites:
        noise = Y.NoiseModel([{"name": n, "sites": [i], "strength": 0.03} for i in noise_sites
                              for n in ("pauli_x", "lowering")], L)
    opts = Y.TJMOptions(local_method="TEBD", long_range_method="TEBD")
    g, gm, gb, st, r, rm, rb, rst = _run_opts(ctx, oracle_lib, L, circ, cfg, noise, opts, seed=5)
    assert st["jumps"] == rst["jumps"]
    if noise_sites:
        assert st["jumps"] > 0
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)


def _long_range_generator_circuit(L, depth, seed):
    G = Y.gates
    rng = np.random.default_rng(seed)
    c = Y.DigitalCircuit(L)
    for q in range(1, L + 1):
        c.add_gate(G.RyGate(rng.uniform(0, np.pi)), [q])
    for _ in range(depth):
        a, b = sorted(int(x) for x in rng.choice(np.arange(1, L + 1), 2, replace=False))
        th = rng.uniform(-np.pi, np.pi)
        op = [G.RxxGate(th), G.RyyGate(th), G.RzzGate(th), G.CZGate(), G.CPhaseGate(th), G.CXGate()][rng.integers(6)]
        c.add_gate(op, [a, b] if rng.random() < 0.5 else [b, a])
        q = int(rng.integers(1, L + 1))
        c.add_gate(G.RxGate(rng.uniform(-1, 1)), [q])
    return c


def _generic_mps(L, chi, seed):
    """Random MPS in right-canonical form whose bonds all have full rank
    (dimension min(2^i, 2^(L-i), chi))."""
    rng = np.random.default_rng(seed)
    dims = [min(2 ** i, 2 ** (L - i), chi) for i in range(L + 1)]
    ts = [rng.standard_normal((dims[i], 2, dims[i + 1])) + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))
          for i in range(L)]
    m = Y.MPS(L, tensors=ts)
    m.normalize()
    return m


@pytest.mark.parametrize("noisy", [False, True])
def test_long_range_tdvp_windows_parity(ctx, oracle_lib, noisy):
    """long_range_method :TDVP (the reference default): the circuit sweep over
    the window of each long-range generator gate (window MPO of bond
    dimension 1, forward 2-site steps with 1-site back-evolution, edge step),
    per trajectory against the oracle's sweep.  The sweep projects onto the
    tangent space of the window, which depends on the basis a split picks for
    exactly zero singular values (the TDVP split keeps at least two): from a
    product state the reference's result depends on its SVD's null-space
    completion.  The run therefore starts from a random MPS whose bonds have
    full rank (dimension 4 inside), where the tangent spaces are fixed by the
    state."""
    print(f"case: digital long-range TDVP windows L=7 noisy={noisy}", flush=True)
    L = 7
    circ = _long_range_generator_circuit(L, 12, seed=31)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
          [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=12 if noisy else 1, max_bond_dim=16, truncation_threshold=1e-9)
    noise = _depolarizing(L, 0.03) if noisy else None
    g, gm, gb, st, r, rm, rb, rst = _run_opts(ctx, oracle_lib, L, circ, cfg, noise, Y.TJMOptions(), seed=9,
                                              mps=_generic_mps(L, 4, 17))
    assert st["jumps"] == rst["jumps"]
    if noisy:
        assert st["jumps"] > 0
    assert _err("g", g, r) < 1e-10
    assert np.array_equal(gb, rb)
