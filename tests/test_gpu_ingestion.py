"""An ingested circuit (src/CircuitIngestion.jl) run on the GPU.

The circuit of every mapped gate from `test_ingestion.every_gate_circuit`
(measures and barriers mixed in, 2-qubit gates listed in both site orders)
is ingested and run in FP64 through the HIP library: with TEBD gates it
equals the dense state vector of the ingested gate list, and with the
reference's default options (TDVP windows) and depolarizing noise it equals
the oracle per trajectory.
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi
from yaqs_julia_amd.ingestion import ingest_qiskit_circuit

from test_ingestion import dense_state, every_gate_circuit, z_expect

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def _gpu(ctx, L, circ, cfg, noise, options, seed=0):
    h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, precision=_abi.PREC_FP64, seed=seed, alg_options=options)
    ncol = int(ctx.lib.yaqs_digital_num_samples(h.job))
    g, gm, _, _ = ctx.wait(ctx.submit_digital(h), len(cfg.observables), ncol, cfg.num_traj)
    return g, gm


@pytest.mark.parametrize("L,seed", [(6, 1), (8, 2)])
def test_ingested_tebd_matches_statevector(ctx, L, seed):
    print(f"case: ingested circuit L={L} seed={seed} TEBD", flush=True)
    circ = ingest_qiskit_circuit(every_gate_circuit(L, seed))
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=1, max_bond_dim=2 ** (L // 2), truncation_threshold=0.0)
    _, gm = _gpu(ctx, L, circ, cfg, None, Y.TJMOptions("TEBD", "TEBD"))
    err = float(np.abs(gm[:, -1] - z_expect(L, dense_state(L, circ.gates))).max())
    print(f"  max|gpu-statevector| = {err:.3e}", flush=True)
    assert err < 1e-10


def test_ingested_default_options_noisy_matches_oracle(ctx, oracle_lib):
    print("case: ingested circuit L=6 TDVP windows + depolarizing", flush=True)
    L, ntraj = 6, 16
    circ = ingest_qiskit_circuit(every_gate_circuit(L, 3, swap=False))
    procs = [{"name": n, "sites": [i], "strength": 0.01} for i in range(1, L + 1)
             for n in ("pauli_x", "pauli_y", "pauli_z")]
    noise = Y.NoiseModel(procs, L)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=8, truncation_threshold=1e-12)
    g, _ = _gpu(ctx, L, circ, cfg, noise, Y.TJMOptions(), seed=4)
    h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, seed=4, alg_options=Y.TJMOptions())
    r, _, _, _ = oracle_lib.run_digital(h, nthreads=8)
    err = float(np.abs(g - r).max())
    print(f"  max|gpu-oracle| = {err:.3e}", flush=True)
    assert err < 1e-10
