"""Writes oracle-generated fixtures for the GPU parity tests at sizes the
oracle cannot finish inside a test's time limit.

c2_oracle_traj0-3.npy: the FP64 oracle's <Z_i>, <X_i> (i = 1..30) of
trajectories 0..3 (job seed 0) of the Heisenberg XXX workload — L=30, Neel
start, lowering + pauli_z gamma=0.01 on every site, 2-site TDVP dt=0.05,
24 steps (25 sample times, bonds saturate at chi_max=64 by the last
steps), truncation threshold 1e-9, Krylov dimension 25.  Shape
[60 observables, 25 times, 4 trajectories], float64.  About three minutes
on four cores.  tests/test_oracle_tjm.py::test_c2_fixture_reproduced reruns
trajectory 0 for its first steps against this file, so the fixture stays
tied to the oracle.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def c2_job(L=30, steps=24, ntraj=4):
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import conftest  # noqa: F401  (registers the package)
    import yaqs_julia_amd as Y
    from yaqs_julia_amd import _abi
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
          [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, steps * 0.05, dt=0.05, num_traj=ntraj, max_bond_dim=64,
                                truncation_threshold=1e-9)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    noise = Y.relaxation_dephasing(L, 0.01)
    return Y, _abi, Y.MPS(L, state="Neel"), H, cfg, noise


def main():
    Y, _abi, mps, H, cfg, noise = c2_job()
    from oracle import oracle as O
    h = _abi.pack_analog(mps, H, cfg, noise, seed=0)
    traj, _, _ = O.run_analog(h, nthreads=4)
    np.save(os.path.join(HERE, "c2_oracle_traj0-3.npy"), traj)


if __name__ == "__main__":
    main()
