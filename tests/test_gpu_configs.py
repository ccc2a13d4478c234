"""The benchmark configurations other than the headline one, on the GPU.

- C1 at full size: Ising L=6 (J=1, g=0.5) from |000000>, lowering + pauli_z
  gamma=0.1 on every site, dt=0.1, T=5, chi_max=16, threshold 1e-10, 100
  trajectories seeds 0..99, FP64: every trajectory against the oracle
  (1e-10), and the ensemble against the exact Lindblad solution within
  sampling error.
- C3: the L=40 Haar SU(4) brickwork, 40 layers, depolarizing p=0.01 as
  pauli_x/y/z with p/3 on each gate's qubits, TEBD threshold 1e-9,
  chi_max=128, 512 trajectories, mixed: finite <Z_i> in [-1, 1], bonds
  <= 128, concurrent waves equal one wave; FP64 trajectories 0..1 of the
  first 7 layers against the oracle (1e-10, bond dimensions equal).
- C4: jobs of the mixed-request generator (bench.c4_jobs, with the largest
  L=64 and chi=256 forced on one of them; trajectories and T cut to 64 and
  0.5 to fit a test) give the same results alone and in a concurrent mix.
- C5: Ising L=100 (J=1, g=1) from |0..0>, pauli_z gamma=0.02, dt=0.05,
  chi_max=512, threshold 1e-10, 256 trajectories, mixed, its first 2 steps
  under an HBM budget that forces several waves (the admission footprint is
  the worst case at chi=512, 2.6 GiB a trajectory): results equal the run
  under the default budget and trajectories 0..3 agree with the FP64
  oracle.
"""
import os
import sys

import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def _zx(L):
    return [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
           [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]


def test_c1_full_size(ctx, oracle_lib):
    print("case: C1 L=6 100 traj T=5 fp64", flush=True)
    from oracle import exact as E
    L = 6
    cfg = Y.TimeEvolutionConfig(_zx(L), 5.0, dt=0.1, num_traj=100, max_bond_dim=16, truncation_threshold=1e-10)
    H = Y.init_ising(L, 1.0, 0.5)
    noise = Y.relaxation_dephasing(L, 0.1)
    h = _abi.pack_analog(Y.MPS(L), H, cfg, noise, seed=0)
    jid = ctx.submit_analog(h)
    g, gm, _, st = ctx.wait(jid, len(cfg.observables), len(cfg.times), 100)
    r, rm, rst = oracle_lib.run_analog(_abi.pack_analog(Y.MPS(L), H, cfg, noise, seed=0), nthreads=8)
    d = np.abs(g - r).max(axis=(0, 1))
    e = float(d.max())
    print(f"  max|gpu-oracle| = {e:.3e}, median over trajectories {np.median(d):.3e}, "
          f"{int((d > 1e-10).sum())} above 1e-10; jumps {st['jumps']} / {int(rst['jumps'])}", flush=True)
    # trajectories a lowering jump leaves nearly a product state keep a
    # second singular value far below 1e-8 (the T-A rule keeps at least two):
    # the engine resolves its singular vectors with a deflated second Gram
    assert st["jumps"] == rst["jumps"] and e < 1e-10
    # ensemble against the exact Lindblad evolution
    jumps = [(0.1, E.site_op(E.LOWER, s, L)) for s in range(1, L + 1)] + \
            [(0.1, E.site_op(E.Z, s, L)) for s in range(1, L + 1)]
    psi = np.zeros(2 ** L, complex)
    psi[0] = 1
    ops = [E.site_op(E.Z, s, L) for s in range(1, L + 1)] + [E.site_op(E.X, s, L) for s in range(1, L + 1)]
    ex = E.lindblad(E.ising(L, 1.0, 0.5), jumps, np.outer(psi, psi.conj()), 0.1, 50, ops)
    # sampling error: the reference's own TJM (500 trajectories) sits up to
    # 4.3e-2 from its exact Lindblad columns (F3); 100 trajectories give
    # about sqrt(5) times that
    dev = np.abs(gm - ex)
    print(f"  ensemble vs Lindblad: max {dev.max():.3e}, mean {dev.mean():.3e}", flush=True)
    assert dev.max() < 0.1 and dev.mean() < 0.02


def _c3(ctx, ntraj, layers, precision, max_batch=0):
    import bench
    h, cfg = bench.c3_job(Y, ntraj, layers=layers, precision=precision)
    c = ctx if max_batch == 0 else Y.Context(0, max_batch=max_batch)
    try:
        ncol = int(c.lib.yaqs_digital_num_samples(h.job))
        jid = c.submit_digital(h)
        return c.wait(jid, len(cfg.observables), ncol, ntraj, want_bonds=True)
    finally:
        if c is not ctx:
            c.close()


def test_c3_full_size_properties_and_waves(ctx):
    print("case: C3 L=40 40 layers 512 traj mixed, concurrent waves vs one wave", flush=True)
    g, gm, gb, st = _c3(ctx, 512, 40, "mixed")
    print(f"  wall {st['wall_s']:.1f}s max bond {int(gb.max())} jumps {st['jumps']}", flush=True)
    assert np.all(np.isfinite(g)) and np.abs(g).max() <= 1 + 1e-6
    assert gb.max() <= 128 and gb.max() >= 64 and st["jumps"] > 0
    o, om, ob, _ = _c3(ctx, 512, 40, "mixed", max_batch=512)
    e = float(np.abs(g - o).max())
    print(f"  concurrent waves vs one wave: max {e:.3e}", flush=True)
    assert e < 1e-4 and np.abs(gm - om).max() < 1e-4


def test_c3_fp64_vs_oracle_first_layers(ctx, oracle_lib):
    print("case: C3 L=40 first 7 layers, trajectories 0..1 fp64 vs oracle", flush=True)
    import bench
    g, gm, gb, st = _c3(ctx, 2, 7, "fp64")
    h, _ = bench.c3_job(Y, 2, layers=7, precision="fp64")
    r, rm, rb, rst = oracle_lib.run_digital(h, nthreads=2)
    e = float(np.abs(g - r).max())
    print(f"  max|gpu-oracle| = {e:.3e}  bonds {gb.max(axis=1)} / {rb.max(axis=1)}", flush=True)
    assert np.array_equal(gb, rb) and e < 1e-10


def test_c4_jobs_same_alone_and_in_mix(ctx):
    print("case: C4 generator jobs alone vs in a mix", flush=True)
    import bench
    specs = bench.c4_jobs(Y, 4, seed=4, T=0.5, ntraj_range=(64, 64), force={"L": 64, "chi": 256, "model": 2})
    alone = []
    for (h, cfg, L, ntraj, desc) in specs:
        jid = ctx.submit_analog(h)
        alone.append(ctx.wait(jid, len(cfg.observables), len(cfg.times), ntraj))
        print(f"  {desc}: alone {alone[-1][3]['wall_s']:.2f}s", flush=True)
    ids = [ctx.submit_analog(h) for (h, _, _, _, _) in specs]
    for (h, cfg, L, ntraj, desc), jid, a in zip(specs, ids, alone):
        m = ctx.wait(jid, len(cfg.observables), len(cfg.times), ntraj)
        e = float(np.abs(m[0] - a[0]).max())
        print(f"  {desc}: in the mix {m[3]['wall_s']:.2f}s, max|mix - alone| = {e:.3e}", flush=True)
        assert e < 1e-4 and np.abs(m[1] - a[1]).max() < 1e-4


def _c5(ntraj, budget, T=0.15, precision=_abi.PREC_MIXED, num_traj_job=None):
    L = 100
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, T, dt=0.05, num_traj=ntraj, max_bond_dim=512, truncation_threshold=1e-10)
    h = _abi.pack_analog(Y.MPS(L), Y.init_ising(L, 1.0, 1.0), cfg, Y.NoiseModel(
        [{"name": "pauli_z", "sites": [i], "strength": 0.02} for i in range(1, L + 1)], L), precision=precision,
        seed=0)
    return h, cfg


def test_c5_full_size_window_under_budget(ctx, oracle_lib):
    print("case: C5 L=100 chi=512 256 traj mixed, 2 steps under a 128 GiB budget", flush=True)
    h, cfg = _c5(256, 0, T=0.1)
    per = ctx.lib.yaqs_job_footprint_bytes(h.job) / 256
    budget = 128 << 30
    assert per * 256 > 4 * budget - (8 << 30)  # the budget forces waves
    bctx = Y.Context(0, hbm_budget_bytes=budget)
    try:
        jid = bctx.submit_analog(h)
        g, gm, _, st = bctx.wait(jid, len(cfg.observables), len(cfg.times), 256)
    finally:
        bctx.close()
    print(f"  footprint {per / 2**30:.2f} GiB per trajectory, wall {st['wall_s']:.1f}s, max bond {st['max_bond']}",
          flush=True)
    assert np.all(np.isfinite(g)) and np.abs(g).max() <= 1 + 1e-6
    jid = ctx.submit_analog(h)
    u, um, _, _ = ctx.wait(jid, len(cfg.observables), len(cfg.times), 256)
    e = float(np.abs(g - u).max())
    print(f"  budget-limited waves vs unconstrained: max {e:.3e}", flush=True)
    assert e < 1e-4
    ho, _ = _c5(4, 0, T=0.1, precision=_abi.PREC_FP64)
    r, _, _ = oracle_lib.run_analog(ho, nthreads=4)
    eo = float(np.abs(g[:, :, :4] - r).max())
    print(f"  trajectories 0..3 vs FP64 oracle: max {eo:.3e}", flush=True)
    assert eo < 1e-4
