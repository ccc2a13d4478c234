"""Serving: jobs submitted together share the device (continuous batching of
trajectory waves) and each result is the same as the job run alone.

Tolerances: FP64 jobs 1e-10 per <O>(t) (a trajectory's result does not
depend on its wave mates, only the zero padding of the wave may change the
rounding order); mixed-precision jobs 1e-4.
"""
import time

import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu


def _zx(L):
    return [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
           [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]


def _jobs():
    """Different models, sizes, noise, precisions and one digital job."""
    out = []
    L = 6
    cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.1, num_traj=24, max_bond_dim=16, truncation_threshold=1e-10)
    out.append(("analog", Y.MPS(L), Y.init_ising(L, 1.0, 0.5), cfg, Y.relaxation_dephasing(L, 0.1),
                _abi.PREC_FP64, 11, 1e-10))
    L = 8
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.05, num_traj=16, max_bond_dim=8, truncation_threshold=1e-9)
    out.append(("analog", Y.MPS(L, state="Neel"), Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0), cfg,
                Y.relaxation_dephasing(L, 0.05), _abi.PREC_FP64, 5, 1e-10))
    L = 10
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.5, dt=0.05, num_traj=20, max_bond_dim=16, truncation_threshold=1e-9)
    out.append(("analog", Y.MPS(L, state="x+"), Y.init_general_hamiltonian(L, 1, 1, 0, 0, 0, 0), cfg,
                Y.relaxation_dephasing(L, 0.02), _abi.PREC_MIXED, 3, 1e-4))
    L = 6
    circ = Y.haar_brickwork(L, 6, seed=2)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=12, max_bond_dim=16, truncation_threshold=1e-9)
    noise = Y.NoiseModel([{"name": n, "sites": [i], "strength": 0.02} for i in range(1, L + 1)
                          for n in ("pauli_x", "pauli_z")], L)
    out.append(("digital", Y.MPS(L), circ, cfg, noise, _abi.PREC_FP64, 9, 1e-10))
    return out


def _submit(ctx, job):
    kind, mps, op, cfg, noise, prec, seed, _ = job
    if kind == "analog":
        h = _abi.pack_analog(mps, op, cfg, noise, precision=prec, seed=seed)
        ncol = len(cfg.times) if cfg.sample_timesteps else 1
        return h, ctx.submit_analog(h), ncol
    h = _abi.pack_digital(mps, op, cfg, noise, precision=prec, seed=seed)
    ncol = int(ctx.lib.yaqs_digital_num_samples(h.job))
    return h, ctx.submit_digital(h), ncol


def _collect(ctx, job, jid, ncol):
    cfg = job[3]
    return ctx.wait(jid, len(cfg.observables), ncol, cfg.num_traj, want_bonds=job[0] == "digital")


def test_same_result_alone_and_in_mix():
    print("case: serving mix of 4 jobs vs alone", flush=True)
    jobs = _jobs()
    alone = []
    ctx = Y.Context(0)
    for job in jobs:
        h, jid, ncol = _submit(ctx, job)
        alone.append(_collect(ctx, job, jid, ncol))
    ctx.close()
    # small waves so every job is split into several waves that interleave
    mix = Y.Context(0, max_batch=4)
    handles = [_submit(mix, job) for job in jobs]
    got = [_collect(mix, job, jid, ncol) for job, (h, jid, ncol) in zip(jobs, handles)]
    mix.close()
    for job, a, g in zip(jobs, alone, got):
        tol = job[7]
        assert a[0].shape == g[0].shape
        assert np.abs(a[0] - g[0]).max() < tol, (job[0], np.abs(a[0] - g[0]).max())
        assert np.abs(a[1] - g[1]).max() < tol
        if job[0] == "digital":
            assert np.array_equal(a[2], g[2])


def test_late_small_job_overtakes_long_job():
    """A small job submitted behind a long one starts on the next free worker
    and finishes first (its waves interleave with the long job's)."""
    print("case: serving late small job", flush=True)
    L = 8
    big_cfg = Y.TimeEvolutionConfig(_zx(L), 1.0, dt=0.05, num_traj=48, max_bond_dim=16)
    small_cfg = Y.TimeEvolutionConfig(_zx(4), 0.2, dt=0.05, num_traj=4, max_bond_dim=4)
    ctx = Y.Context(0, max_batch=6)
    hb = _abi.pack_analog(Y.MPS(L, state="Neel"), Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0), big_cfg,
                          Y.relaxation_dephasing(L, 0.05), seed=1)
    jb = ctx.submit_analog(hb)
    time.sleep(0.05)
    hs = _abi.pack_analog(Y.MPS(4), Y.init_ising(4, 1.0, 0.5), small_cfg, Y.relaxation_dephasing(4, 0.1), seed=2)
    js = ctx.submit_analog(hs)
    rs = ctx.wait(js, 8, len(small_cfg.times), 4)
    rb = ctx.wait(jb, 2 * L, len(big_cfg.times), 48)
    ctx.close()
    t_small, t_big = rs[3]["wall_s"], rb[3]["wall_s"]
    assert t_small < t_big, (t_small, t_big)
    assert np.isfinite(rb[1]).all() and np.isfinite(rs[1]).all()


def test_concurrent_waves_match_one_wave():
    """A job that fits one wave runs as concurrent waves of at least 256
    trajectories (three at most) on the context's workers; every
    trajectory's result equals the one-wave run (max_batch = num_traj) to
    the padding-order tolerance (FP64 1e-10), and the job statistics count
    the same trajectories."""
    print("case: 768 trajectories as concurrent waves vs one wave", flush=True)
    L = 6
    ntraj = 768
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.3, dt=0.05, num_traj=ntraj, max_bond_dim=8, truncation_threshold=1e-10)
    H = Y.init_ising(L, 1.0, 0.5)
    out = []
    for mb in (0, ntraj):
        ctx = Y.Context(0, max_batch=mb)
        h = _abi.pack_analog(Y.MPS(L), H, cfg, Y.relaxation_dephasing(L, 0.1), precision=_abi.PREC_FP64, seed=21)
        out.append(ctx.wait(ctx.submit_analog(h), len(cfg.observables), len(cfg.times), ntraj))
        ctx.close()
    split, one = out
    assert np.isfinite(split[0]).all()
    assert np.abs(split[0] - one[0]).max() < 1e-10, np.abs(split[0] - one[0]).max()
    assert np.abs(split[1] - one[1]).max() < 1e-10
    assert split[3]["jumps"] == one[3]["jumps"]


def test_digital_concurrent_waves_match_one_wave():
    """The same for a digital job (512 trajectories: two concurrent waves):
    results, bond dimensions and jumps equal the one-wave run."""
    print("case: digital 512 trajectories as concurrent waves vs one wave", flush=True)
    L = 6
    ntraj = 512
    circ = Y.haar_brickwork(L, 4, seed=5)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=8, truncation_threshold=1e-9)
    noise = Y.NoiseModel([{"name": n, "sites": [i], "strength": 0.02} for i in range(1, L + 1)
                          for n in ("pauli_x", "pauli_z")], L)
    out = []
    for mb in (0, ntraj):
        ctx = Y.Context(0, max_batch=mb)
        h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, precision=_abi.PREC_FP64, seed=13)
        ncol = int(ctx.lib.yaqs_digital_num_samples(h.job))
        out.append(ctx.wait(ctx.submit_digital(h), len(obs), ncol, ntraj, want_bonds=True))
        ctx.close()
    split, one = out
    assert np.abs(split[0] - one[0]).max() < 1e-10, np.abs(split[0] - one[0]).max()
    assert np.abs(split[1] - one[1]).max() < 1e-10
    assert np.array_equal(split[2], one[2])
    assert split[3]["jumps"] == one[3]["jumps"]
