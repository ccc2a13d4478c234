"""Failure handling and sharding through the C-ABI.

- NaN/Inf guard before a bond split (src/Algorithms.jl:414-417): a test hook
  stores NaN into one trajectory's two-site tensor before a chosen split.
  In mixed precision the trajectory is flagged and rerun alone in FP64: the
  job succeeds, reports one FP64 retry, the flagged trajectory equals the
  FP64 run's and every other trajectory is untouched (to FP32 rounding).  In FP64 the job
  fails with YAQS_ENUMERIC.  The reference replaces NaN/Inf by zeros there
  and goes on ("simulation may be compromised"); the engine does not return
  such results.
- Trajectory sharding (the multi-GPU layout): a job split into two
  traj_offset shards returns the whole job's trajectories (bit for bit
  where the shards are the whole job's waves).
"""
import ctypes as ct

import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    c.lib.yaqs_debug_poison_theta.argtypes = [ct.c_int64, ct.c_int64, ct.c_int]
    yield c
    c.lib.yaqs_debug_poison_theta(-1, -1, -1)
    c.close()


def _zx(L):
    return [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
           [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]


def _analog(ctx, precision, ntraj=12, traj_offset=0, L=8, seed=6):
    cfg = Y.TimeEvolutionConfig(_zx(L), 0.4, dt=0.05, num_traj=ntraj, max_bond_dim=16)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    h = _abi.pack_analog(Y.MPS(L, state="Neel"), H, cfg, Y.relaxation_dephasing(L, 0.1), precision=precision,
                         seed=seed, traj_offset=traj_offset)
    jid = ctx.submit_analog(h)
    return ctx.wait(jid, len(cfg.observables), len(cfg.times), ntraj)


def test_nan_guard_mixed_retries_in_fp64(ctx):
    print("case: NaN in trajectory 5 before split 40, mixed", flush=True)
    clean, _, _, _ = _analog(ctx, _abi.PREC_MIXED)
    ref64, _, _, _ = _analog(ctx, _abi.PREC_FP64)
    ctx.lib.yaqs_debug_poison_theta(5, 40, _abi.PREC_MIXED)
    try:
        got, _, _, st = _analog(ctx, _abi.PREC_MIXED)
    finally:
        ctx.lib.yaqs_debug_poison_theta(-1, -1, -1)
    print(f"  fp64 retries {st['fp64_retries']}", flush=True)
    assert st["fp64_retries"] == 1
    e5 = float(np.abs(got[:, :, 5] - ref64[:, :, 5]).max())
    others = [t for t in range(12) if t != 5]
    e_oth = float(np.abs(got[:, :, others] - clean[:, :, others]).max())
    print(f"  trajectory 5 vs FP64 run {e5:.3e}, others vs clean mixed run {e_oth:.3e}", flush=True)
    # the other trajectories agree with the clean run to FP32 rounding: equal
    # bits in most runs, 2.1e-7 in one of six (a timing-dependent difference
    # at the rounding level of the mixed path, cause not identified); a
    # corrupted trajectory would be off by far more
    assert e5 < 1e-10 and e_oth < 1e-5


def test_nan_guard_fp64_fails_with_enumeric(ctx):
    print("case: NaN in trajectory 2 before split 10, fp64", flush=True)
    ctx.lib.yaqs_debug_poison_theta(2, 10, _abi.PREC_FP64)
    try:
        with pytest.raises(Y.YaqsError, match="NaN/Inf") as ei:
            _analog(ctx, _abi.PREC_FP64)
    finally:
        ctx.lib.yaqs_debug_poison_theta(-1, -1, -1)
    assert ei.value.code == _abi.YAQS_ENUMERIC
    # the context keeps serving after the failed job
    _, mean, _, _ = _analog(ctx, _abi.PREC_FP64, ntraj=2)
    assert np.all(np.isfinite(mean))


def test_nan_guard_digital_mixed_retries(ctx):
    print("case: NaN in a digital trajectory, mixed", flush=True)
    L = 8
    circ = Y.haar_brickwork(L, 6, seed=2)
    cfg = Y.StrongMeasurementConfig([Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)],
                                    num_traj=6, max_bond_dim=16)
    procs = [{"name": n, "sites": [i], "strength": 0.01} for i in range(1, L + 1) for n in ("pauli_x", "pauli_z")]
    noise = Y.NoiseModel(procs, L)

    def run(prec):
        h = _abi.pack_digital(Y.MPS(L), circ, cfg, noise, precision=prec, seed=1)
        ncol = int(ctx.lib.yaqs_digital_num_samples(h.job))
        jid = ctx.submit_digital(h)
        return ctx.wait(jid, L, ncol, 6, want_bonds=True)

    ref64 = run(_abi.PREC_FP64)
    ctx.lib.yaqs_debug_poison_theta(1, 7, _abi.PREC_MIXED)
    try:
        got = run(_abi.PREC_MIXED)
    finally:
        ctx.lib.yaqs_debug_poison_theta(-1, -1, -1)
    assert got[3]["fp64_retries"] == 1
    e1 = float(np.abs(got[0][:, :, 1] - ref64[0][:, :, 1]).max())
    print(f"  trajectory 1 vs FP64 run {e1:.3e}", flush=True)
    assert e1 < 1e-10
    assert np.array_equal(got[2][1], ref64[2][1])


@pytest.mark.parametrize("precision,tol", [(_abi.PREC_FP64, 1e-10), (_abi.PREC_MIXED, 1e-4)])
def test_traj_offset_shards_equal_whole_job(ctx, precision, tol):
    """Each trajectory's draws depend on (seed, global index) only, so shards
    reproduce the whole job.  Where a shard is exactly one of the whole job's
    waves (a context with max_batch = 10) the bits agree; against the whole
    job run as one wave of 20 they agree to the rounding of the wider
    zero-padded working dimensions (batch invariance)."""
    print(f"case: two traj_offset shards vs the whole job, precision={precision}", flush=True)
    a, _, _, _ = _analog(ctx, precision, ntraj=10, traj_offset=0)
    b, _, _, _ = _analog(ctx, precision, ntraj=10, traj_offset=10)
    both = np.concatenate([a, b], axis=2)
    wctx = Y.Context(0, max_batch=10)
    try:
        waves, _, _, _ = _analog(wctx, precision, ntraj=20)
    finally:
        wctx.close()
    whole, _, _, _ = _analog(ctx, precision, ntraj=20)
    e_w = float(np.abs(both - waves).max())
    e = float(np.abs(both - whole).max())
    print(f"  max|shards - waves of 10| = {e_w:.3e}  max|shards - one wave of 20| = {e:.3e}", flush=True)
    assert e_w == 0.0
    assert e < tol
