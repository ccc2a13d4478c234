"""The allocation-free analog entry point (yaqs_run_analog_noalloc) against
yaqs_run_analog, through the C-ABI: two calls in a row on the device's
scratch array, the second with a smaller job (the scratch keeps the first
call's contents), FP64 within 1e-10 per <O>(t) and mixed precision within
1e-4; a job of which one trajectory exceeds the scratch is rejected."""
import ctypes as ct

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


def _job(L, ntraj, T, chi, precision, seed):
    cfg = Y.TimeEvolutionConfig(_zx(L), T, dt=0.05, num_traj=ntraj, max_bond_dim=chi, truncation_threshold=1e-10)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0) if L % 2 == 0 else Y.init_ising(L, 1.0, 0.6)
    noise = Y.relaxation_dephasing(L, 0.1)
    h = _abi.pack_analog(Y.MPS(L, state="Neel"), H, cfg, noise, precision=precision, seed=seed)
    return h, cfg


def _noalloc(ctx, h, cfg):
    nobs, ncol, ntraj = len(cfg.observables), len(cfg.times), cfg.num_traj
    traj = np.full(nobs * ncol * ntraj, np.nan)
    mean = np.full(nobs * ncol, np.nan)
    st = _abi.JobStats()
    rc = ctx.lib.yaqs_run_analog_noalloc(ctx.handle, ct.byref(h.job), traj.ctypes.data_as(_abi.c_dp),
                                         mean.ctypes.data_as(_abi.c_dp), ct.byref(st))
    ctx._check(rc)
    return traj.reshape(nobs, ncol, ntraj), mean.reshape(nobs, ncol), st.as_dict()


def _regular(ctx, h, cfg):
    jid = ctx.submit_analog(h)
    traj, mean, _, st = ctx.wait(jid, len(cfg.observables), len(cfg.times), cfg.num_traj)
    return traj, mean, st


@pytest.mark.parametrize("precision,tol", [(_abi.PREC_FP64, 1e-10), (_abi.PREC_MIXED, 1e-4)])
def test_noalloc_matches_run_analog(ctx, precision, tol):
    for (L, ntraj, T, chi, seed) in ((10, 40, 0.5, 16, 3), (6, 9, 0.3, 8, 7)):  # second call smaller
        print(f"case: noalloc precision={precision} L={L} ntraj={ntraj}", flush=True)
        h, cfg = _job(L, ntraj, T, chi, precision, seed)
        a, am, ast = _noalloc(ctx, h, cfg)
        b, bm, bst = _regular(ctx, h, cfg)
        err = float(np.abs(a - b).max())
        print(f"  max|noalloc - run_analog| = {err:.3e}  jumps {ast['jumps']} / {bst['jumps']}", flush=True)
        assert np.all(np.isfinite(a))
        assert err < tol
        assert np.abs(am - bm).max() < tol
        assert ast["jumps"] == bst["jumps"] and ast["max_bond"] == bst["max_bond"]


def test_noalloc_rejects_jobs_beyond_scratch(ctx):
    print("case: noalloc rejects a job larger than the scratch", flush=True)
    cap = ctx.lib.yaqs_noalloc_scratch_bytes()
    assert cap == 256 << 20
    L = 100
    obs = [Y.Observable("z1", Y.gates.ZGate(), 1)]
    cfg = Y.TimeEvolutionConfig(obs, 0.1, dt=0.05, num_traj=4, max_bond_dim=512)
    h = _abi.pack_analog(Y.MPS(L), Y.init_ising(L, 1.0, 1.0), cfg, Y.relaxation_dephasing(L, 0.02))
    assert ctx.lib.yaqs_job_footprint_bytes(ct.byref(h.job)) > 4 * cap
    with pytest.raises(Y.YaqsError, match="scratch"):
        _noalloc(ctx, h, cfg)
