"""Launch-variant tuner: the tile shape of each matrix-core GEMM and the
fused / unfused MPO step are chosen by measured speed per contraction
shape.  All variants give the same bits, so a job gives identical results
with the tuner off (default variants) and on (measured choices)."""
import ctypes as ct

import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu


def _job(prec):
    L = 8
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, 0.4, dt=0.05, num_traj=8, max_bond_dim=24, truncation_threshold=1e-10)
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    return _abi.pack_analog(Y.MPS(L, state="Neel"), H, cfg, Y.relaxation_dephasing(L, 0.05), precision=prec,
                            seed=3), cfg


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_tuner_on_off_identical(prec):
    print(f"case: tuner on/off prec={prec}", flush=True)
    lib = _abi.load_library()
    lib.yaqs_debug_set_tuning.argtypes = [ct.c_int]
    lib.yaqs_debug_tuning_entries.restype = ct.c_int
    out = []
    try:
        for on in (0, 1):
            lib.yaqs_debug_set_tuning(on)
            ctx = Y.Context(0)
            h, cfg = _job(prec)
            jid = ctx.submit_analog(h)
            out.append(ctx.wait(jid, len(cfg.observables), len(cfg.times), cfg.num_traj))
            ctx.close()
    finally:
        lib.yaqs_debug_set_tuning(1)
    assert lib.yaqs_debug_tuning_entries() > 0
    for k in (0, 1):
        a, b = out[0][k], out[1][k]
        diff = np.argwhere(a != b)
        assert diff.size == 0, (k, len(diff), diff[:8].tolist(), np.abs(a - b).max())
