"""C-ABI library: loads, exports every symbol of include/yaqs_hip.h, and its
host-only entry points agree with the oracle (no GPU needed)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libyaqs_hip.so not built")
    return _abi.load_library()


def test_header_symbols_exported():
    lib = _lib()
    header = open(os.path.join(ROOT, "include", "yaqs_hip.h")).read()
    declared = set(re.findall(r"\b(yaqs_[a-z_0-9]+)\s*\(", header))
    assert declared == set(_abi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts():
    """ctypes mirrors match the C layout (natural alignment, 8-byte fields)."""
    assert ct.sizeof(_abi.MpsTensor) == 32
    assert ct.sizeof(_abi.MpoTensor) == 40
    assert ct.sizeof(_abi.NoiseProcess) == 40 + 256
    assert ct.sizeof(_abi.ObservableS) == 24 + 256
    assert ct.sizeof(_abi.Gate) == 24 + 256
    assert ct.sizeof(_abi.JobStats) == 64


def test_abi_version_and_rng_contract(oracle_lib):
    lib = _lib()
    assert lib.yaqs_abi_version() == 4
    for seed, t, k in [(0, 0, 0), (7, 3, 5), (2 ** 40 + 3, 999, 12), (123, 2 ** 33, 2 ** 35)]:
        assert lib.yaqs_rng_uniform(seed, t, k) == oracle_lib.rng_uniform(seed, t, k)


def test_num_samples_and_footprint():
    lib = _lib()
    L = 6
    obs = [Y.Observable("z1", Y.gates.ZGate(), 1)]
    cfg = Y.TimeEvolutionConfig(obs, 5.0, dt=0.1, num_traj=10, max_bond_dim=16)
    h = _abi.pack_analog(Y.MPS(L), Y.init_ising(L, 1.0, 0.5), cfg, Y.relaxation_dephasing(L, 0.1))
    assert lib.yaqs_analog_num_samples(ct.byref(h.job)) == 51
    fp = lib.yaqs_job_footprint_bytes(ct.byref(h.job))
    assert fp > 0
    h.job.num_traj = 20
    assert abs(lib.yaqs_job_footprint_bytes(ct.byref(h.job)) - 2 * fp) < 1e-6 * fp + 16
    c = Y.haar_brickwork(8, 6, seed=1, sample_every_layer=True, sample_at_start=True)
    hd = _abi.pack_digital(Y.MPS(8), c, Y.StrongMeasurementConfig(obs, num_traj=4), None)
    assert lib.yaqs_digital_num_samples(ct.byref(hd.job)) == 7


def test_invalid_jobs_rejected_without_device():
    lib = _lib()
    h = ct.c_void_p()
    rc = lib.yaqs_create(ct.byref(h), 0, None)
    if rc == 0:
        lib.yaqs_destroy(h)
        pytest.skip("a device is present")
    assert rc == _abi.YAQS_EDEVICE
    assert b"device" in lib.yaqs_last_error()
