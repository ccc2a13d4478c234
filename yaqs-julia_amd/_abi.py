"""ctypes mirror of include/yaqs_hip.h and the job packers.

The structures here match the C header field for field; `pack_analog` /
`pack_digital` turn the host-side Yaqs objects into a job struct whose
buffers stay alive as long as the returned holder.
"""
from __future__ import annotations

import ctypes as ct
import os
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libyaqs_hip.so")

YAQS_OK, YAQS_EINVAL, YAQS_ENOMEM_DEVICE, YAQS_EDEVICE, YAQS_ENUMERIC, YAQS_ENOTFOUND, YAQS_EAGAIN = range(7)
PREC_FP64, PREC_MIXED = 0, 1
GATE_1Q, GATE_2Q, GATE_SAMPLE = 1, 2, 3
GATE_EXACT, GATE_WINDOW_TDVP = 0, 1  # yaqs_gate_method
NOISE_KIND = {"lowering": 0, "raising": 1, "pauli_x": 2, "pauli_y": 3, "pauli_z": 4}
NOISE_CUSTOM = 5
NOISE_LONG_RANGE = 6

c_dp = ct.POINTER(ct.c_double)


class MpsTensor(ct.Structure):
    _fields_ = [("dl", ct.c_int64), ("d", ct.c_int64), ("dr", ct.c_int64), ("data", c_dp)]


class MpoTensor(ct.Structure):
    _fields_ = [("dl", ct.c_int64), ("dout", ct.c_int64), ("din", ct.c_int64), ("dr", ct.c_int64),
                ("data", c_dp)]


class NoiseProcess(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("nsites", ct.c_int32), ("site0", ct.c_int64), ("site1", ct.c_int64),
                ("strength", ct.c_double), ("pauli_name", ct.c_int32), ("_pad", ct.c_int32),
                ("matrix", ct.c_double * 32)]


class ObservableS(ct.Structure):
    _fields_ = [("nsites", ct.c_int32), ("_pad", ct.c_int32), ("site0", ct.c_int64), ("site1", ct.c_int64),
                ("op", ct.c_double * 32)]


class AnalogJob(ct.Structure):
    _fields_ = [("L", ct.c_int64), ("mps", ct.POINTER(MpsTensor)), ("mpo", ct.POINTER(MpoTensor)),
                ("has_noise_model", ct.c_int32), ("tdvp_sites", ct.c_int32), ("n_noise", ct.c_int64),
                ("noise", ct.POINTER(NoiseProcess)), ("n_obs", ct.c_int64), ("obs", ct.POINTER(ObservableS)),
                ("dt", ct.c_double), ("num_times", ct.c_int64), ("num_traj", ct.c_int64),
                ("max_bond_dim", ct.c_int64), ("truncation_threshold", ct.c_double), ("order", ct.c_int32),
                ("sample_timesteps", ct.c_int32), ("krylov_dim", ct.c_int32), ("precision", ct.c_int32),
                ("krylov_tol", ct.c_double), ("seed", ct.c_uint64), ("traj_offset", ct.c_uint64)]


class Gate(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("method", ct.c_int32), ("site0", ct.c_int64), ("site1", ct.c_int64),
                ("matrix", ct.c_double * 32)]


class DigitalJob(ct.Structure):
    _fields_ = [("L", ct.c_int64), ("mps", ct.POINTER(MpsTensor)), ("n_gates", ct.c_int64),
                ("gates", ct.POINTER(Gate)), ("has_noise_model", ct.c_int32), ("_pad0", ct.c_int32),
                ("n_noise", ct.c_int64), ("noise", ct.POINTER(NoiseProcess)), ("n_obs", ct.c_int64),
                ("obs", ct.POINTER(ObservableS)), ("num_traj", ct.c_int64), ("max_bond_dim", ct.c_int64),
                ("truncation_threshold", ct.c_double), ("precision", ct.c_int32), ("_pad1", ct.c_int32),
                ("seed", ct.c_uint64), ("traj_offset", ct.c_uint64)]


class JobStats(ct.Structure):
    _fields_ = [("wall_s", ct.c_double), ("queue_s", ct.c_double), ("flop_model", ct.c_double),
                ("matvecs", ct.c_int64), ("jumps", ct.c_int64), ("max_bond", ct.c_int64),
                ("num_samples", ct.c_int64), ("fp64_retries", ct.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BenchReport(ct.Structure):
    _fields_ = [("timed_s", ct.c_double), ("timed_steps", ct.c_int64), ("gemm_ms", ct.c_double),
                ("gemm_launches", ct.c_int64), ("gemm_flop", ct.c_double), ("flop_model", ct.c_double),
                ("matvecs", ct.c_int64), ("max_bond", ct.c_int64), ("concurrent_waves", ct.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CtxOpts(ct.Structure):
    _fields_ = [("hbm_budget_bytes", ct.c_int64), ("max_batch", ct.c_int32), ("verbose", ct.c_int32)]


def _cm_interleaved(a: np.ndarray) -> np.ndarray:
    """ComplexF64 column-major buffer as interleaved doubles."""
    f = np.asfortranarray(np.asarray(a, dtype=np.complex128))
    return np.ascontiguousarray(f.ravel(order="F")).view(np.float64)


def _fill_matrix(dst, m: np.ndarray) -> None:
    flat = _cm_interleaved(m)
    for i, v in enumerate(flat):
        dst[i] = v


class _Holder:
    """Keeps numpy buffers and ctypes arrays referenced by a job alive."""

    def __init__(self):
        self.keep: List[object] = []
        self.job = None


def _pack_mps(h: _Holder, mps):
    arr = (MpsTensor * mps.length)()
    for i, t in enumerate(mps.tensors):
        buf = _cm_interleaved(t)
        h.keep.append(buf)
        arr[i] = MpsTensor(t.shape[0], t.shape[1], t.shape[2], buf.ctypes.data_as(c_dp))
    h.keep.append(arr)
    return arr


def _pack_noise(h: _Holder, noise_model):
    procs = [] if noise_model is None else noise_model.processes
    arr = (NoiseProcess * max(1, len(procs)))()
    for i, p in enumerate(procs):
        q = arr[i]
        if hasattr(p, "mpo"):
            # MPONoiseProcess: O = a I + b sigma tau as (sigma, tau, a, b)
            p.check_dissipation_name()
            q.kind = NOISE_LONG_RANGE
            q.nsites = 2
            q.site0, q.site1 = min(p.sites), max(p.sites)
            q.strength = p.strength
            q.pauli_name = 1 if p.pauli_long_range() else 0
            vals = list(_cm_interleaved(p.sigma)) + list(_cm_interleaved(p.tau))
            vals += [complex(p.a).real, complex(p.a).imag, complex(p.b).real, complex(p.b).imag]
            for k, v in enumerate(vals):
                q.matrix[k] = v
            continue
        q.kind = NOISE_KIND.get(p.name, NOISE_CUSTOM) if len(p.sites) == 1 else NOISE_CUSTOM
        q.nsites = len(p.sites)
        q.site0 = p.sites[0]
        q.site1 = p.sites[1] if len(p.sites) == 2 else p.sites[0]
        q.strength = p.strength
        q.pauli_name = 1 if p.pauli_name else 0
        _fill_matrix(q.matrix, p.matrix)
    h.keep.append(arr)
    return arr, len(procs)


def _pack_obs(h: _Holder, observables):
    arr = (ObservableS * max(1, len(observables)))()
    for i, o in enumerate(observables):
        q = arr[i]
        q.nsites = len(o.sites)
        q.site0 = o.sites[0]
        q.site1 = o.sites[1] if len(o.sites) == 2 else o.sites[0]
        _fill_matrix(q.op, o.op.matrix)
    h.keep.append(arr)
    return arr, len(observables)


def pack_analog(mps, mpo, cfg, noise_model, *, precision: int = PREC_FP64, seed: int = 0, traj_offset: int = 0,
                krylov_dim: int = 25, krylov_tol: float = 1e-10, num_traj: Optional[int] = None,
                tdvp_sites: int = 2) -> _Holder:
    """`tdvp_sites` = 1 runs single_site_tdvp! where the reference's analog
    path calls two_site_tdvp! (bond dimensions then stay those of `mps`)."""
    h = _Holder()
    mps_arr = _pack_mps(h, mps)
    mpo_arr = (MpoTensor * mpo.length)()
    for i, w in enumerate(mpo.tensors):
        buf = _cm_interleaved(w)
        h.keep.append(buf)
        mpo_arr[i] = MpoTensor(w.shape[0], w.shape[1], w.shape[2], w.shape[3], buf.ctypes.data_as(c_dp))
    h.keep.append(mpo_arr)
    noise_arr, n_noise = _pack_noise(h, noise_model)
    obs_arr, n_obs = _pack_obs(h, cfg.observables)
    job = AnalogJob()
    job.L = mps.length
    job.mps = mps_arr
    job.mpo = mpo_arr
    job.has_noise_model = 0 if noise_model is None else 1
    job.n_noise = n_noise
    job.noise = noise_arr
    job.n_obs = n_obs
    job.obs = obs_arr
    job.dt = cfg.dt
    job.num_times = len(cfg.times)
    job.num_traj = cfg.num_traj if num_traj is None else num_traj
    job.max_bond_dim = cfg.max_bond_dim
    job.truncation_threshold = cfg.truncation_threshold
    job.order = cfg.order
    job.tdvp_sites = tdvp_sites
    job.sample_timesteps = 1 if cfg.sample_timesteps else 0
    job.krylov_dim = krylov_dim
    job.precision = precision
    job.krylov_tol = krylov_tol
    job.seed = seed
    job.traj_offset = traj_offset
    h.job = job
    return h


def pack_digital(mps, circuit, cfg, noise_model, *, precision: int = PREC_FP64, seed: int = 0,
                 traj_offset: int = 0, num_traj: Optional[int] = None, alg_options=None) -> _Holder:
    """`alg_options` (TJMOptions) picks how 2-qubit gates are applied; None
    applies every gate matrix (local_method TEBD)."""
    from .circuit import window_hamiltonian
    from .gates import Barrier
    tdvp = alg_options is not None and alg_options.local_method == "TDVP"
    tdvp_long = alg_options is not None and alg_options.long_range_method == "TDVP"
    h = _Holder()
    mps_arr = _pack_mps(h, mps)
    gates = circuit.gates
    garr = (Gate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        q = garr[i]
        if isinstance(g.op, Barrier):
            if g.op.label.upper() != "SAMPLE_OBSERVABLES":
                raise NotImplementedError("only SAMPLE_OBSERVABLES barriers are supported")
            q.kind = GATE_SAMPLE
            continue
        q.kind = GATE_1Q if len(g.sites) == 1 else GATE_2Q
        q.site0 = g.sites[0]
        q.site1 = g.sites[1] if len(g.sites) == 2 else g.sites[0]
        long_range = q.kind == GATE_2Q and abs(q.site1 - q.site0) > 1
        if q.kind == GATE_2Q and (tdvp_long if long_range else tdvp):
            q.method = GATE_WINDOW_TDVP
            _fill_matrix(q.matrix, window_hamiltonian(g))
        else:
            _fill_matrix(q.matrix, g.op.matrix)
    h.keep.append(garr)
    noise_arr, n_noise = _pack_noise(h, noise_model)
    obs_arr, n_obs = _pack_obs(h, cfg.observables)
    job = DigitalJob()
    job.L = mps.length
    job.mps = mps_arr
    job.n_gates = len(gates)
    job.gates = garr
    job.has_noise_model = 0 if noise_model is None else 1
    job.n_noise = n_noise
    job.noise = noise_arr
    job.n_obs = n_obs
    job.obs = obs_arr
    job.num_traj = cfg.num_traj if num_traj is None else num_traj
    job.max_bond_dim = cfg.max_bond_dim
    job.truncation_threshold = cfg.truncation_threshold
    job.precision = precision
    job.seed = seed
    job.traj_offset = traj_offset
    h.job = job
    return h


_lib = None


def load_library(path: Optional[str] = None):
    """Load the HIP C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libyaqs_hip.so not built ({p}); run __graft_entry__.build()")
    lib = ct.CDLL(p)
    lib.yaqs_abi_version.restype = ct.c_int
    lib.yaqs_last_error.restype = ct.c_char_p
    lib.yaqs_create.argtypes = [ct.POINTER(ct.c_void_p), ct.c_int, ct.POINTER(CtxOpts)]
    lib.yaqs_destroy.argtypes = [ct.c_void_p]
    lib.yaqs_submit_analog.argtypes = [ct.c_void_p, ct.POINTER(AnalogJob), ct.POINTER(ct.c_uint64)]
    lib.yaqs_submit_digital.argtypes = [ct.c_void_p, ct.POINTER(DigitalJob), ct.POINTER(ct.c_uint64)]
    wait_args = [ct.c_void_p, ct.c_uint64, c_dp, c_dp, ct.POINTER(ct.c_int64), ct.POINTER(JobStats)]
    lib.yaqs_wait.argtypes = wait_args
    lib.yaqs_poll.argtypes = wait_args
    lib.yaqs_run_analog.argtypes = [ct.c_void_p, ct.POINTER(AnalogJob), c_dp, c_dp, ct.POINTER(JobStats)]
    lib.yaqs_run_digital.argtypes = [ct.c_void_p, ct.POINTER(DigitalJob), c_dp, c_dp, ct.POINTER(ct.c_int64),
                                     ct.POINTER(JobStats)]
    lib.yaqs_analog_num_samples.argtypes = [ct.POINTER(AnalogJob)]
    lib.yaqs_analog_num_samples.restype = ct.c_int64
    lib.yaqs_digital_num_samples.argtypes = [ct.POINTER(DigitalJob)]
    lib.yaqs_digital_num_samples.restype = ct.c_int64
    lib.yaqs_job_footprint_bytes.argtypes = [ct.POINTER(AnalogJob)]
    lib.yaqs_job_footprint_bytes.restype = ct.c_int64
    lib.yaqs_engine_device_bytes.argtypes = [ct.c_int, ct.POINTER(AnalogJob), ct.c_int64, ct.POINTER(ct.c_int64),
                                             ct.POINTER(ct.c_int64)]
    lib.yaqs_bench_analog.argtypes = [ct.c_void_p, ct.POINTER(AnalogJob), ct.c_int64, c_dp, c_dp,
                                      ct.POINTER(BenchReport)]
    lib.yaqs_run_analog_noalloc.argtypes = [ct.c_void_p, ct.POINTER(AnalogJob), c_dp, c_dp, ct.POINTER(JobStats)]
    lib.yaqs_noalloc_scratch_bytes.restype = ct.c_int64
    lib.yaqs_measure_shots.argtypes = [ct.c_void_p, ct.c_int64, ct.POINTER(MpsTensor), ct.c_int64, ct.c_uint64,
                                       ct.POINTER(ct.c_int64)]
    lib.yaqs_rng_uniform.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_uint64]
    lib.yaqs_rng_uniform.restype = ct.c_double
    if path is None:
        _lib = lib
    return lib


EXPORTED_SYMBOLS = [
    "yaqs_abi_version", "yaqs_last_error", "yaqs_create", "yaqs_destroy", "yaqs_submit_analog",
    "yaqs_submit_digital", "yaqs_wait", "yaqs_poll", "yaqs_run_analog", "yaqs_run_digital",
    "yaqs_analog_num_samples", "yaqs_digital_num_samples", "yaqs_job_footprint_bytes", "yaqs_bench_analog",
    "yaqs_rng_uniform", "yaqs_run_analog_noalloc", "yaqs_noalloc_scratch_bytes", "yaqs_measure_shots",
    "yaqs_engine_device_bytes",
]
