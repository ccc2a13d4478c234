"""Simulator.run — the drop-in entry point (src/Simulator.jl:127-141).

The whole trajectory loop of `_run_analog` / `_run_digital`
(src/Simulator.jl:22-125) becomes one C-ABI call into libyaqs_hip.so, which
runs every trajectory batched on the GPU.  There is no CPU fallback: if the
library is missing or the device fails, this raises.
"""
from __future__ import annotations

import ctypes as ct
import threading
from typing import Optional, Union

import numpy as np

from . import _abi
from .circuit import DigitalCircuit, TJMOptions
from .configs import StrongMeasurementConfig, TimeEvolutionConfig
from .networks import MPO, MPS
from .noise import NoiseModel

_PREC = {"fp64": _abi.PREC_FP64, "mixed": _abi.PREC_MIXED}


class YaqsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"yaqs error {code}: {msg}")
        self.code = code


class Context:
    """One device context (yaqs_create/yaqs_destroy)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_batch: int = 0, verbose: int = 0):
        self.lib = _abi.load_library()
        opts = _abi.CtxOpts(hbm_budget_bytes, max_batch, verbose)
        h = ct.c_void_p()
        self._check(self.lib.yaqs_create(ct.byref(h), device, ct.byref(opts)))
        self.handle = h

    def _check(self, rc: int) -> None:
        if rc != _abi.YAQS_OK:
            raise YaqsError(rc, (self.lib.yaqs_last_error() or b"").decode())

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.yaqs_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def measure_shots_bits(self, mps: MPS, shots: int, seed: int = 0) -> np.ndarray:
        """yaqs_measure_shots: per-shot bitstrings (site i in bit i-1)."""
        h = _abi._Holder()
        arr = _abi._pack_mps(h, mps)
        out = np.zeros(max(1, shots), dtype=np.int64)
        self._check(self.lib.yaqs_measure_shots(self.handle, mps.length, arr, shots, seed,
                                                out.ctypes.data_as(ct.POINTER(ct.c_int64))))
        return out[:shots]

    # -- submission --------------------------------------------------------
    def submit_analog(self, holder) -> int:
        jid = ct.c_uint64()
        self._check(self.lib.yaqs_submit_analog(self.handle, ct.byref(holder.job), ct.byref(jid)))
        return jid.value

    def submit_digital(self, holder) -> int:
        jid = ct.c_uint64()
        self._check(self.lib.yaqs_submit_digital(self.handle, ct.byref(holder.job), ct.byref(jid)))
        return jid.value

    def _out(self, n_obs, ncol, ntraj, want_bonds):
        traj = np.zeros(max(1, n_obs * ncol * ntraj))
        mean = np.zeros(max(1, n_obs * ncol))
        bonds = np.zeros(max(1, ntraj * ncol), dtype=np.int64) if want_bonds else None
        return traj, mean, bonds

    def wait(self, jid: int, n_obs: int, ncol: int, ntraj: int, want_bonds: bool = False, poll: bool = False):
        traj, mean, bonds = self._out(n_obs, ncol, ntraj, want_bonds)
        st = _abi.JobStats()
        fn = self.lib.yaqs_poll if poll else self.lib.yaqs_wait
        rc = fn(self.handle, jid, traj.ctypes.data_as(_abi.c_dp), mean.ctypes.data_as(_abi.c_dp),
                bonds.ctypes.data_as(ct.POINTER(ct.c_int64)) if bonds is not None else None, ct.byref(st))
        if poll and rc == _abi.YAQS_EAGAIN:
            return None
        self._check(rc)
        traj = traj[: n_obs * ncol * ntraj].reshape(n_obs, ncol, ntraj)
        mean = mean[: n_obs * ncol].reshape(n_obs, ncol)
        if bonds is not None:
            bonds = bonds[: ntraj * ncol].reshape(ntraj, ncol)
        return traj, mean, bonds, st.as_dict()


_default_ctx: Optional[Context] = None
_ctx_lock = threading.Lock()


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def _all_zero(noise_model: Optional[NoiseModel]) -> bool:
    return noise_model is None or all(p.strength == 0 for p in noise_model.processes)


def run(initial_state: MPS, operator_or_circuit: Union[MPO, DigitalCircuit], sim_params,
        noise_model: Optional[NoiseModel] = None, *, precision: str = "fp64", seed: int = 0,
        ctx: Optional[Context] = None, parallel: bool = True, alg_options: Optional[TJMOptions] = None,
        tdvp_sites: int = 2):
    """Simulator.run(ψ0, H_or_circuit, cfg, noise; alg_options).  Fills
    `obs.trajectories` (num_traj x num_cols, real parts stored as complex)
    and `obs.results` for every observable, as src/Simulator.jl:45-68 does.
    Returns job statistics (analog) or per-trajectory bond dimensions
    (digital, src/Simulator.jl:82-124).  Digital runs default to
    TJMOptions() (TDVP windows, src/DigitalTJM.jl:366), as the reference's
    run_digital_tjm does.  Analog runs with `tdvp_sites=1` evolve with
    single_site_tdvp! (src/Algorithms.jl:449-452) where the reference calls
    two_site_tdvp!; the bond dimensions then stay those of the initial
    state."""
    initial_state.normalize()  # src/Simulator.jl:128
    ctx = ctx or default_context()
    prec = _PREC[precision]
    if isinstance(sim_params, TimeEvolutionConfig) and isinstance(operator_or_circuit, MPO):
        if _all_zero(noise_model):
            sim_params.num_traj = 1  # src/Simulator.jl:27-29
        h = _abi.pack_analog(initial_state, operator_or_circuit, sim_params, noise_model, precision=prec,
                             seed=seed, tdvp_sites=tdvp_sites)
        ncol = len(sim_params.times) if sim_params.sample_timesteps else 1
        for o in sim_params.observables:
            o.initialize(sim_params.num_traj, ncol)
        jid = ctx.submit_analog(h)
        traj, mean, _, stats = ctx.wait(jid, len(sim_params.observables), ncol, sim_params.num_traj)
        for q, o in enumerate(sim_params.observables):
            o.trajectories[:, :] = traj[q].T
            o.results = mean[q].copy()
        return stats
    if isinstance(operator_or_circuit, DigitalCircuit) and isinstance(
            sim_params, (TimeEvolutionConfig, StrongMeasurementConfig)):
        if _all_zero(noise_model):
            sim_params.num_traj = 1
        h = _abi.pack_digital(initial_state, operator_or_circuit, sim_params, noise_model, precision=prec,
                              seed=seed, alg_options=alg_options if alg_options is not None else TJMOptions())
        ncol = int(ctx.lib.yaqs_digital_num_samples(ct.byref(h.job)))
        for o in sim_params.observables:
            o.initialize(sim_params.num_traj, ncol)
        jid = ctx.submit_digital(h)
        traj, mean, bonds, stats = ctx.wait(jid, len(sim_params.observables), ncol, sim_params.num_traj,
                                            want_bonds=True)
        for q, o in enumerate(sim_params.observables):
            o.trajectories[:, :] = traj[q].T
            o.results = mean[q].copy()
        return [list(b) for b in bonds]
    raise TypeError("Simulation requires MPO (Analog) or DigitalCircuit (Digital).")


def measure_shots(mps: MPS, shots: int, *, seed: int = 0, ctx: Optional[Context] = None) -> dict:
    """measure_shots(mps, shots) (src/MPS.jl:742-749): counts of the sampled
    bitstrings (site i in bit i-1), sampled on the device by
    yaqs_measure_shots.  Shot t draws the counter-RNG uniforms (seed, t,
    site); the reference draws Julia's global rand()."""
    ctx = ctx or default_context()
    bits = ctx.measure_shots_bits(mps, shots, seed)
    vals, counts = np.unique(bits, return_counts=True)
    return {int(v): int(c) for v, c in zip(vals, counts)}


def single_shot_measure(mps: MPS, *, seed: int = 0, shot: int = 0, ctx: Optional[Context] = None) -> int:
    """single_shot_measure(mps) (src/MPS.jl:683-738): one bitstring, the
    shot `shot` of the stream under `seed`."""
    ctx = ctx or default_context()
    if shot == 0:
        return int(ctx.measure_shots_bits(mps, 1, seed)[0])
    return int(ctx.measure_shots_bits(mps, shot + 1, seed)[shot])


def run_shots(mps: MPS, config, *, seed: int = 0, ctx: Optional[Context] = None) -> dict:
    """Fills a MeasurementConfig: `config.measurements` holds one
    {bitstring: 1} dict per shot and `config.results` their aggregate
    (aggregate_measurements!, src/SimulationConfigs.jl:198-209)."""
    ctx = ctx or default_context()
    bits = ctx.measure_shots_bits(mps, config.shots, seed)
    config.measurements = [{int(b): 1} for b in bits]
    config.aggregate_measurements()
    return config.results
