"""Observable and simulation configurations (src/SimulationConfigs.jl:20-182)."""
from __future__ import annotations

from typing import List, Sequence, Union

import numpy as np

from .gates import AbstractOperator


class Observable:
    """Observable(name, op, sites) with per-trajectory and mean buffers
    (src/SimulationConfigs.jl:20-34)."""

    def __init__(self, name: str, op: AbstractOperator, sites: Union[int, Sequence[int]]):
        self.name = name
        self.op = op
        self.sites = [int(sites)] if np.ndim(sites) == 0 else [int(s) for s in sites]
        self.results = np.zeros(0)
        self.trajectories = np.zeros((0, 0), dtype=np.complex128)

    def initialize(self, num_traj: int, num_cols: int) -> None:
        """initialize! (src/SimulationConfigs.jl:146-154)."""
        self.trajectories = np.zeros((num_traj, num_cols), dtype=np.complex128)
        self.results = np.zeros(num_cols)


class TimeEvolutionConfig:
    """TimeEvolutionConfig (src/SimulationConfigs.jl:68-92).  `times` follows
    Julia's `collect(0.0:dt:T)`: n = floor(T/dt + tiny) + 1 points."""

    def __init__(self, observables: List[Observable], total_time: float, *, dt: float = 0.1,
                 num_traj: int = 1000, max_bond_dim: int = 4096, min_bond_dim: int = 2,
                 truncation_threshold: float = 1e-9, sample_timesteps: bool = True, order: int = 2):
        self.observables = list(observables)
        self.total_time = float(total_time)
        self.dt = float(dt)
        n = int(np.floor(self.total_time / self.dt * (1 + 1e-12) + 1e-9)) + 1
        self.times = np.arange(n) * self.dt
        self.num_traj = int(num_traj)
        self.max_bond_dim = int(max_bond_dim)
        self.min_bond_dim = int(min_bond_dim)
        self.truncation_threshold = float(truncation_threshold)
        self.sample_timesteps = bool(sample_timesteps)
        self.order = int(order)

    def aggregate_trajectories(self) -> None:
        """aggregate_trajectories! (src/SimulationConfigs.jl:168-182)."""
        for obs in self.observables:
            if obs.trajectories.size:
                obs.results = obs.trajectories.real.sum(axis=0) / self.num_traj


class MeasurementConfig:
    """MeasurementConfig (src/SimulationConfigs.jl:99-114): shots, bond
    limits, per-shot measurement dicts and their aggregate."""

    def __init__(self, shots: int, *, max_bond_dim: int = 4096, min_bond_dim: int = 2,
                 truncation_threshold: float = 1e-9):
        self.shots = int(shots)
        self.max_bond_dim = int(max_bond_dim)
        self.min_bond_dim = int(min_bond_dim)
        self.truncation_threshold = float(truncation_threshold)
        self.measurements = [None] * self.shots
        self.results = {}

    def aggregate_measurements(self) -> None:
        """aggregate_measurements! (src/SimulationConfigs.jl:198-209)."""
        total = {}
        for m in self.measurements:
            if m is not None:
                for k, v in m.items():
                    total[k] = total.get(k, 0) + v
        self.results = total


class StrongMeasurementConfig:
    """StrongMeasurementConfig (src/SimulationConfigs.jl:121-137), the
    configuration of digital runs."""

    def __init__(self, observables: List[Observable], *, num_traj: int = 1000, max_bond_dim: int = 4096,
                 min_bond_dim: int = 2, truncation_threshold: float = 1e-9, sample_layers: bool = False):
        self.observables = list(observables)
        self.num_traj = int(num_traj)
        self.max_bond_dim = int(max_bond_dim)
        self.min_bond_dim = int(min_bond_dim)
        self.truncation_threshold = float(truncation_threshold)
        self.sample_layers = bool(sample_layers)

    def aggregate_trajectories(self) -> None:
        for obs in self.observables:
            if obs.trajectories.size:
                obs.results = obs.trajectories.real.sum(axis=0) / self.num_traj
