"""yaqs-julia_amd — MI355X TJM trajectory engine behind the Yaqs API.

Host-side mirror of Yaqs.jl's simulator surface (MPS/MPO, NoiseModel,
Observable, TimeEvolutionConfig, Simulator.run).  All trajectory work runs in
the HIP C-ABI library `build/libyaqs_hip.so`; this package only builds job
descriptions and reads results back.
"""
from . import gates
from .circuit import DigitalCircuit, TJMOptions, haar_brickwork
from .configs import MeasurementConfig, Observable, StrongMeasurementConfig, TimeEvolutionConfig
from .networks import MPO, MPS, init_general_hamiltonian, init_ising
from .noise import LocalNoiseProcess, NoiseModel, relaxation_dephasing
from .simulator import Context, YaqsError, measure_shots, run, run_shots, single_shot_measure

__all__ = [
    "gates", "DigitalCircuit", "TJMOptions", "haar_brickwork", "MeasurementConfig", "Observable",
    "StrongMeasurementConfig", "measure_shots", "run_shots", "single_shot_measure",
    "TimeEvolutionConfig",
    "MPO", "MPS", "init_general_hamiltonian", "init_ising", "LocalNoiseProcess", "NoiseModel",
    "relaxation_dephasing", "Context", "YaqsError", "run",
]
