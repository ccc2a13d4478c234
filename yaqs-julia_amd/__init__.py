"""yaqs-julia_amd — MI355X TJM trajectory engine behind the Yaqs API.

Host-side mirror of Yaqs.jl's simulator surface (MPS/MPO, NoiseModel,
Observable, TimeEvolutionConfig, Simulator.run).  All trajectory work runs in
the HIP C-ABI library `build/libyaqs_hip.so`; this package only builds job
descriptions and reads results back.
"""
from . import gates
from .circuit import DigitalCircuit, TJMOptions, haar_brickwork
from .ingestion import convert_instruction_to_gate, ingest_qiskit_circuit, map_qiskit_name
from .configs import MeasurementConfig, Observable, StrongMeasurementConfig, TimeEvolutionConfig
from .networks import MPO, MPS, init_general_hamiltonian, init_ising
from .noise import LocalNoiseProcess, NoiseModel, relaxation_dephasing
from .simulator import Context, YaqsError, measure_shots, run, run_shots, single_shot_measure

__all__ = [
    "gates", "DigitalCircuit", "ingest_qiskit_circuit", "map_qiskit_name", "convert_instruction_to_gate", "TJMOptions", "haar_brickwork", "MeasurementConfig", "Observable",
    "StrongMeasurementConfig", "measure_shots", "run_shots", "single_shot_measure",
    "TimeEvolutionConfig",
    "MPO", "MPS", "init_general_hamiltonian", "init_ising", "LocalNoiseProcess", "NoiseModel",
    "relaxation_dephasing", "Context", "YaqsError", "run",
]
