"""NoiseModel for the standard (local) unraveling (src/Noise.jl:15-37, 299-420)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

from .gates import LoweringGate, RaisingGate, XGate, YGate, ZGate

C = np.complex128

_PAULI = {"x": XGate, "y": YGate, "z": ZGate}
# name-based Pauli test of src/Dissipation.jl:14-21
_PAULI_NAMES = {"pauli_x", "pauli_y", "pauli_z", "crosstalk_xx", "crosstalk_yy", "crosstalk_zz",
                "crosstalk_xy", "crosstalk_yx", "crosstalk_zy", "crosstalk_zx", "crosstalk_yz",
                "crosstalk_xz"}


def get_operator(name: str) -> np.ndarray:
    """src/Noise.jl:41-61."""
    if name == "pauli_x":
        return XGate().matrix
    if name == "pauli_y":
        return YGate().matrix
    if name == "pauli_z":
        return ZGate().matrix
    if name == "raising":
        return RaisingGate().matrix
    if name == "lowering":
        return LoweringGate().matrix
    if name == "raising_two":
        return np.kron(RaisingGate().matrix, RaisingGate().matrix)
    if name == "lowering_two":
        return np.kron(LoweringGate().matrix, LoweringGate().matrix)
    if name.startswith("crosstalk_") and len(name) == 12:
        a, b = name[10], name[11]
        return np.kron(_PAULI[a]().matrix, _PAULI[b]().matrix)
    raise ValueError(f"Unknown operator: {name}")


@dataclass
class LocalNoiseProcess:
    name: str
    sites: List[int]
    strength: float
    matrix: np.ndarray

    @property
    def pauli_name(self) -> bool:
        return self.name in _PAULI_NAMES


class NoiseModel:
    """NoiseModel(processes, num_qubits) with the "standard" unraveling.  The
    projector / unitary unravelings and long-range (MPO) processes of
    src/Noise.jl:134-295 are not part of the batched TJM path and raise."""

    def __init__(self, processes: Sequence[Any] = (), num_qubits: Optional[int] = None):
        procs: List[LocalNoiseProcess] = []
        for p in processes:
            if isinstance(p, LocalNoiseProcess):
                procs.append(p)
                continue
            p = dict(p)
            name = p["name"]
            sites = list(p["sites"])
            strength = float(p["strength"])
            unravel = p.get("unraveling", "standard")
            if unravel not in ("standard",):
                raise NotImplementedError(f"unraveling {unravel!r} is not supported by the GPU path")
            if len(sites) == 2:
                sites = sorted(sites)
            if not sites or len(sites) > 2:
                raise ValueError("Invalid sites")
            if len(sites) == 2 and abs(sites[1] - sites[0]) > 1:
                raise NotImplementedError("long-range (MPO) noise processes are not supported by the GPU path")
            mat = np.asarray(p["matrix"], dtype=C) if "matrix" in p else get_operator(name)
            procs.append(LocalNoiseProcess(name, sites, strength, mat))
        self.processes = procs


def relaxation_dephasing(L: int, gamma: float) -> NoiseModel:
    """Lowering (|0><1|) and pauli_z on every site, strength gamma each."""
    ps: List[Dict[str, Any]] = []
    for i in range(1, L + 1):
        ps.append({"name": "lowering", "sites": [i], "strength": gamma})
        ps.append({"name": "pauli_z", "sites": [i], "strength": gamma})
    return NoiseModel(ps, L)
