"""Digital circuits (src/DigitalTJM.jl:35-66) and the Haar brickwork generator."""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .gates import AbstractOperator, Barrier, MatrixGate, haar_su4


@dataclass
class DigitalGate:
    op: AbstractOperator
    sites: List[int]


class DigitalCircuit:
    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self.gates: List[DigitalGate] = []

    def add_gate(self, op: AbstractOperator, sites) -> None:
        """add_gate! (src/DigitalTJM.jl:64-66)."""
        sites = [int(sites)] if np.ndim(sites) == 0 else [int(s) for s in sites]
        self.gates.append(DigitalGate(op, sites))

    def add_sample_barrier(self) -> None:
        self.gates.append(DigitalGate(Barrier("SAMPLE_OBSERVABLES"), []))


def haar_brickwork(L: int, depth: int, seed: int, sample_every_layer: bool = False,
                   sample_at_start: bool = False) -> DigitalCircuit:
    """Brickwork of Haar-random SU(4) gates: odd layers act on bonds
    (1,2),(3,4),..., even layers on (2,3),(4,5),...; the gates are drawn in
    order from numpy's PCG64 seeded with `seed`.  A SAMPLE_OBSERVABLES barrier
    closes the circuit (and every layer if sample_every_layer)."""
    rng = np.random.default_rng(seed)
    c = DigitalCircuit(L)
    if sample_at_start:
        c.add_sample_barrier()
    for layer in range(depth):
        start = 1 if layer % 2 == 0 else 2
        for s in range(start, L, 2):
            c.add_gate(MatrixGate(haar_su4(rng)), [s, s + 1])
        if sample_every_layer or layer == depth - 1:
            c.add_sample_barrier()
    return c
