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
    generator: Optional[list] = None  # overrides op.generator() (src/DigitalTJM.jl:40-44)


@dataclass
class TJMOptions:
    """TJMOptions (src/DigitalTJM.jl:70-76): how 2-qubit gates are applied.
    "TDVP" (the reference default) evolves the gate's window under its
    generator (for a long-range gate: the circuit sweep over the window
    s1..s2); "TEBD" applies the gate matrix (for a long-range gate: the SWAP
    network, src/DigitalTJM.jl:314-330)."""
    local_method: str = "TDVP"
    long_range_method: str = "TDVP"

    def __post_init__(self):
        for m in (self.local_method, self.long_range_method):
            if str(m).lstrip(":").upper() not in ("TDVP", "TEBD"):
                raise ValueError(f"unknown method {m!r} (TDVP or TEBD)")
        self.local_method = str(self.local_method).lstrip(":").upper()
        self.long_range_method = str(self.long_range_method).lstrip(":").upper()


def window_hamiltonian(gate: DigitalGate) -> np.ndarray:
    """h = coeff * g_a (x) g_b of construct_window_mpo (src/DigitalTJM.jl:192-236)
    for a 2-qubit gate, g_a on the lower site: generator element 1 belongs to
    gate.sites[0], so the factors swap when the sites are given in
    descending order.  Indexed like gate matrices (p_a + 2 p_b)."""
    gen = gate.generator if gate.generator is not None else gate.op.generator()
    if gen is None or len(gen) != 2:
        raise ValueError("a 2-qubit gate needs a two-factor generator for the TDVP window")
    ga, gb = (np.asarray(gen[0], dtype=complex), np.asarray(gen[1], dtype=complex))
    if gate.sites[0] != min(gate.sites):
        ga, gb = gb, ga
    return gate.op.hamiltonian_coeff() * np.kron(gb, ga)


class DigitalCircuit:
    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self.gates: List[DigitalGate] = []
        # processed layers (src/DigitalTJM.jl:41-45): filled by circuit
        # ingestion for inspection; runs layer `gates` themselves
        self.layers: List[List[DigitalGate]] = []

    def add_gate(self, op: AbstractOperator, sites, generator=None) -> None:
        """add_gate! (src/DigitalTJM.jl:64-66)."""
        sites = [int(sites)] if np.ndim(sites) == 0 else [int(s) for s in sites]
        self.gates.append(DigitalGate(op, sites, generator))

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
