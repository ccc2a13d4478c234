"""Circuit ingestion (src/CircuitIngestion.jl): a Qiskit QuantumCircuit to a
DigitalCircuit.

`ingest_qiskit_circuit` (src/CircuitIngestion.jl:15-164) turns the circuit
into its dependency DAG and repeatedly removes the DAG's front layer: the
operations of the front layer that are gates become one layer (1-qubit gates
first, then 2-qubit gates), `measure` and `barrier` operations are dropped,
and a gate on more than two qubits is an error.  Every gate is also appended
to `circ.gates`, in layer order, which is what a run simulates (the run
layers `circ.gates` again, src/DigitalTJM.jl:86-145).

The DAG is computed here from the circuit's instruction list instead of
`qiskit.converters.circuit_to_dag`, so the front layers do not need Qiskit:
an instruction depends on the previous instruction on each of its wires
(its qubits and its classical bits), the edges Qiskit's DAG has, and the
front layer is every remaining instruction whose predecessors have all been
removed.  Within a front layer, operations come in the order Qiskit's layer
walk discovers them: by the last of their wires in wire order (qubits, then
classical bits).  The circuit is read through the attributes the reference
reads: `qc.num_qubits`, `qc.data` (items with `.operation`, `.qubits`,
`.clbits`, or legacy `(operation, qargs, cargs)` tuples), `operation.name`,
`.params`, `.label`, a qubit's `_index` (else `qc.find_bit(q).index`).
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import gates as G
from .circuit import DigitalCircuit, DigitalGate


def map_qiskit_name(name: str, params: Sequence[float]) -> G.AbstractOperator:
    """src/CircuitIngestion.jl:270-310.  `sx` maps to Rx(pi/2) (sqrt(X) up to
    a global phase); a `cx` that carries parameters is unsupported."""
    params = [float(p) for p in params]
    if name == "x" or (name == "cx" and len(params) == 0):
        return G.CXGate() if name == "cx" else G.XGate()
    if name == "y":
        return G.YGate()
    if name in ("z", "cz"):
        return G.CZGate() if name == "cz" else G.ZGate()
    if name == "h":
        return G.HGate()
    if name == "id":
        return G.IdGate()
    if name == "s":
        return G.SGate()
    if name == "t":
        return G.TGate()
    if name == "sx":
        return G.RxGate(np.pi / 2)
    if name == "rx":
        return G.RxGate(params[0])
    if name == "ry":
        return G.RyGate(params[0])
    if name == "rz":
        return G.RzGate(params[0])
    if name in ("p", "phase"):
        return G.PhaseGate(params[0])
    if name in ("u", "u3"):
        return G.UGate(params[0], params[1], params[2])
    if name == "swap":
        return G.SWAPGate()
    if name == "rxx":
        return G.RxxGate(params[0])
    if name == "ryy":
        return G.RyyGate(params[0])
    if name == "rzz":
        return G.RzzGate(params[0])
    raise ValueError(f"Unsupported Qiskit gate: {name}")


def _parts(instr):
    """(operation, qubits, clbits) of a CircuitInstruction or a legacy tuple."""
    if isinstance(instr, tuple):
        op, qargs = instr[0], instr[1]
        cargs = instr[2] if len(instr) > 2 else ()
        return op, list(qargs), list(cargs)
    return instr.operation, list(instr.qubits), list(getattr(instr, "clbits", ()))


def _bit_index(bit, circuit) -> int:
    """0-based index of a qubit or clbit: `_index` if the bit has one
    (src/CircuitIngestion.jl:185-193), else `circuit.find_bit(bit).index`."""
    if isinstance(bit, (int, np.integer)):
        return int(bit)
    idx = getattr(bit, "_index", None)
    if idx is not None:
        return int(idx)
    return int(circuit.find_bit(bit).index)


def _generator(op: G.AbstractOperator):
    """GateLibrary.generator(op), or None where the gate defines none
    (src/CircuitIngestion.jl:227-233)."""
    try:
        return op.generator()
    except NotImplementedError:
        return None


def _params(op) -> List[float]:
    return [float(p) for p in getattr(op, "params", ())]


def convert_instruction_to_gate(instr, circuit) -> Optional[DigitalGate]:
    """src/CircuitIngestion.jl:166-234: one circuit instruction as a
    DigitalGate (1-based sites).  A barrier becomes Barrier(label), with
    SAMPLE_OBSERVABLES when it has no label; an unsupported gate gives None."""
    op, qubits, _ = _parts(instr)
    sites = [_bit_index(q, circuit) + 1 for q in qubits]
    name = str(op.name)
    if name == "barrier":
        label = getattr(op, "label", None)
        label = "" if label is None else str(label)
        return DigitalGate(G.Barrier(label or "SAMPLE_OBSERVABLES"), sites, None)
    try:
        jop = map_qiskit_name(name, _params(op))
    except (ValueError, IndexError):
        return None
    return DigitalGate(jop, sites, _generator(jop))


def _front_layers(qc):
    """Yield the DAG's front layers as lists of (operation, qubit indices),
    each in the layer walk's discovery order, removing each layer before the
    next is formed."""
    nq = int(qc.num_qubits)
    nodes = []  # (op, qubit indices, wires)
    for instr in qc.data:
        op, qubits, clbits = _parts(instr)
        qi = [_bit_index(q, qc) for q in qubits]
        wires = qi + [nq + _bit_index(c, qc) for c in clbits]
        nodes.append((op, qi, wires))
    # per wire, the queue of nodes on it (in circuit order)
    queues = {}
    for n, (_, _, wires) in enumerate(nodes):
        for w in dict.fromkeys(wires):
            queues.setdefault(w, []).append(n)
    heads = {w: 0 for w in queues}
    remaining = len(nodes)
    while remaining:
        front = []
        for w in sorted(queues):
            q, h = queues[w], heads[w]
            if h >= len(q):
                continue
            n = q[h]
            wires = dict.fromkeys(nodes[n][2])
            # a node is in the front layer when it heads every one of its
            # wires; it is found when its last wire in wire order is reached
            if w == max(wires) and all(queues[x][heads[x]] == n for x in wires):
                front.append(n)
        assert front, "dependency cycle in the circuit's instruction list"
        yield [(nodes[n][0], nodes[n][1]) for n in front]
        for n in front:
            for w in dict.fromkeys(nodes[n][2]):
                heads[w] += 1
        remaining -= len(front)


def ingest_qiskit_circuit(qc) -> DigitalCircuit:
    """src/CircuitIngestion.jl:15-164: the circuit's DAG front layers as
    `circ.layers` and, in the same order, `circ.gates`."""
    circ = DigitalCircuit(int(qc.num_qubits))
    for layer in _front_layers(qc):
        single, double = [], []
        for op, qi in layer:
            name = str(op.name)
            if name in ("measure", "barrier"):
                continue
            if len(qi) == 1:
                single.append((op, qi))
            elif len(qi) == 2:
                double.append((op, qi))
            else:
                raise ValueError("Gates with >2 qubits not supported")
        gates_here = []
        for op, qi in single + double:
            jop = map_qiskit_name(str(op.name), _params(op))
            g = DigitalGate(jop, [q + 1 for q in qi], _generator(jop))
            gates_here.append(g)
            circ.gates.append(g)
        if gates_here:
            circ.layers.append(gates_here)
    return circ
