"""Gate and operator matrices (GateLibrary).

Matrices are numpy arrays indexed [row, col] with the reference's values
(src/GateLibrary.jl:44-302).  A two-site 4x4 matrix acts on the combined
physical index p1 + 2*p2 (the first site is the least significant), which is
how the reference reshapes it to (2,2,2,2) when applying it
(src/DigitalTJM.jl:265-268, src/Dissipation.jl:86-88).
"""
from __future__ import annotations

import numpy as np

C = np.complex128


def _cm(vals, n):
    """Build an n x n matrix from column-major values (the SMatrix literal order)."""
    return np.array(vals, dtype=C).reshape(n, n, order="F")


class AbstractOperator:
    nsites = 1
    unitary = True

    @property
    def matrix(self) -> np.ndarray:
        raise NotImplementedError

    def generator(self):
        """The factors [g_1(, g_2)] with gate = exp(-i coeff g_1 (x) g_2)
        (src/GateLibrary.jl:24-26: gates without one raise)."""
        raise NotImplementedError(f"generator() not implemented for {type(self).__name__}")

    def hamiltonian_coeff(self) -> float:
        """src/GateLibrary.jl:28-36: 1.0 unless the gate defines it."""
        return 1.0


class _Fixed(AbstractOperator):
    _vals: tuple = ()

    @property
    def matrix(self):
        n = 2 if len(self._vals) == 4 else 4
        return _cm(self._vals, n)


class IdGate(_Fixed):
    _vals = (1, 0, 0, 1)

    def generator(self):
        return [IdGate().matrix]


class XGate(_Fixed):  # src/GateLibrary.jl:48-50
    _vals = (0, 1, 1, 0)

    def generator(self):
        return [XGate().matrix]


class YGate(_Fixed):  # src/GateLibrary.jl:52-54
    _vals = (0, 1j, -1j, 0)

    def generator(self):
        return [YGate().matrix]


class ZGate(_Fixed):  # src/GateLibrary.jl:56-58
    _vals = (1, 0, 0, -1)

    def generator(self):
        return [ZGate().matrix]


class HGate(_Fixed):
    _vals = (1 / np.sqrt(2), 1 / np.sqrt(2), 1 / np.sqrt(2), -1 / np.sqrt(2))


class SGate(_Fixed):
    _vals = (1, 0, 0, 1j)


class SdgGate(_Fixed):
    _vals = (1, 0, 0, -1j)


class TGate(_Fixed):
    _vals = (1, 0, 0, np.exp(1j * np.pi / 4))


class TdgGate(_Fixed):
    _vals = (1, 0, 0, np.exp(-1j * np.pi / 4))


class RaisingGate(_Fixed):  # |1><0|, src/GateLibrary.jl:77-78
    _vals = (0, 1, 0, 0)
    unitary = False


class LoweringGate(_Fixed):  # |0><1|, src/GateLibrary.jl:80-81
    _vals = (0, 0, 1, 0)
    unitary = False


class _Param(AbstractOperator):
    def __init__(self, theta: float):
        self.theta = float(theta)


class RxGate(_Param):
    @property
    def matrix(self):
        c, s = np.cos(self.theta / 2), -1j * np.sin(self.theta / 2)
        return _cm((c, s, s, c), 2)

    def generator(self):  # src/GateLibrary.jl:90-120
        return [XGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class RyGate(_Param):
    @property
    def matrix(self):
        c, s = np.cos(self.theta / 2), np.sin(self.theta / 2)
        return _cm((c, s, -s, c), 2)

    def generator(self):  # src/GateLibrary.jl:90-120
        return [YGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class RzGate(_Param):
    @property
    def matrix(self):
        return _cm((np.exp(-1j * self.theta / 2), 0, 0, np.exp(1j * self.theta / 2)), 2)

    def generator(self):  # src/GateLibrary.jl:90-120
        return [ZGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class PhaseGate(_Param):
    @property
    def matrix(self):
        return _cm((1, 0, 0, np.exp(1j * self.theta)), 2)


class UGate(AbstractOperator):
    def __init__(self, theta, phi, lam):
        self.theta, self.phi, self.lam = float(theta), float(phi), float(lam)

    @property
    def matrix(self):
        c, s = np.cos(self.theta / 2), np.sin(self.theta / 2)
        return _cm((c, np.exp(1j * self.phi) * s, -np.exp(1j * self.lam) * s,
                    np.exp(1j * (self.phi + self.lam)) * c), 2)


class _Two(_Fixed):
    nsites = 2


class CXGate(_Two):  # src/GateLibrary.jl:144-150
    _vals = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0)

    def generator(self):  # pi/4 (I - Z) (x) (I - X), src/GateLibrary.jl:151-158
        return [_cm((0, 0, 0, 2), 2), _cm((1, -1, -1, 1), 2)]

    def hamiltonian_coeff(self):
        return np.pi / 4


class CYGate(_Two):
    _vals = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1j, 0, 0, -1j, 0)


class CZGate(_Two):
    _vals = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1)

    def generator(self):  # pi/4 (I - Z) (x) (I - Z), src/GateLibrary.jl:176-188
        return [_cm((0, 0, 0, 2), 2), _cm((0, 0, 0, 2), 2)]

    def hamiltonian_coeff(self):
        return np.pi / 4


class SWAPGate(_Two):
    _vals = (1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1)


class iSWAPGate(_Two):
    _vals = (1, 0, 0, 0, 0, 0, 1j, 0, 0, 1j, 0, 0, 0, 0, 0, 1)


class _Two_Param(_Param):
    nsites = 2


class CPhaseGate(_Two_Param):
    @property
    def matrix(self):
        return np.diag([1, 1, 1, np.exp(1j * self.theta)]).astype(C)

    def generator(self):  # -theta/4 (I - Z) (x) (I - Z), src/GateLibrary.jl:229-233
        return [_cm((0, 0, 0, 2), 2), _cm((0, 0, 0, 2), 2)]

    def hamiltonian_coeff(self):
        return -self.theta / 4


class RxxGate(_Two_Param):  # src/GateLibrary.jl:260-272
    @property
    def matrix(self):
        c, s = np.cos(self.theta / 2), -1j * np.sin(self.theta / 2)
        return _cm((c, 0, 0, s, 0, c, s, 0, 0, s, c, 0, s, 0, 0, c), 4)

    def generator(self):  # src/GateLibrary.jl:267-301
        return [XGate().matrix, XGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class RyyGate(_Two_Param):
    @property
    def matrix(self):
        c, s = np.cos(self.theta / 2), -1j * np.sin(self.theta / 2)
        return _cm((c, 0, 0, -s, 0, c, s, 0, 0, s, c, 0, -s, 0, 0, c), 4)

    def generator(self):  # src/GateLibrary.jl:267-301
        return [YGate().matrix, YGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class RzzGate(_Two_Param):  # exp(-i θ/2 Z⊗Z), src/GateLibrary.jl:286-300
    @property
    def matrix(self):
        a, b = np.exp(-1j * self.theta / 2), np.exp(1j * self.theta / 2)
        return np.diag([a, b, b, a]).astype(C)

    def generator(self):  # src/GateLibrary.jl:267-301
        return [ZGate().matrix, ZGate().matrix]

    def hamiltonian_coeff(self):
        return self.theta / 2


class MatrixGate(AbstractOperator):
    """An arbitrary 1- or 2-site unitary given as a matrix (used for Haar-random
    brickwork circuits; the reference's GateLibrary has no such type)."""

    def __init__(self, m):
        m = np.asarray(m, dtype=C)
        assert m.shape in ((2, 2), (4, 4))
        self._m = m
        self.nsites = 1 if m.shape == (2, 2) else 2

    @property
    def matrix(self):
        return self._m


class Barrier(AbstractOperator):
    """Barrier(label); label SAMPLE_OBSERVABLES requests a measurement
    (src/DigitalTJM.jl:147-170)."""

    def __init__(self, label: str = ""):
        self.label = label

    @property
    def matrix(self):
        raise ValueError("Barrier has no matrix")


def matrix(op: AbstractOperator) -> np.ndarray:
    return op.matrix


def haar_su4(rng: np.random.Generator) -> np.ndarray:
    """Haar-random SU(4): QR of a complex Ginibre matrix with the phases of
    R's diagonal moved into Q (Mezzadri 2007), then det normalised to 1."""
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    q = q / np.linalg.det(q) ** 0.25
    return q.astype(C)
