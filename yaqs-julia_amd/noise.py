"""NoiseModel with the standard, projector, two-point unitary and Gaussian
unitary unravelings of local and long-range processes (src/Noise.jl:15-37,
77-295, 299-433)."""
from __future__ import annotations

import warnings
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


def get_pauli(c: str) -> np.ndarray:
    """src/Noise.jl:63-68."""
    if c not in _PAULI:
        raise ValueError(f"Unknown Pauli: {c}")
    return _PAULI[c]().matrix


def _is_pauli_matrix(op: np.ndarray) -> bool:
    """is_pauli (src/StochasticProcess.jl:337-342): P^+ P = I."""
    op = np.asarray(op, dtype=C)
    return bool(np.allclose(op.conj().T @ op, np.eye(op.shape[0]), atol=1e-10, rtol=0))


def build_mpo_phys(L: int, i: int, j: int, sigma, tau, a, b):
    """build_mpo_phys (src/Noise.jl:77-128): the MPO of O = a I + b sigma_i
    tau_j (1-based i < j), bond dimension 2 between i and j: site i the row
    [I, sigma], sites between diag(I, I), site j the column [a I; b tau],
    identities elsewhere; tensors (Left, Out, In, Right)."""
    from .networks import MPO
    if not 1 <= i < j <= L:
        raise ValueError("Require 1 <= i < j <= L")
    eye = np.eye(2, dtype=C)
    ts = []
    for k in range(1, L + 1):
        if k < i or k > j:
            T = np.zeros((1, 2, 2, 1), dtype=C)
            T[0, :, :, 0] = eye
        elif k == i:
            T = np.zeros((1, 2, 2, 2), dtype=C)
            T[0, :, :, 0] = eye
            T[0, :, :, 1] = sigma
        elif k < j:
            T = np.zeros((2, 2, 2, 2), dtype=C)
            T[0, :, :, 0] = eye
            T[1, :, :, 1] = eye
        else:
            T = np.zeros((2, 2, 2, 1), dtype=C)
            T[0, :, :, 0] = a * eye
            T[1, :, :, 0] = b * np.asarray(tau, dtype=C)
        ts.append(T)
    return MPO(L, ts)


@dataclass
class MPONoiseProcess:
    """MPONoiseProcess (src/Noise.jl:22-33): a process on two non-adjacent
    sites whose jump operator is the MPO of O = a I + b sigma_i tau_j
    (build_mpo_phys).  `factors` ([sigma, tau]) is set for the standard
    (pass-through) long-range process only, as the reference does."""
    name: str
    sites: List[int]
    strength: float
    mpo: Any
    factors: List[np.ndarray]
    sigma: np.ndarray
    tau: np.ndarray
    a: complex
    b: complex

    @property
    def pauli_name(self) -> bool:
        return self.name in _PAULI_NAMES

    def is_pauli_proc(self) -> bool:
        """is_pauli_proc (src/StochasticProcess.jl:344-358): by name (pauli /
        crosstalk, not projector), else all factors unitary."""
        n = self.name
        if ("pauli" in n or "crosstalk" in n) and "projector" not in n:
            return True
        return bool(self.factors) and all(_is_pauli_matrix(f) for f in self.factors)

    def pauli_long_range(self) -> bool:
        """The candidate create_probability_distribution tags
        "pauli_long_range" (src/StochasticProcess.jl:168-181): a Pauli
        process that carries factors."""
        return self.is_pauli_proc() and bool(self.factors)

    def check_dissipation_name(self) -> None:
        """handle_longrange_scalar! (src/Dissipation.jl:161-204) accepts
        these names only."""
        if not self.name.startswith(("projector_", "unitary2pt_", "unitary_gauss_", "crosstalk_", "pauli_")):
            raise ValueError(f"Non-Pauli long-range processes not fully implemented for: {self.name}")


def _factors_of(p: Dict[str, Any]):
    """The (sigma, tau) of a long-range process: "factors", else the last
    two letters of the name (src/Noise.jl:159-167)."""
    if "factors" in p:
        sig, tau = p["factors"]
        return np.asarray(sig, dtype=C), np.asarray(tau, dtype=C)
    suffix = p["name"].split("_")[-1]
    if len(suffix) < 2:
        raise ValueError(f"cannot parse the Pauli factors of {p['name']!r}")
    return get_pauli(suffix[0]), get_pauli(suffix[1])


def _add_longrange(procs, name, sites, strength, L, sig, tau, a, b, factors=()):
    i, j = sites
    procs.append(MPONoiseProcess(name, [i, j], float(strength), build_mpo_phys(L, i, j, sig, tau, a, b),
                                 [np.asarray(f, dtype=C) for f in factors], sig, tau, complex(a), complex(b)))


def add_projector_expansion_longrange(procs, p, sites, L, gamma) -> None:
    """src/Noise.jl:156-178: the MPOs of (I +- sigma_i tau_j) at gamma/2."""
    sig, tau = _factors_of(p)
    for comp, sgn in (("plus", 1.0), ("minus", -1.0)):
        _add_longrange(procs, f"projector_{comp}_{p['name']}", sites, gamma / 2.0, L, sig, tau, 1.0, sgn)


def add_unitary_2pt_expansion_longrange(procs, p, sites, L, gamma, theta0) -> None:
    """src/Noise.jl:195-220: cos(theta0) I +- i sin(theta0) sigma_i tau_j at
    lambda/2, lambda = gamma / sin^2(theta0)."""
    sig, tau = _factors_of(p)
    lam = gamma / np.sin(theta0) ** 2
    for comp, sgn in (("plus", 1.0), ("minus", -1.0)):
        _add_longrange(procs, f"unitary2pt_{comp}_{p['name']}", sites, lam / 2.0, L, sig, tau,
                       np.cos(theta0), 1j * sgn * np.sin(theta0))


def add_unitary_gauss_expansion_longrange(procs, p, sites, L, gamma, sigma, gauss_M, gauss_k) -> None:
    """src/Noise.jl:254-295: cos(theta_k) I + i sin(theta_k) sigma_i tau_j
    at lambda w_k on the Gaussian angle grid."""
    sig, tau = _factors_of(p)
    M = int(p.get("M", gauss_M))
    theta_max = float(p.get("theta_max", gauss_k * sigma))
    thetas, w = gauss_angles(sigma, M, theta_max)
    lam = gamma / float(np.sum(w * np.sin(thetas) ** 2))
    for idx, (wk, th) in enumerate(zip(w, thetas)):
        if wk <= 0.0:
            continue
        _add_longrange(procs, f"unitary_gauss_{idx}_{p['name']}", sites, lam * float(wk), L, sig, tau,
                       np.cos(th), 1j * np.sin(th))


def _expm_i(theta: float, P: np.ndarray) -> np.ndarray:
    """exp(i theta P) for a 2x2 or 4x4 operator (the matrix exponential of
    src/Noise.jl:183,263)."""
    from scipy.linalg import expm
    return expm(1j * theta * np.asarray(P, dtype=C))


def _operator_of(p: Dict[str, Any]) -> np.ndarray:
    """src/Noise.jl:423-433: an explicit "matrix" wins over the name."""
    return np.asarray(p["matrix"], dtype=C) if "matrix" in p else get_operator(p["name"])


def add_projector_expansion(procs: List[LocalNoiseProcess], p: Dict[str, Any], sites: List[int],
                            P: np.ndarray, gamma: float) -> None:
    """src/Noise.jl:134-154: L_pm = sqrt(gamma/2) (I +- P)."""
    eye = np.eye(P.shape[0], dtype=C)
    for comp, sgn in (("plus", 1.0), ("minus", -1.0)):
        procs.append(LocalNoiseProcess(f"projector_{comp}_{p['name']}", list(sites), gamma / 2.0, eye + sgn * P))


def add_unitary_2pt_expansion(procs: List[LocalNoiseProcess], p: Dict[str, Any], sites: List[int],
                              P: np.ndarray, gamma: float, theta0: float) -> None:
    """src/Noise.jl:180-193: U_pm = exp(+-i theta0 P) at rate
    lambda/2 each, lambda = gamma / sin(theta0)^2."""
    s_val = np.sin(theta0) ** 2
    if not s_val > 0:
        raise ValueError("theta0 too small")
    lam = gamma / s_val
    for comp, sgn in (("plus", 1.0), ("minus", -1.0)):
        procs.append(LocalNoiseProcess(f"unitary2pt_{comp}_{p['name']}", list(sites), lam / 2.0,
                                       _expm_i(sgn * theta0, P)))


def gauss_angles(sigma: float, M: int, theta_max: float):
    """Angles and weights of the Gaussian unraveling (src/Noise.jl:226-237):
    (M+1)//2 equidistant angles on [0, theta_max] mirrored to the negative
    side, Gaussian weights normalised to 1 and symmetrised."""
    pos = np.linspace(0.0, theta_max, (M + 1) // 2)
    thetas = np.concatenate([-pos[1:][::-1], pos])
    w = np.exp(-0.5 * (thetas / sigma) ** 2)
    w = w / w.sum()
    w = 0.5 * (w + w[::-1])
    return thetas, w


def add_unitary_gauss_expansion(procs: List[LocalNoiseProcess], p: Dict[str, Any], sites: List[int],
                                P: np.ndarray, gamma: float, sigma: float, gauss_M: int, gauss_k: float) -> None:
    """src/Noise.jl:222-252: U_k = exp(i theta_k P) at rate lambda w_k,
    lambda = gamma / sum_k w_k sin^2(theta_k); zero-weight angles are
    skipped but keep their index in the name."""
    M = int(p.get("M", gauss_M))
    theta_max = float(p.get("theta_max", gauss_k * sigma))
    thetas, w = gauss_angles(sigma, M, theta_max)
    s_weight = float(np.sum(w * np.sin(thetas) ** 2))
    if not s_weight > 1e-12:
        raise ValueError("E[sin^2 theta] too small")
    lam = gamma / s_weight
    for idx, (wk, th) in enumerate(zip(w, thetas)):
        if wk <= 0.0:
            continue
        procs.append(LocalNoiseProcess(f"unitary_gauss_{idx}_{p['name']}", list(sites), lam * float(wk),
                                       _expm_i(float(th), P)))


class NoiseModel:
    """NoiseModel(processes, num_qubits; theta0, dt, sigma, gauss_M, gauss_k)
    (src/Noise.jl:299-420).  Each process dict names the operator ("name",
    or an explicit "matrix"), its 1-based "sites" (one site or an adjacent
    pair), "strength" and an optional "unraveling":

    - "standard": the process itself;
    - "projector": the two processes (I +- P) at strength gamma/2;
    - "unitary_2pt": exp(+- i theta0 P) at strength gamma / (2 sin^2 theta0)
      (per-process "theta0" overrides the keyword);
    - "unitary_gauss": exp(i theta_k P) on the Gaussian angle grid
      (per-process "sigma", "M", "gauss_k", "theta_max");
    - "analog_auto": deprecated, warns and falls back to "standard"; any
      other value is taken as "standard", as in the reference.

    All of these average to the same Lindblad dissipator.  A pair of
    non-adjacent sites is a long-range process: an MPONoiseProcess whose
    jump operator is the MPO of O = a I + b sigma_i tau_j with the factors
    from "factors" or the name's last two letters (standard: a = 0, b = 1,
    factors kept; projector: a = 1, b = +-1; unitary_2pt / unitary_gauss:
    a = cos(theta), b = +-i sin(theta)).  Adjacent pairs with a non-Pauli
    operator (raising_two, lowering_two, explicit 4x4 matrices, the
    expansions of crosstalk processes) run on the analog GPU path as
    merged-pair dissipation and jumps; long-range processes run on the
    analog path; the digital path rejects both."""

    def __init__(self, processes: Sequence[Any] = (), num_qubits: Optional[int] = None, *,
                 theta0: float = np.pi / 2, dt: Optional[float] = None, sigma: float = 1.0,
                 gauss_M: int = 11, gauss_k: float = 4.0):
        procs: List[Any] = []
        for p in processes:
            if isinstance(p, (LocalNoiseProcess, MPONoiseProcess)):
                procs.append(p)
                continue
            p = dict(p)
            name = p["name"]
            sites = list(p["sites"])
            strength = float(p["strength"])
            unravel = p.get("unraveling", "standard")
            if len(sites) == 2:
                sites = sorted(sites)
            if not sites or len(sites) > 2:
                raise ValueError("Invalid sites")
            long_range = len(sites) == 2 and abs(sites[1] - sites[0]) > 1
            if long_range and num_qubits is None:
                raise ValueError("long-range processes need num_qubits")
            if unravel == "projector":
                if long_range:
                    add_projector_expansion_longrange(procs, p, sites, int(num_qubits), strength)
                else:
                    add_projector_expansion(procs, p, sites, _operator_of(p), strength)
                continue
            if unravel == "unitary_2pt":
                th = float(p.get("theta0", theta0))
                s_val = np.sin(th) ** 2
                if s_val > 1e-12 and dt is not None and strength / s_val * dt > 0.1:
                    warnings.warn(f"Stability Warning: theta0={th} is too small for dt={dt} "
                                  f"(lambda*dt = {strength / s_val * dt} > 0.1). Consider increasing theta0 "
                                  "or decreasing dt.")
                if long_range:
                    add_unitary_2pt_expansion_longrange(procs, p, sites, int(num_qubits), strength, th)
                else:
                    add_unitary_2pt_expansion(procs, p, sites, _operator_of(p), strength, th)
                continue
            if unravel == "unitary_gauss":
                args = (float(p.get("sigma", sigma)), int(p.get("M", gauss_M)), float(p.get("gauss_k", gauss_k)))
                if long_range:
                    add_unitary_gauss_expansion_longrange(procs, p, sites, int(num_qubits), strength, *args)
                else:
                    add_unitary_gauss_expansion(procs, p, sites, _operator_of(p), strength, *args)
                continue
            if unravel == "analog_auto":
                warnings.warn("Unraveling mode 'analog_auto' is deprecated. Please choose 'unitary_2pt' or "
                              "'unitary_gauss' manually.")
            if long_range and "mpo" not in p:
                # standard long-range: O = sigma_i tau_j, factors kept (src/Noise.jl:372-408)
                sig, tau = _factors_of(p)
                _add_longrange(procs, name, sites, strength, int(num_qubits), sig, tau, 0.0, 1.0, (sig, tau))
                continue
            procs.append(LocalNoiseProcess(name, sites, strength, _operator_of(p)))
        self.processes = procs


def relaxation_dephasing(L: int, gamma: float) -> NoiseModel:
    """Lowering (|0><1|) and pauli_z on every site, strength gamma each."""
    ps: List[Dict[str, Any]] = []
    for i in range(1, L + 1):
        ps.append({"name": "lowering", "sites": [i], "strength": gamma})
        ps.append({"name": "pauli_z", "sites": [i], "strength": gamma})
    return NoiseModel(ps, L)
