"""NoiseModel with the standard, projector, two-point unitary and Gaussian
unitary unravelings of local processes (src/Noise.jl:15-37, 134-295,
299-433)."""
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

    All of these average to the same Lindblad dissipator.  Long-range
    (non-adjacent, MPO) processes are not part of the batched TJM path and
    raise NotImplementedError; adjacent pairs with a non-Pauli operator
    (raising_two, lowering_two, explicit 4x4 matrices, the expansions of
    crosstalk processes) run on the analog GPU path as merged-pair
    dissipation and jumps, and are rejected by the digital path."""

    def __init__(self, processes: Sequence[Any] = (), num_qubits: Optional[int] = None, *,
                 theta0: float = np.pi / 2, dt: Optional[float] = None, sigma: float = 1.0,
                 gauss_M: int = 11, gauss_k: float = 4.0):
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
            if len(sites) == 2:
                sites = sorted(sites)
            if not sites or len(sites) > 2:
                raise ValueError("Invalid sites")
            if len(sites) == 2 and abs(sites[1] - sites[0]) > 1:
                raise NotImplementedError("long-range (MPO) noise processes are not supported by the GPU path")
            if unravel == "projector":
                add_projector_expansion(procs, p, sites, _operator_of(p), strength)
                continue
            if unravel == "unitary_2pt":
                th = float(p.get("theta0", theta0))
                s_val = np.sin(th) ** 2
                if s_val > 1e-12 and dt is not None and strength / s_val * dt > 0.1:
                    warnings.warn(f"Stability Warning: theta0={th} is too small for dt={dt} "
                                  f"(lambda*dt = {strength / s_val * dt} > 0.1). Consider increasing theta0 "
                                  "or decreasing dt.")
                add_unitary_2pt_expansion(procs, p, sites, _operator_of(p), strength, th)
                continue
            if unravel == "unitary_gauss":
                add_unitary_gauss_expansion(procs, p, sites, _operator_of(p), strength,
                                            float(p.get("sigma", sigma)), int(p.get("M", gauss_M)),
                                            float(p.get("gauss_k", gauss_k)))
                continue
            if unravel == "analog_auto":
                warnings.warn("Unraveling mode 'analog_auto' is deprecated. Please choose 'unitary_2pt' or "
                              "'unitary_gauss' manually.")
            procs.append(LocalNoiseProcess(name, sites, strength, _operator_of(p)))
        self.processes = procs


def relaxation_dephasing(L: int, gamma: float) -> NoiseModel:
    """Lowering (|0><1|) and pauli_z on every site, strength gamma each."""
    ps: List[Dict[str, Any]] = []
    for i in range(1, L + 1):
        ps.append({"name": "lowering", "sites": [i], "strength": gamma})
        ps.append({"name": "pauli_z", "sites": [i], "strength": gamma})
    return NoiseModel(ps, L)
