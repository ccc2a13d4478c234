"""MPS and MPO types and builders (host side of the boundary).

Mirrors src/MPS.jl:33-144 (MPS, product-state constructors, normalize!) and
src/MPO.jl:32-260 (MPO, init_ising, init_general_hamiltonian).  Tensors are
numpy arrays with the reference's index order: MPS (Dl, d, Dr), MPO
(Dl, d_out, d_in, Dr); they are handed to the device in Julia's
column-major memory order.
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np

from .gates import XGate, YGate, ZGate

C = np.complex128


class MPS:
    """Matrix product state (src/MPS.jl:33-44)."""

    def __init__(self, length: int, tensors: Optional[Sequence[np.ndarray]] = None,
                 physical_dimensions: Union[int, Sequence[int], None] = None,
                 state: str = "zeros", pad: Optional[int] = None,
                 basis_string: Optional[str] = None):
        if physical_dimensions is None:
            pd = [2] * length
        elif isinstance(physical_dimensions, int):
            pd = [physical_dimensions] * length
        else:
            pd = list(physical_dimensions)
        self.length = int(length)
        self.phys_dims = pd
        self.orth_center = 1
        if tensors is not None:
            assert len(tensors) == length
            self.tensors = [np.array(t, dtype=C) for t in tensors]
            return
        ts = []
        if state == "basis":  # src/MPS.jl:78-89
            assert basis_string is not None and len(basis_string) == length
            for i in range(length):
                v = np.zeros((1, pd[i], 1), dtype=C)
                v[0, int(basis_string[i]), 0] = 1.0
                ts.append(v)
        elif state == "random":  # src/MPS.jl:91-101
            rng = np.random.default_rng()
            for i in range(length):
                cl = 1 if i == 0 else 4
                cr = 1 if i == length - 1 else 4
                ts.append(rng.standard_normal((cl, pd[i], cr)) + 1j * rng.standard_normal((cl, pd[i], cr)))
            self.tensors = ts
            self.normalize()
            return
        else:  # product states, src/MPS.jl:103-130
            s2 = 1 / np.sqrt(2)
            for i in range(length):
                v = np.zeros(pd[i], dtype=C)
                if state == "zeros":
                    v[0] = 1
                elif state == "ones":
                    v[1 if pd[i] > 1 else 0] = 1
                elif state == "x+":
                    v[0], v[1] = s2, s2
                elif state == "x-":
                    v[0], v[1] = s2, -s2
                elif state == "y+":
                    v[0], v[1] = s2, 1j * s2
                elif state == "y-":
                    v[0], v[1] = s2, -1j * s2
                elif state == "Neel":  # odd sites (1-based) |0>, even |1>
                    v[0 if (i + 1) % 2 != 0 else 1] = 1
                elif state == "wall":
                    v[0 if (i + 1) <= length // 2 else 1] = 1
                else:
                    raise ValueError(f"Unknown state: {state}")
                ts.append(v.reshape(1, pd[i], 1))
        self.tensors = ts
        if pad is not None:  # src/MPS.jl:139-141
            self.pad_bond_dimension(pad)

    def pad_bond_dimension(self, target_dim: int, noise_scale: float = 1e-8, rng=None) -> None:
        """pad_bond_dimension! (src/MPS.jl:453-485): normalize, grow every
        internal bond by appending columns (site i) and rows (site i+1)
        filled with complex normal noise of amplitude `noise_scale`, then
        normalize again.  The state changes by O(noise_scale) and every
        Schmidt value of the padded bonds is non-zero, which the 1-site TDVP
        sweep needs to grow entanglement.

        Bond b is padded to min(target_dim, 2^min(b, L-b)) for qubits (the
        largest rank the bond can carry): the reference pads every bond to
        target_dim, and its final normalize! already trims the right half
        back to that rank, while the extra left-half dimensions carry only
        noise of rank <= 2^b; the engine sizes its buffers by that rank.
        The noise comes from numpy's generator, not Julia's default_rng.
        The padded Schmidt values are products of the seeds (noise_scale^2
        and below: near 1e-16 with the default).  The engine's analog runs
        never move the centre through the padded directions before a sweep
        evolves them, so the 1-site sweep grows entanglement from them as in
        the reference (tests/test_gpu_tdvp1.py: the exact evolution of a
        padded Néel state to 1e-12).  The oracle restates the reference's
        measurement and dissipation centre moves, whose gauge steps drop
        directions below 1e-12 of their block as exactly rank-deficient ones
        (LAPACK in the reference keeps them), so per-trajectory comparisons
        of 1-site runs use states whose Schmidt values all lie above that."""
        assert target_dim > 0
        rng = rng or np.random.default_rng()
        self.normalize()
        L = self.length
        for i in range(L - 1):
            A, B = self.tensors[i], self.tensors[i + 1]
            cap = min(int(np.prod(self.phys_dims[:i + 1])), int(np.prod(self.phys_dims[i + 1:])))
            chi = A.shape[2]
            new = min(target_dim, cap)
            if chi >= new:
                continue
            nA = np.zeros((A.shape[0], A.shape[1], new), dtype=C)
            nA[:, :, :chi] = A
            nA[:, :, chi:] = noise_scale * (rng.standard_normal(nA[:, :, chi:].shape)
                                            + 1j * rng.standard_normal(nA[:, :, chi:].shape))
            nB = np.zeros((new, B.shape[1], B.shape[2]), dtype=C)
            nB[:chi] = B
            nB[chi:] = noise_scale * (rng.standard_normal(nB[chi:].shape) + 1j * rng.standard_normal(nB[chi:].shape))
            self.tensors[i], self.tensors[i + 1] = nA, nB
        self.normalize()

    def copy(self) -> "MPS":
        m = MPS(self.length, tensors=[t.copy() for t in self.tensors], physical_dimensions=self.phys_dims)
        m.orth_center = self.orth_center
        return m

    def normalize(self) -> None:
        """normalize! (src/MPS.jl:274-311): LQ sweep L->2, then scale site 1."""
        for i in range(self.length - 1, 0, -1):
            A = self.tensors[i]
            Dl, d, Dr = A.shape
            mat = A.reshape(Dl, d * Dr, order="F")
            q, r = np.linalg.qr(mat.conj().T)  # mat = r^H q^H
            Lm, Q = r.conj().T, q.conj().T
            rn = Lm.shape[1]
            self.tensors[i] = Q.reshape(rn, d, Dr, order="F")
            P = self.tensors[i - 1]
            lp, dp, _ = P.shape
            Pm = P.reshape(lp * dp, Dl, order="F") @ Lm
            self.tensors[i - 1] = Pm.reshape(lp, dp, rn, order="F")
        nrm = np.linalg.norm(self.tensors[0])
        if nrm > 1e-13:
            self.tensors[0] = self.tensors[0] / nrm
        self.orth_center = 1

    def max_bond_dim(self) -> int:
        return max(max(t.shape[0], t.shape[2]) for t in self.tensors)

    def to_vec(self) -> np.ndarray:
        """Dense state vector, site 1 most significant (for tests at small L)."""
        v = self.tensors[0][0]  # (d, Dr) since Dl = 1
        for t in self.tensors[1:]:
            v = np.einsum("xa,asb->xsb", v, t).reshape(-1, t.shape[2])
        return v[:, 0]


class MPO:
    """Matrix product operator (src/MPO.jl:32-46); tensors (Dl, d_out, d_in, Dr)."""

    def __init__(self, length: int, tensors: Sequence[np.ndarray], phys_dims=None):
        self.length = int(length)
        self.tensors = [np.array(t, dtype=C) for t in tensors]
        self.phys_dims = phys_dims or [2] * length

    @property
    def bond_dim(self) -> int:
        return max(max(t.shape[0], t.shape[3]) for t in self.tensors)


def _block_to_tensor(W):
    rows, cols = len(W), len(W[0])
    T = np.zeros((rows, 2, 2, cols), dtype=C)
    for r in range(rows):
        for c in range(cols):
            T[r, :, :, c] = W[r][c]
    return T


def init_ising(length: int, J, g) -> MPO:
    """H = -J Σ Z_i Z_{i+1} - g Σ X_i, bond dimension 3 (src/MPO.jl:92-163)."""
    X, Z, I, O = XGate().matrix, ZGate().matrix, np.eye(2, dtype=C), np.zeros((2, 2), dtype=C)

    def gJ(i):  # 1-based
        if np.ndim(J):
            return J[i - 1] if i <= len(J) else 0.0
        return J

    def gg(i):
        return g[i - 1] if np.ndim(g) else g

    if length == 1:
        return MPO(1, [(-gg(1) * X).reshape(1, 2, 2, 1)])
    ts = [_block_to_tensor([[I, -gJ(1) * Z, -gg(1) * X]])]
    for i in range(2, length):
        W = [[O] * 3 for _ in range(3)]
        W[0][0], W[0][1], W[0][2] = I, -gJ(i) * Z, -gg(i) * X
        W[1][2], W[2][2] = Z, I
        ts.append(_block_to_tensor(W))
    ts.append(_block_to_tensor([[-gg(length) * X], [Z], [I]]))
    return MPO(length, ts)


def init_general_hamiltonian(length: int, Jxx, Jyy, Jzz, hx, hy, hz) -> MPO:
    """H = Σ (Jxx XX + Jyy YY + Jzz ZZ) + Σ (hx X + hy Y + hz Z), bond dimension 5
    (src/MPO.jl:175-260)."""
    X, Y, Z = XGate().matrix, YGate().matrix, ZGate().matrix
    I, O = np.eye(2, dtype=C), np.zeros((2, 2), dtype=C)

    def gp(p, i):
        if np.ndim(p):
            return p[i - 1] if i <= len(p) else 0.0
        return p

    ts = []
    for i in range(1, length + 1):
        jx = gp(Jxx, i) if i < length else 0.0
        jy = gp(Jyy, i) if i < length else 0.0
        jz = gp(Jzz, i) if i < length else 0.0
        field = gp(hx, i) * X + gp(hy, i) * Y + gp(hz, i) * Z
        W = [[O] * 5 for _ in range(5)]
        W[0] = [I, jx * X, jy * Y, jz * Z, field]
        W[1][4], W[2][4], W[3][4], W[4][4] = X, Y, Z, I
        if i == 1:
            if length == 1:
                ts.append(field.reshape(1, 2, 2, 1))
            else:
                ts.append(_block_to_tensor([W[0]]))
        elif i == length:
            ts.append(_block_to_tensor([[W[r][4]] for r in range(5)]))
        else:
            ts.append(_block_to_tensor(W))
    return MPO(length, ts)
