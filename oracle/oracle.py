"""ctypes wrapper of the FP64 CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product path.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
c_dp = ct.POINTER(ct.c_double)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        L.oracle_rng_uniform.restype = ct.c_double
        L.oracle_rng_uniform.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_uint64]
        L.oracle_run_analog.argtypes = [ct.c_void_p, c_dp, c_dp, ct.c_int, c_dp]
        L.oracle_run_digital.argtypes = [ct.c_void_p, c_dp, c_dp, ct.POINTER(ct.c_int64), ct.c_int, c_dp]
        L.oracle_analog_num_samples.argtypes = [ct.c_void_p]
        L.oracle_analog_num_samples.restype = ct.c_int64
        L.oracle_digital_num_samples.argtypes = [ct.c_void_p]
        L.oracle_digital_num_samples.restype = ct.c_int64
        L.oracle_svd.argtypes = [ct.c_int, ct.c_int, c_dp, c_dp, c_dp, c_dp]
        L.oracle_qr.argtypes = [ct.c_int, ct.c_int, c_dp, c_dp, c_dp]
        L.oracle_lq.argtypes = [ct.c_int, ct.c_int, c_dp, c_dp, c_dp]
        L.oracle_expm_krylov_dense.argtypes = [ct.c_int, c_dp, c_dp, ct.c_double, ct.c_int, ct.c_double, c_dp]
        L.oracle_tridiag_exp.argtypes = [ct.c_int, c_dp, c_dp, ct.c_double, c_dp]
        L.oracle_project_site.argtypes = [ct.c_int] * 5 + [c_dp] * 5
        L.oracle_env_left.argtypes = [ct.c_int] * 5 + [c_dp] * 4
        L.oracle_env_right.argtypes = [ct.c_int] * 5 + [c_dp] * 4
        L.oracle_stochastic_factor.restype = ct.c_double
        L.oracle_stochastic_factor.argtypes = [ct.c_int64, ct.c_void_p]
        L.oracle_prob_distribution.argtypes = [ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_double,
                                               c_dp, ct.POINTER(ct.c_int)]
        L.oracle_set_window.argtypes = [ct.c_int]
        L.oracle_mpo_jump.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int, ct.c_int,
                                      ct.POINTER(ct.c_int64), c_dp, ct.c_int64]
        L.oracle_mpo_jump.restype = ct.c_int64
        L.oracle_last_window_seconds.restype = ct.c_double
        L.oracle_tdvp_evolve.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int, ct.c_int,
                                         ct.c_double, ct.c_int, ct.c_void_p, c_dp]
        L.oracle_tdvp_overlap.argtypes = [ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int, ct.c_int,
                                          ct.c_int, ct.c_double, c_dp]
        L.oracle_measure_shots.argtypes = [ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_uint64,
                                           ct.POINTER(ct.c_int64)]
        L.oracle_use_blas.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_char_p]
        L.oracle_blas_bound.restype = ct.c_int
        L.oracle_use_lapack_svd.argtypes = [ct.c_char_p, ct.c_char_p]
        _lib = L
    return _lib


def openblas_path():
    """numpy's bundled OpenBLAS (0.3.29, 64-bit integers: the BLAS version of
    the reference's Manifest), or None."""
    import glob
    d = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    hits = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    return hits[0] if hits else None


def use_blas(on: bool = True) -> bool:
    """Timing baseline only: route the oracle's larger products through
    OpenBLAS zgemm and its SVDs through LAPACK zgesdd (the library and
    version the reference links, one thread per call), or back to its own
    loops and Jacobi SVD, which the parity tests pin (the default).  Returns
    whether BLAS is bound."""
    L = lib()
    path = openblas_path() if on else None
    if on and path:
        L.oracle_use_blas(path.encode(), b"scipy_zgemm_64_", b"scipy_openblas_set_num_threads64_")
        L.oracle_use_lapack_svd(path.encode(), b"scipy_zgesdd_64_")
    else:
        L.oracle_use_blas(None, None, None)
    return bool(L.oracle_blas_bound())


def _ptr(a):
    return a.ctypes.data_as(c_dp)


def cm(a):
    """Complex array -> interleaved column-major doubles."""
    return np.ascontiguousarray(np.asfortranarray(np.asarray(a, np.complex128)).ravel(order="F")).view(np.float64)


def from_cm(buf, shape):
    return buf.view(np.complex128).reshape(shape, order="F")


def run_analog(holder, nthreads: int = 0):
    """Returns (traj [obs, col, traj], mean [obs, col], stats dict)."""
    L = lib()
    job = holder.job
    ncol = int(L.oracle_analog_num_samples(ct.byref(job)))
    nobs, ntraj = job.n_obs, job.num_traj
    traj = np.zeros(max(1, nobs * ncol * ntraj))
    mean = np.zeros(max(1, nobs * ncol))
    st = np.zeros(6)
    rc = L.oracle_run_analog(ct.byref(job), _ptr(traj), _ptr(mean), nthreads, _ptr(st))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    stats = dict(zip(["flop", "matvec2", "matvec1", "svd", "env", "jumps"], st.tolist()))
    return traj[: nobs * ncol * ntraj].reshape(nobs, ncol, ntraj), mean[: nobs * ncol].reshape(nobs, ncol), stats


def run_analog_window(holder, warmup: int, nthreads: int = 0) -> float:
    """Runs the job and returns the wall time of steps warmup+1.. of its
    slowest trajectory (trajectories run concurrently, one per thread)."""
    L = lib()
    L.oracle_set_window(int(warmup))
    try:
        run_analog(holder, nthreads=nthreads)
    finally:
        L.oracle_set_window(-1)
    return float(L.oracle_last_window_seconds())


def run_digital(holder, nthreads: int = 0):
    L = lib()
    job = holder.job
    ncol = int(L.oracle_digital_num_samples(ct.byref(job)))
    nobs, ntraj = job.n_obs, job.num_traj
    traj = np.zeros(max(1, nobs * ncol * ntraj))
    mean = np.zeros(max(1, nobs * ncol))
    bonds = np.zeros(max(1, ntraj * ncol), dtype=np.int64)
    st = np.zeros(6)
    rc = L.oracle_run_digital(ct.byref(job), _ptr(traj), _ptr(mean), bonds.ctypes.data_as(ct.POINTER(ct.c_int64)),
                              nthreads, _ptr(st))
    if rc:
        raise RuntimeError(f"oracle error {rc}")
    return (traj[: nobs * ncol * ntraj].reshape(nobs, ncol, ntraj), mean[: nobs * ncol].reshape(nobs, ncol),
            bonds[: ntraj * ncol].reshape(ntraj, ncol), {"jumps": st[0]})


def rng_uniform(seed, traj, k):
    return lib().oracle_rng_uniform(seed, traj, k)


def svd(A):
    m, n = A.shape
    r = min(m, n)
    U, S, Vt = np.zeros(2 * m * r), np.zeros(r), np.zeros(2 * r * n)
    lib().oracle_svd(m, n, _ptr(cm(A)), _ptr(U), _ptr(S), _ptr(Vt))
    return from_cm(U, (m, r)), S, from_cm(Vt, (r, n))


def qr(A):
    m, n = A.shape
    r = min(m, n)
    Q, R = np.zeros(2 * m * r), np.zeros(2 * r * n)
    lib().oracle_qr(m, n, _ptr(cm(A)), _ptr(Q), _ptr(R))
    return from_cm(Q, (m, r)), from_cm(R, (r, n))


def lq(A):
    m, n = A.shape
    r = min(m, n)
    Lm, Q = np.zeros(2 * m * r), np.zeros(2 * r * n)
    lib().oracle_lq(m, n, _ptr(cm(A)), _ptr(Lm), _ptr(Q))
    return from_cm(Lm, (m, r)), from_cm(Q, (r, n))


def expm_krylov_dense(H, v, dt, kmax=25, tol=1e-10):
    n = H.shape[0]
    out = np.zeros(2 * n)
    it = lib().oracle_expm_krylov_dense(n, _ptr(cm(H)), _ptr(cm(v)), dt, kmax, tol, _ptr(out))
    return out.view(np.complex128).copy(), it


def tridiag_exp(alpha, beta, tau):
    k = len(alpha)
    out = np.zeros(2 * k)
    b = np.zeros(k)
    b[1:] = beta[1:k]
    lib().oracle_tridiag_exp(k, _ptr(np.asarray(alpha, float)), _ptr(b), tau, _ptr(out))
    return out.view(np.complex128).copy()


def project_site(x, Lenv, W, Renv):
    kl, p, kr = x.shape
    ml, mr = W.shape[0], W.shape[3]
    out = np.zeros(2 * kl * p * kr)
    lib().oracle_project_site(kl, p, kr, ml, mr, _ptr(cm(x)), _ptr(cm(Lenv)), _ptr(cm(W)), _ptr(cm(Renv)),
                              _ptr(out))
    return from_cm(out, (kl, p, kr))


def env_left(A, W, E):
    al, p, ar = A.shape
    ml, mr = W.shape[0], W.shape[3]
    out = np.zeros(2 * ar * mr * ar)
    lib().oracle_env_left(al, p, ar, ml, mr, _ptr(cm(A)), _ptr(cm(W)), _ptr(cm(E)), _ptr(out))
    return from_cm(out, (ar, mr, ar))


def env_right(A, W, E):
    al, p, ar = A.shape
    ml, mr = W.shape[0], W.shape[3]
    out = np.zeros(2 * al * ml * al)
    lib().oracle_env_right(al, p, ar, ml, mr, _ptr(cm(A)), _ptr(cm(W)), _ptr(cm(E)), _ptr(out))
    return from_cm(out, (al, ml, al))


def tdvp_evolve(holder_analog, steps):
    """Noise-free 2-site TDVP sweeps with per-step 1-site observables."""
    job = holder_analog.job
    nobs = job.n_obs
    out = np.zeros(nobs * (steps + 1))
    lib().oracle_tdvp_evolve(job.L, ct.cast(job.mps, ct.c_void_p), ct.cast(job.mpo, ct.c_void_p), job.dt, steps,
                             int(job.max_bond_dim), job.truncation_threshold, nobs, ct.cast(job.obs, ct.c_void_p),
                             _ptr(out))
    return out.reshape(nobs, steps + 1)


def tdvp_overlap(holder_analog, steps, sites):
    """<ψ0|ψ(t)> after `steps` noise-free sweeps of two_site_tdvp! (sites=2)
    or single_site_tdvp! (sites=1)."""
    job = holder_analog.job
    out = np.zeros(2)
    lib().oracle_tdvp_overlap(job.L, ct.cast(job.mps, ct.c_void_p), ct.cast(job.mpo, ct.c_void_p), job.dt, steps,
                              sites, int(job.max_bond_dim), job.truncation_threshold, _ptr(out))
    return complex(out[0], out[1])


def measure_shots_bits(mps, shots, seed):
    """Per-shot bitstrings of single_shot_measure with the Philox draws
    (seed, shot, site)."""
    from yaqs_julia_amd import _abi
    h = _abi._Holder()
    arr = _abi._pack_mps(h, mps)
    out = np.zeros(max(1, shots), dtype=np.int64)
    lib().oracle_measure_shots(mps.length, ct.cast(arr, ct.c_void_p), shots, seed,
                               out.ctypes.data_as(ct.POINTER(ct.c_int64)))
    return out[:shots]


def stochastic_factor(holder):
    job = holder.job
    return lib().oracle_stochastic_factor(job.L, ct.cast(job.mps, ct.c_void_p))


def mpo_jump(holder, proc_index, thr, max_bond, mode=0):
    """apply_mpo_jump! of noise process `proc_index` (mode 0) or MPS.truncate!
    (mode 1) on the holder's MPS (centre field at L); returns the site
    tensors (dl, d, dr) as numpy arrays."""
    job = holder.job
    L = job.L
    dims = np.zeros(3 * L, dtype=np.int64)
    cap = 1 << 22
    data = np.zeros(cap)
    procp = ct.cast(ct.byref(job.noise[proc_index]), ct.c_void_p) if mode == 0 else None
    n = lib().oracle_mpo_jump(L, ct.cast(job.mps, ct.c_void_p), procp, thr, max_bond, mode,
                              dims.ctypes.data_as(ct.POINTER(ct.c_int64)), _ptr(data), cap)
    assert n >= 0
    out, off = [], 0
    for i in range(L):
        dl, d, dr = (int(x) for x in dims[3 * i:3 * i + 3])
        m = dl * d * dr
        out.append(from_cm(data[off:off + 2 * m].copy(), (dl, d, dr)))
        off += 2 * m
    return out


def prob_distribution(holder, dt):
    job = holder.job
    n = max(1, 4 * job.L * max(1, job.n_noise))
    probs = np.zeros(n)
    cand = np.zeros(3 * n, dtype=np.int32)
    k = lib().oracle_prob_distribution(job.L, ct.cast(job.mps, ct.c_void_p), job.n_noise,
                                       ct.cast(job.noise, ct.c_void_p), dt, _ptr(probs),
                                       cand.ctypes.data_as(ct.POINTER(ct.c_int)))
    return probs[:k], cand[: 3 * k].reshape(k, 3)
