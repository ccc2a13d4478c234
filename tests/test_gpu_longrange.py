"""Long-range (MPO) noise processes on the analog GPU path against the FP64
oracle, per trajectory (src/Noise.jl:22-33, 156-295, 372-408;
src/Dissipation.jl:161-204; src/StochasticProcess.jl:46-82, 168-198,
255-327; src/MPS.jl:320-441).

Ising L=6 (J=1, g=0.5), lowering on every site and a long-range process on
sites (2, 5) and (1, 6), dt=0.05, T=1, from a random MPS whose bonds have
full rank: an MPO jump on a product state (|000000> at t=0) keeps exactly
zero singular values under truncate!'s keep-two rule, and the directions
completing them -- which LAPACK's SVD picks in the reference -- then steer
the TDVP that follows, as for the TDVP from product states.  FP64 per
trajectory within 1e-10 with equal jump counts for the standard crosstalk
("pauli_long_range": scalar dissipation, weight dt gamma |psi|^2, no operator
applied), its projector, two-point and Gaussian unitary unravelings (MPO
jumps: O applied on doubled bonds, truncate!, normalize!); the mixed
precision ensemble within 2e-2 of the FP64 ensemble on the same seeds (it
follows the same jumps unless a draw falls within rounding of a boundary).
Digital circuits with a long-range gate on (L, 1) and a crosstalk on that
pair (the reference's paper experiments), its gates as SWAP networks or
TDVP windows: FP64 per trajectory within 1e-10, jump counts equal.
"""
import numpy as np
import pytest

import yaqs_julia_amd as Y
from yaqs_julia_amd import _abi

pytestmark = pytest.mark.gpu

L, DT, STEPS = 6, 0.05, 20


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def _noise(unravel, gamma=0.6):
    procs = [{"name": "lowering", "sites": [i], "strength": 0.05} for i in range(1, L + 1)]
    for sites in ([2, 5], [6, 1]):
        d = {"name": "crosstalk_xz", "sites": sites, "strength": gamma}
        if unravel != "standard":
            d["unraveling"] = unravel
        if unravel == "unitary_2pt":
            d["theta0"] = 0.9
        if unravel == "unitary_gauss":
            d.update({"sigma": 0.8, "M": 5})
        procs.append(d)
    return Y.NoiseModel(procs, L)


def _generic_mps(seed=17):
    """Random MPS with full-rank bonds (dimension min(2^i, 2^(L-i)))."""
    rng = np.random.default_rng(seed)
    dims = [min(2 ** i, 2 ** (L - i)) for i in range(L + 1)]
    ts = [rng.standard_normal((dims[i], 2, dims[i + 1])) + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))
          for i in range(L)]
    m = Y.MPS(L, tensors=ts)
    m.normalize()
    return m


def _job(noise, ntraj, precision, seed=11):
    obs = [Y.Observable(f"z{s}", Y.gates.ZGate(), s) for s in range(1, L + 1)] + \
          [Y.Observable(f"x{s}", Y.gates.XGate(), s) for s in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, DT * STEPS, dt=DT, num_traj=ntraj, max_bond_dim=8, truncation_threshold=1e-10)
    h = _abi.pack_analog(_generic_mps(), Y.init_ising(L, 1.0, 0.5), cfg, noise, precision=precision, seed=seed)
    return h, cfg


def _gpu(ctx, h, cfg):
    ncol = len(cfg.times)
    return ctx.wait(ctx.submit_analog(h), len(cfg.observables), ncol, cfg.num_traj)


@pytest.mark.parametrize("unravel", ["standard", "projector", "unitary_2pt", "unitary_gauss"])
def test_longrange_fp64_parity(ctx, oracle_lib, unravel):
    print(f"case: long-range {unravel} L={L} fp64", flush=True)
    ntraj = 24
    noise = _noise(unravel)
    h, cfg = _job(noise, ntraj, _abi.PREC_FP64)
    g, gm, _, st = _gpu(ctx, h, cfg)
    hr, _ = _job(noise, ntraj, _abi.PREC_FP64)
    r, rm, rst = oracle_lib.run_analog(hr, nthreads=8)
    err = float(np.abs(g - r).max())
    print(f"  jumps gpu {st['jumps']} oracle {rst['jumps']}  max|gpu-oracle| = {err:.3e}", flush=True)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert err < 1e-10


@pytest.mark.parametrize("unravel", ["projector", "unitary_2pt"])
def test_longrange_mixed_ensemble(ctx, unravel):
    print(f"case: long-range {unravel} L={L} mixed vs fp64 ensemble", flush=True)
    ntraj = 128
    noise = _noise(unravel)
    h, cfg = _job(noise, ntraj, _abi.PREC_FP64)
    _, m64, _, _ = _gpu(ctx, h, cfg)
    h, cfg = _job(noise, ntraj, _abi.PREC_MIXED)
    _, m32, _, _ = _gpu(ctx, h, cfg)
    err = float(np.abs(m64 - m32).max())
    print(f"  max|mixed-fp64| ensemble = {err:.3e}", flush=True)
    assert err < 2e-2


def _lr_circuit(depth, seed):
    """Brickwork of Haar gates plus, per layer, a long-range gate on (L, 1)
    (the periodic bond of the reference's paper experiments,
    01_PaperExps/experiments_file.jl:540-552) and single-qubit rotations."""
    rng = np.random.default_rng(seed)
    c = Y.DigitalCircuit(L)
    for layer in range(depth):
        for s in range(1 + layer % 2, L, 2):
            c.add_gate(Y.gates.MatrixGate(Y.gates.haar_su4(rng)), [s, s + 1])
        c.add_gate(Y.gates.RzzGate(rng.uniform(-1, 1)), [L, 1])
        c.add_gate(Y.gates.RxGate(rng.uniform(-1, 1)), [int(rng.integers(1, L + 1))])
    c.add_sample_barrier()
    return c


def _digital_noise(unravel):
    procs = [{"name": n, "sites": [i], "strength": 0.01} for i in range(1, L + 1) for n in ("pauli_x", "pauli_z")]
    d = {"name": "crosstalk_zz", "sites": [L, 1], "strength": 0.08}
    if unravel != "standard":
        d["unraveling"] = unravel
    if unravel == "unitary_2pt":
        d["theta0"] = np.pi / 3
    procs.append(d)
    return Y.NoiseModel(procs, L)


@pytest.mark.parametrize("method", ["TEBD", "TDVP"])
@pytest.mark.parametrize("unravel", ["standard", "projector", "unitary_2pt", "unitary_gauss"])
def test_longrange_digital_fp64_parity(ctx, oracle_lib, unravel, method):
    """Digital circuits with a long-range gate on (L, 1) and a long-range
    crosstalk on the same pair (applied after that gate only,
    create_local_noise_model, src/DigitalTJM.jl:172-188): FP64 per
    trajectory against the oracle, jump counts equal."""
    print(f"case: digital long-range noise {unravel} gates {method}", flush=True)
    ntraj = 16
    circ = _lr_circuit(4, 5)
    noise = _digital_noise(unravel)
    obs = [Y.Observable(f"z{s}", Y.gates.ZGate(), s) for s in range(1, L + 1)] + \
          [Y.Observable(f"x{s}", Y.gates.XGate(), s) for s in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=8, truncation_threshold=1e-10)
    opts = Y.TJMOptions("TEBD", method)  # the Haar gates have no generator
    hg = _abi.pack_digital(_generic_mps(), circ, cfg, noise, precision=_abi.PREC_FP64, seed=3, alg_options=opts)
    ncol = int(ctx.lib.yaqs_digital_num_samples(hg.job))
    g, _, _, st = ctx.wait(ctx.submit_digital(hg), len(obs), ncol, ntraj)
    hr = _abi.pack_digital(_generic_mps(), circ, cfg, noise, seed=3, alg_options=opts)
    r, _, _, rst = oracle_lib.run_digital(hr, nthreads=8)
    err = float(np.abs(g - r).max())
    print(f"  jumps gpu {st['jumps']} oracle {rst['jumps']}  max|gpu-oracle| = {err:.3e}", flush=True)
    assert st["jumps"] == rst["jumps"] and st["jumps"] > 0
    assert err < 1e-10
