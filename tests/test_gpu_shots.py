"""yaqs_measure_shots (measure_shots / single_shot_measure, src/MPS.jl:
683-749) on the GPU.

- Right-canonical random MPS (L=10, chi=16): every shot's bitstring equals
  the oracle's for the same (seed, shot, site) draws (20000 shots).
- A non-canonical, unnormalised MPS: the library right-normalises a copy
  first, so the histogram still matches the exact |<b|ψ>|^2 (chi-square).
- L=30, chi=64, 2^20 shots (the shape of the headline workload's states):
  the first 2000 shots equal the oracle's; the shot rate is printed.
- MeasurementConfig and measure_shots' counts through the Python mirror.
"""
import time

import numpy as np
import pytest

import yaqs_julia_amd as Y

from test_oracle_shots import _probs_by_bits, chi2_ok, random_mps

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Y.Context(0)
    yield c
    c.close()


def test_shots_equal_oracle(ctx, oracle_lib):
    print("case: shots L=10 chi=16 canonical, 20000 shots vs oracle", flush=True)
    m = random_mps(10, 16, seed=1)
    g = ctx.measure_shots_bits(m, 20000, 9)
    r = oracle_lib.measure_shots_bits(m, 20000, 9)
    n_diff = int((g != r).sum())
    print(f"  shots differing from the oracle: {n_diff}", flush=True)
    assert n_diff == 0


def test_shots_non_canonical_input(ctx):
    print("case: shots L=8 chi=8 non-canonical unnormalised input, 100000 shots vs exact", flush=True)
    m = random_mps(8, 8, seed=2, normalize=False)
    n = 100000
    b = ctx.measure_shots_bits(m, n, 3)
    p = _probs_by_bits(m)
    p /= p.sum()
    counts = np.bincount(b, minlength=2 ** 8).astype(float)
    stat, dof, ok = chi2_ok(counts, p, n)
    print(f"  chi2 {stat:.1f} for {dof} dof", flush=True)
    assert ok


def test_shots_large(ctx, oracle_lib):
    print("case: shots L=30 chi=64, 2^20 shots", flush=True)
    m = random_mps(30, 64, seed=3)
    n = 1 << 20
    ctx.measure_shots_bits(m, 1024, 0)  # code objects and buffers
    t0 = time.perf_counter()
    g = ctx.measure_shots_bits(m, n, 11)
    dt = time.perf_counter() - t0
    print(f"  {n} shots in {dt:.3f} s: {n / dt:.3g} shots/s (L=30, chi=64)", flush=True)
    r = oracle_lib.measure_shots_bits(m, 2000, 11)
    assert np.array_equal(g[:2000], r)
    # every site's outcome occurs (a random state of this size has no
    # deterministic site)
    freq = np.array([((g >> i) & 1).mean() for i in range(30)])
    assert np.all((freq > 0.01) & (freq < 0.99))


def test_measurement_config_and_counts(ctx):
    print("case: MeasurementConfig / measure_shots counts, GHZ-like product", flush=True)
    cfg = Y.MeasurementConfig(500)
    res = Y.run_shots(Y.MPS(6, state="Neel"), cfg, seed=1, ctx=ctx)
    assert res == {0b101010: 500} and len(cfg.measurements) == 500
    counts = Y.measure_shots(Y.MPS(4, state="x+"), 4000, seed=2, ctx=ctx)
    assert sum(counts.values()) == 4000 and len(counts) == 16
    assert Y.single_shot_measure(Y.MPS(3, state="ones"), ctx=ctx) == 0b111
