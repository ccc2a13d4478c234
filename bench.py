"""Throughput benchmark of the batched TJM engine on one MI355X.

Workload (the headline configuration): Heisenberg XXX chain L=30 from the
Neel state, lowering + pauli_z noise gamma=0.01 on every site, 2-site TDVP
dt=0.05, chi_max=64, threshold 1e-9, Krylov dim <= 25, observables Z_i and
X_i, 1000 trajectories, complex-FP32 contractions with FP64 Lanczos / Gram /
norms ("mixed").  One step = one TJM time step of every trajectory (the
step_through and sample of analog_tjm_2: two TDVP sweeps, dissipation,
jumps, 60 observables).  The job starts at t=0; the timed window is K steps
(--steps) that start at saturated bonds: the untimed warm-up is
max(--warmup, 19) steps, since the bulk bonds reach chi=64 by step 19 and
every later step of the T=4 job costs the same.  Metric:
trajectory-site-steps per second = ntraj * L * K / window seconds.

Also reported, on the same steps: batched FP64 on the GPU (the precision
gain, separated from the batching gain), the plain per-trajectory FP64 GPU
path (one trajectory per batch, on a trajectory subset), the ensemble
<Z_i> and <X_i> error of the mixed run against batched FP64 over all
trajectories, and the FP64 CPU oracle (a bounded trajectory subset, one
trajectory per thread, OpenBLAS zgemm / LAPACK zgesdd as the reference
links them), plus the roofline of the dominant kernel (the complex-FP32
MFMA GEMM) from HIP events around every launch in the window, in a run of
the job as one wave before the timed run (the timed run's concurrent waves
share the device; the one-wave run also lets the launch tuner time its
variants alone, and the timed run takes the choices it cached).
"""
import argparse
import ctypes as ct
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "yaqs-julia_amd")
FP32_MFMA_PEAK = 157.3e12  # MI355X dense FP32 matrix peak, FLOP/s
SATURATED_WARMUP = 19      # steps from t=0 until the bulk bonds reach chi_max=64
METRIC = "traj-site-steps/s (L=30,\u03c7=64, \u27e8Z\u27e9 err\u22641e-4) vs plain GPU; p50/p99 job latency"


def _pkg():
    if "yaqs_julia_amd" in sys.modules:
        return sys.modules["yaqs_julia_amd"]
    spec = importlib.util.spec_from_file_location("yaqs_julia_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["yaqs_julia_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _pmc_traffic():
    """HBM bytes per window launch of the GEMM kernel from the committed PMC
    passes (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration);
    None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "c2_gemm_mfma_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def c2_job(Y, ntraj, nsteps, precision, L=30, chi=64):
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)] + \
          [Y.Observable(f"x{i}", Y.gates.XGate(), i) for i in range(1, L + 1)]
    cfg = Y.TimeEvolutionConfig(obs, nsteps * 0.05, dt=0.05, num_traj=ntraj, max_bond_dim=chi,
                                truncation_threshold=1e-9)
    assert len(cfg.times) == nsteps + 1
    H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
    noise = Y.relaxation_dephasing(L, 0.01)
    mps = Y.MPS(L, state="Neel")
    mps.normalize()
    prec = Y._abi.PREC_MIXED if precision == "mixed" else Y._abi.PREC_FP64
    return Y._abi.pack_analog(mps, H, cfg, noise, precision=prec, seed=0), cfg


def gpu_window(Y, lib, ctx, ntraj, warmup, steps, precision):
    h, cfg = c2_job(Y, ntraj, warmup + steps, precision)
    nobs, ncol = len(cfg.observables), len(cfg.times)
    traj = np.zeros(nobs * ncol * ntraj)
    mean = np.zeros(nobs * ncol)
    rep = Y._abi.BenchReport()
    rc = lib.yaqs_bench_analog(ctx.handle, ct.byref(h.job), warmup, traj.ctypes.data_as(Y._abi.c_dp),
                               mean.ctypes.data_as(Y._abi.c_dp), ct.byref(rep))
    if rc != 0:
        raise RuntimeError(f"yaqs_bench_analog failed ({rc}): {lib.yaqs_last_error().decode()}")
    out = rep.as_dict()
    out["traj"] = traj.reshape(nobs, ncol, ntraj)
    return out, mean.reshape(nobs, ncol)


def host_info():
    """Host CPU model, logical CPUs of the machine and CPUs this process may
    run on (the box's share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "cpus_allowed": allowed}


def cpu_window(Y, ntraj, warmup, steps, nthreads):
    """FP64 oracle on the same job restricted to `ntraj` trajectories (one per
    OpenMP thread) with the reference's BLAS/LAPACK (OpenBLAS zgemm, LAPACK
    zgesdd, one thread per call): wall time of steps warmup+1 ..
    warmup+steps, measured inside the oracle for every trajectory (longest
    one reported)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O
    blas = O.use_blas(True)
    try:
        h, _ = c2_job(Y, ntraj, warmup + steps, "fp64")
        return O.run_analog_window(h, warmup, nthreads=nthreads), blas
    finally:
        O.use_blas(False)


def c3_job(Y, ntraj, layers=40, L=40, chi=128, precision="mixed", p=0.01):
    """Noisy brickwork circuit: L=40 qubits, `layers` layers of Haar-random
    SU(4) gates (odd layers on bonds (1,2),(3,4)..., even on (2,3)...),
    depolarizing p as pauli_x/y/z with strength p/3 on every qubit (applied
    after each gate on its two qubits, dt=1), TEBD threshold 1e-9, chi_max=128,
    <Z_i> at the end."""
    circ = Y.haar_brickwork(L, layers, seed=0)
    obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
    cfg = Y.StrongMeasurementConfig(obs, num_traj=ntraj, max_bond_dim=chi, truncation_threshold=1e-9)
    procs = [{"name": n, "sites": [i], "strength": p / 3} for i in range(1, L + 1)
             for n in ("pauli_x", "pauli_y", "pauli_z")]
    noise = Y.NoiseModel(procs, L)
    prec = Y._abi.PREC_MIXED if precision == "mixed" else Y._abi.PREC_FP64
    return Y._abi.pack_digital(Y.MPS(L), circ, cfg, noise, precision=prec, seed=0), cfg


def c3_main(Y, args):
    """Digital workload line: trajectory-site-layers per second of the whole
    job (all trajectories in one call, from |0...0>)."""
    L, layers = 40, args.layers
    ctx = Y.Context(0)
    h, cfg = c3_job(Y, args.ntraj, layers=layers, precision=args.precision)
    ncol = int(ctx.lib.yaqs_digital_num_samples(h.job))
    print(f"[bench] C3: L={L} layers={layers} ntraj={args.ntraj} {args.precision}", flush=True)
    t0 = time.perf_counter()
    jid = ctx.submit_digital(h)
    _, mean, bonds, st = ctx.wait(jid, len(cfg.observables), ncol, args.ntraj, want_bonds=True)
    wall = time.perf_counter() - t0
    ctx.close()
    value = args.ntraj * L * layers / st["wall_s"]
    print(f"[bench] C3 GPU {st['wall_s']:.2f}s -> {value:.4g} traj-site-layers/s, max bond {int(bonds.max())}",
          flush=True)
    out = {
        "metric": "traj-site-layers/s (digital TJM, L=40 brickwork)", "value": value,
        "unit": "traj-site-layers/s", "n_gpus": 1, "steps": layers, "warmup": 0,
        "ms_per_step": 1e3 * st["wall_s"] / layers, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 contractions + f64 Gram/eigensolver" if args.precision == "mixed" else "c128",
        "data": "synthetic (Haar SU(4) brickwork from seed 0, seeded jumps)",
        "config": {"workload": "C3: L=40, Haar SU(4) brickwork, depolarizing p=0.01 as pauli_x/y/z p/3 per gate "
                               "qubit, TEBD thr 1e-9, chi_max=128, <Z_i> at the end",
                   "ntraj": args.ntraj, "L": L, "layers": layers, "chi_max": 128, "precision": args.precision,
                   "max_bond_reached": int(bonds.max()), "host_wall_s": wall},
        "roofline": None, "cpu_baseline": None,
    }
    if not args.no_cpu:
        # bounded sample: the oracle follows the reference's full-chain noise
        # sweeps in FP64 and costs about a minute per saturated layer and
        # trajectory, so it runs the first cpu_layers layers of the same
        # circuit (bond dimension still growing: an upper bound of its rate)
        if ROOT not in sys.path:
            sys.path.insert(0, ROOT)
        from oracle import oracle as O
        nthreads = args.cpu_traj or min(16, os.cpu_count() or 1)
        lay = args.cpu_layers
        hc, _ = c3_job(Y, nthreads, layers=lay, precision="fp64")
        print(f"[bench] CPU oracle, {nthreads} trajectories on {nthreads} threads, layers 1..{lay}", flush=True)
        t0 = time.perf_counter()
        O.run_digital(hc, nthreads=nthreads)
        t = time.perf_counter() - t0
        cv = nthreads * L * lay / t
        out["cpu_baseline"] = {"value": cv, "unit": "traj-site-layers/s", "cores": nthreads, "kind": "port",
                               "sample": f"FP64 oracle, {nthreads} trajectories (one per thread), the first {lay} "
                                         "layers of the same circuit"}
        print(f"[bench] CPU oracle: {cv:.4g} traj-site-layers/s on {nthreads} threads", flush=True)
    print(json.dumps(out), flush=True)


def c4_jobs(Y, njobs, seed=0, T=2.0, ntraj_range=(64, 1024), force=None):
    """Jobs of the mixed-request workload, drawn from `seed`: model in {Ising
    g ~ U[0.2, 1.5] (D=3), XY = init_general_hamiltonian(L,1,1,0,0,0,0),
    Heisenberg (D=5)}; L ~ U{10..64}; chi_max in {16, 32, 64, 128, 256};
    relaxation + dephasing gamma log-uniform in [1e-3, 1e-1]; trajectories
    ~ U{ntraj_range}; T (2 in the workload), dt=0.05; Neel start for
    Heisenberg, |0..0> otherwise; observables Z_i; mixed precision; seeds
    drawn from the stream.  `force` overrides fields of job 0 (e.g. the
    largest L and chi).  Returns (holder, cfg, L, ntraj, desc) per job."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(njobs):
        L = int(rng.integers(10, 65))
        model = int(rng.integers(0, 3))
        g = float(rng.uniform(0.2, 1.5))
        chi = int(rng.choice([16, 32, 64, 128, 256]))
        gamma = float(10 ** rng.uniform(-3, -1))
        ntraj = int(rng.integers(ntraj_range[0], ntraj_range[1] + 1))
        jseed = int(rng.integers(0, 2 ** 31))
        if k == 0 and force:
            L = force.get("L", L)
            chi = force.get("chi", chi)
            model = force.get("model", model)
        if model == 0:
            H = Y.init_ising(L, 1.0, g)
        elif model == 1:
            H = Y.init_general_hamiltonian(L, 1, 1, 0, 0, 0, 0)
        else:
            H = Y.init_general_hamiltonian(L, 1, 1, 1, 0, 0, 0)
        obs = [Y.Observable(f"z{i}", Y.gates.ZGate(), i) for i in range(1, L + 1)]
        cfg = Y.TimeEvolutionConfig(obs, T, dt=0.05, num_traj=ntraj, max_bond_dim=chi, truncation_threshold=1e-9)
        mps = Y.MPS(L, state="Neel" if model == 2 else "zeros")
        mps.normalize()
        h = Y._abi.pack_analog(mps, H, cfg, Y.relaxation_dephasing(L, gamma), precision=Y._abi.PREC_MIXED,
                               seed=jseed)
        desc = f"{['Ising', 'XY', 'Heisenberg'][model]} L={L} chi={chi} gamma={gamma:.1e} ntraj={ntraj}"
        out.append((h, cfg, L, ntraj, desc))
    return out


def serve_stream(Y, njobs, seed=0, mean_gap_s=0.25, T=0.5, ntraj_range=(64, 256)):
    """Poisson arrivals (mean gap `mean_gap_s`, short enough to keep the
    device saturated) of `njobs` jobs of the mixed-request workload
    (c4_jobs: models, L 10-64, chi 16-256 and rates as in the workload; by
    default T=0.5 and 64-256 trajectories instead of T=2 and 64-1024, which
    keeps the stream to about a minute: one Heisenberg L=40 chi=128 job of
    the full workload alone takes about a minute); latency is submit -> done
    as measured by the library."""
    rng = np.random.default_rng(seed + 1)
    ctx = Y.Context(0)
    specs = c4_jobs(Y, njobs, seed=seed, T=T, ntraj_range=ntraj_range)
    gaps = rng.exponential(mean_gap_s, njobs)
    t0 = time.perf_counter()
    ids = []
    for (h, cfg, L, ntraj, _), gap in zip(specs, gaps):
        time.sleep(gap)
        ids.append(ctx.submit_analog(h))
    lat, work = [], 0
    for (h, cfg, L, ntraj, desc), jid in zip(specs, ids):
        _, _, _, st = ctx.wait(jid, len(cfg.observables), len(cfg.times), ntraj)
        lat.append(st["wall_s"])
        work += ntraj * L * (len(cfg.times) - 1)
        print(f"[bench] serving: {desc} done, latency {st['wall_s']:.2f}s", flush=True)
    span = time.perf_counter() - t0
    ctx.close()
    lat = np.array(lat)
    return {"jobs": njobs, "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "max_s": float(lat.max()), "throughput": work / span, "unit": "traj-site-steps/s",
            "makespan_s": span, "mean_interarrival_s": mean_gap_s,
            "stream": f"Poisson; Ising/XY/Heisenberg, L 10-64, chi 16-256, gamma 1e-3..1e-1, "
                      f"{ntraj_range[0]}-{ntraj_range[1]} traj, T={T} dt=0.05, mixed",
            "job_list": [d for (_, _, _, _, d) in specs]}


def _window_rows(rows, last, kernel):
    """rows: (order key, name, ...) sorted by the key; the window starts at
    the first of the last `last` dispatches whose name contains `kernel`."""
    idx = [i for i, r in enumerate(rows) if kernel in r[1]]
    if not idx:
        return [], []
    sel = idx[-last:] if last > 0 else idx
    return rows[sel[0]:], [rows[i] for i in sel]


def trace_summary(path, last, kernel="gemm_mfma_kernel<float"):
    """rocprofv3 --kernel-trace CSV: average duration of the last `last`
    dispatches of `kernel` (the timed window is the end of the run), every
    kernel's calls and time from the first of them on, and the device's idle
    time there (gaps over 2 us between the union of dispatch intervals)."""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), r["Kernel_Name"].split("(")[0], int(r["End_Timestamp"])))
    rows.sort()
    win, sel = _window_rows(rows, last, kernel)
    dur = [e - s for s, _, e in sel]
    per = {}
    for s, name, e in win:
        c = per.setdefault(name, [0, 0])
        c[0] += 1
        c[1] += e - s
    idle, end = 0, win[0][0] if win else 0
    for s, _, e in win:
        if s > end + 2000:
            idle += s - end
        end = max(end, e)
    span = (end - win[0][0]) if win else 0
    top = sorted(per.items(), key=lambda kv: -kv[1][1])[:24]
    return {"kernel": kernel, "window_dispatches": len(sel),
            "avg_us_window": sum(dur) / max(1, len(dur)) / 1e3,
            "window_span_s": span / 1e9, "idle_s": idle / 1e9,
            "window_kernels": [{"name": k, "calls": v[0], "total_s": v[1] / 1e9, "avg_us": v[1] / v[0] / 1e3,
                                "share": v[1] / max(1, span)} for k, v in top]}


def pmc_summary(path, last, kernel="gemm_mfma_kernel<float", trace=None):
    """rocprofv3 --pmc CSV: per kernel name, the average per dispatch of
    every counter over the window (the dispatches from the first of the last
    `last` dispatches of `kernel` on); with the kernel trace of the same
    command, the average duration per dispatch and counter-bytes / duration
    (FETCH_SIZE and WRITE_SIZE are in KiB per dispatch)."""
    import csv
    disp = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = disp.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"].split("(")[0], "c": {}})
            d["c"][r["Counter_Name"]] = d["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    rows = sorted((k, v["name"], v["c"]) for k, v in disp.items())
    win, _ = _window_rows(rows, last, kernel)
    per = {}
    for _, name, c in win:
        a = per.setdefault(name, {"calls": 0})
        a["calls"] += 1
        for k, v in c.items():
            a[k] = a.get(k, 0.0) + v
    durs = {}
    if trace:
        t = trace_summary(trace, last, kernel)
        durs = {w["name"]: w["avg_us"] for w in t["window_kernels"]}
    out = []
    for name, a in sorted(per.items(), key=lambda kv: -kv[1]["calls"]):
        e = {"name": name, "calls": a["calls"]}
        for k, v in a.items():
            if k != "calls":
                e[k] = v / a["calls"]
        if name in durs:
            e["avg_us"] = durs[name]
            kib = e.get("FETCH_SIZE", 0.0) + e.get("WRITE_SIZE", 0.0)
            if kib:
                e["raw_GBps"] = kib * 1024 / (durs[name] * 1e-6) / 1e9
        out.append(e)
    return {"kernel": kernel, "window_dispatches_from": last, "per_kernel": out[:30]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=SATURATED_WARMUP,
                    help=f"untimed steps from t=0 (at least {SATURATED_WARMUP}: the window starts at saturated chi=64)")
    ap.add_argument("--ntraj", type=int, default=1000)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--plain-traj", type=int, default=2, help="trajectories for the plain FP64 per-trajectory path")
    ap.add_argument("--cpu-traj", type=int, default=0, help="oracle trajectories (0 = host threads, max 16)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle timed steps (bounded CPU sample)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the batched FP64 leg")
    ap.add_argument("--only", default="all", choices=["all", "roofline", "serve"],
                    help="roofline: run the one-wave roofline window alone; serve: the serving stream alone "
                         "(both for profiling)")
    ap.add_argument("--serve-jobs", type=int, default=12, help="jobs of the serving stream (0 = skip)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3"],
                    help="c2: the analog headline line; c3: the digital brickwork line")
    ap.add_argument("--layers", type=int, default=40, help="c3: circuit layers")
    ap.add_argument("--cpu-layers", type=int, default=7, help="c3: layers of the bounded oracle sample")
    ap.add_argument("--trace-summary", metavar="CSV", help="summarise a rocprofv3 kernel trace and exit")
    ap.add_argument("--trace-last", type=int, default=0, help="dispatches of the timed window (from the bench line)")
    ap.add_argument("--pmc-summary", metavar="CSV", help="summarise a rocprofv3 --pmc counter CSV and exit")
    ap.add_argument("--trace", metavar="CSV", help="with --pmc-summary: the kernel trace of the same command")
    ap.add_argument("--kernel", default="gemm_mfma_kernel<float", help="kernel that marks the window")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="launch a plain copy kernel of BYTES (4 times) and exit (counter calibration)")
    args = ap.parse_args()
    if args.trace_summary:
        print(json.dumps(trace_summary(args.trace_summary, args.trace_last, args.kernel)))
        return
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary, args.trace_last, args.kernel, args.trace)))
        return
    if args.calibrate_copy:
        lib = _pkg()._abi.load_library()
        lib.yaqs_debug_copy_bytes.argtypes = [ct.c_int64, ct.c_int]
        rc = lib.yaqs_debug_copy_bytes(args.calibrate_copy, 4)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "launches": 4, "rc": rc}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); trajectories shard across GPUs "
              "with a single reduction at the end, which this single-GPU build does not launch.", file=sys.stderr)
        sys.exit(2)
    Y = _pkg()
    if args.workload == "c3":
        if args.ntraj == 1000:
            args.ntraj = 512
        c3_main(Y, args)
        return
    lib = Y._abi.load_library()
    L = 30
    K = args.steps
    win0 = max(args.warmup, SATURATED_WARMUP)
    Kaux = min(K, 4)  # steps of the FP64, plain and roofline legs
    window = {"first_step": win0 + 1, "last_step": win0 + K, "t": [win0 * 0.05, (win0 + K) * 0.05],
              "rule": f"timed window starts at saturated bonds: warm-up = max(--warmup, {SATURATED_WARMUP}) steps"}
    if args.only == "serve":
        rep = serve_stream(Y, args.serve_jobs)
        print(json.dumps({"serving": rep}), flush=True)
        return
    if args.only == "roofline":
        # the roofline window alone (one wave), as profiled under rocprofv3
        rctx = Y.Context(0, max_batch=args.ntraj)
        rrep, _ = gpu_window(Y, lib, rctx, args.ntraj, win0, K, args.precision)
        rctx.close()
        print(json.dumps({"roofline_window": window, "timed_s": rrep["timed_s"],
                          "gemm_launches": int(rrep["gemm_launches"]), "gemm_ms": rrep["gemm_ms"],
                          "gemm_flop": rrep["gemm_flop"],
                          "achieved_tflops": rrep["gemm_flop"] / (rrep["gemm_ms"] * 1e-3) / 1e12}), flush=True)
        return
    # The roofline comes from a window where the GEMMs have the device to
    # their wave: with concurrent waves, the HIP events around a launch also
    # span kernels of the other waves' streams.  So the job runs first as one
    # wave (max_batch = ntraj) over the first Kaux steps of the window, for
    # the roofline; this also lets the launch tuner time its variants with
    # the device to itself, and the timed run below (several concurrent
    # waves) takes the choices it cached for the process.
    rctx = Y.Context(0, max_batch=args.ntraj)
    print(f"[bench] roofline window: steps {win0 + 1}..{win0 + Kaux} as one wave", flush=True)
    rrep, _ = gpu_window(Y, lib, rctx, args.ntraj, win0, Kaux, args.precision)
    rctx.close()
    rv = args.ntraj * L * Kaux / rrep["timed_s"]
    print(f"[bench] one wave: {rrep['timed_s']:.2f}s -> {rv:.4g} traj-site-steps/s", flush=True)
    ctx = Y.Context(0)
    print(f"[bench] C2 window: steps {win0 + 1}..{win0 + K} (warm-up {win0}, asked {args.warmup}) "
          f"ntraj={args.ntraj} {args.precision}", flush=True)
    rep, mean = gpu_window(Y, lib, ctx, args.ntraj, win0, K, args.precision)
    value = args.ntraj * L * K / rep["timed_s"]
    print(f"[bench] GPU window {rep['timed_s']:.2f}s -> {value:.4g} traj-site-steps/s, max bond {rep['max_bond']}",
          flush=True)
    roof_note = (f"steps {win0 + 1}..{win0 + Kaux} run as one wave of {args.ntraj} trajectories before the timed "
                 f"run ({rv:.0f} traj-site-steps/s; the timed run used {int(rep['concurrent_waves'])} concurrent "
                 f"wave(s)); reproduce with: python bench.py --only roofline --steps {Kaux}")
    achieved = rrep["gemm_flop"] / (rrep["gemm_ms"] * 1e-3) if rrep["gemm_ms"] > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "traj-site-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * rep["timed_s"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 contractions + f64 Lanczos/Gram/norms" if args.precision == "mixed" else "c128",
        "data": "synthetic (Neel initial state, seeded jumps)",
        "config": {
            "workload": "C2: Heisenberg XXX L=30 Neel, lowering+pauli_z gamma=0.01, 2-site TDVP dt=0.05, "
                        "chi_max=64, thr 1e-9, Krylov<=25, 60 observables",
            "ntraj": args.ntraj, "L": L, "chi_max": 64, "precision": args.precision,
            "window": window,
            "max_bond_reached": int(rep["max_bond"]),
            "concurrent_waves": int(rep["concurrent_waves"]),
            "flop_model_per_traj_site_step": rep["flop_model"] / (args.ntraj * L * K),
            "matvecs": int(rep["matvecs"]),
        },
        "roofline": {
            "bound": "mfma", "kernel": "gemm_mfma_kernel<float, opA, opB> (complex-FP32 real-split MFMA)",
            "achieved": achieved / 1e12, "peak": FP32_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": achieved / FP32_MFMA_PEAK, "traffic": _pmc_traffic(),
            "launches": int(rrep["gemm_launches"]),
            "flop_per_launch": rrep["gemm_flop"] / max(1, rrep["gemm_launches"]),
            "avg_launch_us": 1e3 * rrep["gemm_ms"] / max(1, rrep["gemm_launches"]),
            "gemm_share_of_window": rrep["gemm_ms"] * 1e-3 / rrep["timed_s"],
            "window": roof_note,
        },
    }
    if args.precision == "mixed" and not args.no_fp64:
        # batched FP64 on the same steps: the precision gain apart from the
        # batching gain, and the reference for the metric's error bound
        fctx = Y.Context(0)
        print(f"[bench] batched FP64 GPU, steps {win0 + 1}..{win0 + Kaux}", flush=True)
        frep, fmean = gpu_window(Y, lib, fctx, args.ntraj, win0, Kaux, "fp64")
        fctx.close()
        fv = args.ntraj * L * Kaux / frep["timed_s"]
        out["batched_fp64_gpu"] = {"value": fv, "unit": "traj-site-steps/s", "ntraj": args.ntraj,
                                   "steps": [win0 + 1, win0 + Kaux], "mixed_speedup": value / fv}
        print(f"[bench] batched FP64 GPU: {fv:.4g} traj-site-steps/s", flush=True)
        nc = win0 + Kaux + 1
        ez = float(np.abs(mean[:L, :nc] - fmean[:L, :nc]).max())
        ex = float(np.abs(mean[L:, :nc] - fmean[L:, :nc]).max())
        dt_traj = float(np.abs(rep["traj"][:, :nc, :] - frep["traj"][:, :nc, :]).max())
        out["accuracy"] = {"max_abs_err_Z": ez, "max_abs_err_X": ex, "max_abs_err_per_trajectory": dt_traj,
                           "trajectories": args.ntraj, "times": [0.0, (nc - 1) * 0.05], "tolerance": 1e-4,
                           "vs": "batched FP64 GPU path on the same seeds (it matches the FP64 oracle within 1e-10 "
                                 "per trajectory in tests/test_gpu_c2.py); ensemble <Z_i>, <X_i> at every time"}
        print(f"[bench] mixed vs FP64 ensemble over {args.ntraj} trajectories: <Z_i> {ez:.2e} <X_i> {ex:.2e}",
              flush=True)
    if not args.no_plain and args.plain_traj > 0:
        pctx = Y.Context(0, max_batch=1)
        print(f"[bench] plain FP64 per-trajectory GPU path, {args.plain_traj} trajectories", flush=True)
        prep, _ = gpu_window(Y, lib, pctx, args.plain_traj, win0, Kaux, "fp64")
        pctx.close()
        pv = args.plain_traj * L * Kaux / prep["timed_s"]
        out["plain_fp64_gpu"] = {"value": pv, "unit": "traj-site-steps/s", "ntraj": args.plain_traj,
                                 "batch": 1, "steps": [win0 + 1, win0 + Kaux], "speedup_of_batched": value / pv}
        print(f"[bench] plain FP64 per-trajectory GPU: {pv:.4g} traj-site-steps/s", flush=True)
    if args.serve_jobs > 0:
        print(f"[bench] serving stream, {args.serve_jobs} jobs", flush=True)
        out["serving"] = serve_stream(Y, args.serve_jobs)
        print(f"[bench] serving p50 {out['serving']['p50_s']:.2f}s p99 {out['serving']['p99_s']:.2f}s, "
              f"{out['serving']['throughput']:.4g} traj-site-steps/s", flush=True)
    ctx.close()
    if not args.no_cpu:
        hi = host_info()
        nthreads = args.cpu_traj or min(16, hi["cpus_allowed"])
        csteps = args.cpu_steps
        print(f"[bench] CPU oracle, {nthreads} trajectories on {nthreads} threads, steps "
              f"{win0 + 1}..{win0 + csteps} after {win0} untimed", flush=True)
        t, blas = cpu_window(Y, nthreads, win0, csteps, nthreads)
        cv = nthreads * L * csteps / t
        out["cpu_baseline"] = {
            "value": cv, "unit": "traj-site-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"FP64 oracle (CPU restatement of Yaqs.jl; OpenBLAS 0.3.29 zgemm/zgesdd, one thread each: "
                      f"{'bound' if blas else 'not found, own loops'}), trajectories 0..{nthreads - 1} one per "
                      f"thread, steps {win0 + 1}..{win0 + csteps} of the same job (the first step(s) of the GPU "
                      f"window), slowest trajectory's time",
            **hi}
        print(f"[bench] CPU oracle: {cv:.4g} traj-site-steps/s on {nthreads} threads ({hi['cpu_model']}, "
              f"nproc {hi['nproc']}, allowed {hi['cpus_allowed']})", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
