/*
 * yaqs_hip.h — C-ABI of the MI355X Tensor Jump Method (TJM) trajectory engine.
 *
 * The library replaces the per-trajectory loop of Yaqs.jl's analog and
 * digital simulators with one call that runs every trajectory of a job
 * batched on one GPU.  Every entry point takes plain pointers and sizes; no
 * C++ or torch types cross this boundary, so Julia (`ccall`), Python
 * (`ctypes`) or C can bind it directly.
 *
 * Memory conventions (identical to Julia's in-memory layout, so a Julia
 * caller passes `pointer(A)` of its own arrays without copying):
 *   - complex numbers are interleaved (re, im) doubles (ComplexF64);
 *   - MPS site tensors are column-major (Dl, d, Dr)            — src/MPS.jl:33-44
 *   - MPO site tensors are column-major (Dl, d_out, d_in, Dr)  — src/MPO.jl:14-25
 *   - 2x2 / 4x4 operator matrices are column-major; a 4x4 two-site matrix
 *     acts on the combined index p1 + 2*p2 (site s1 least significant),
 *     which is how src/Dissipation.jl:86-88, src/StochasticProcess.jl:268-269
 *     and src/DigitalTJM.jl:265-268 reshape it to (2,2,2,2);
 *   - sites are 1-based, as in the reference.
 *
 * Ownership: inputs are caller-owned and deep-copied at submit; outputs are
 * caller-owned and written only inside yaqs_wait / yaqs_run_*.
 * Threading: a context owns four device worker threads (one HIP stream
 * each; YAQS_WORKERS=1..8 overrides) that run admitted trajectory waves
 * concurrently; every call is
 * thread-safe.  Errors: every call returns a yaqs_status; the message of the
 * last failure on the calling thread is yaqs_last_error().
 */
#ifndef YAQS_HIP_H
#define YAQS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YAQS_ABI_VERSION 4

/* Return codes.  Mirrors the reference's failure modes: Julia `error(...)` /
 * `@assert` on bad input (src/Simulator.jl:136,139), plus device failures. */
typedef enum {
    YAQS_OK = 0,
    YAQS_EINVAL = 1,         /* malformed job (bad sizes, sites, dims)        */
    YAQS_ENOMEM_DEVICE = 2,  /* job can never fit the HBM budget              */
    YAQS_EDEVICE = 3,        /* HIP runtime error                             */
    YAQS_ENUMERIC = 4,       /* NaN/Inf that survived the FP64 retry          */
    YAQS_ENOTFOUND = 5,      /* unknown job id                                */
    YAQS_EAGAIN = 6          /* yaqs_poll: job still running                  */
} yaqs_status;

typedef enum {
    YAQS_PREC_FP64 = 0,   /* complex FP64 everywhere (the "plain" path)          */
    YAQS_PREC_MIXED = 1   /* complex FP32 contractions, FP64 Lanczos/Gram/norms  */
} yaqs_precision;

/* Noise process kinds: the standard processes of src/Noise.jl:41-61. */
typedef enum {
    YAQS_NOISE_LOWERING = 0,
    YAQS_NOISE_RAISING = 1,
    YAQS_NOISE_PAULI_X = 2,
    YAQS_NOISE_PAULI_Y = 3,
    YAQS_NOISE_PAULI_Z = 4,
    YAQS_NOISE_CUSTOM = 5,   /* matrix given explicitly (1 or 2 adjacent sites) */
    YAQS_NOISE_LONG_RANGE = 6 /* MPONoiseProcess on two non-adjacent sites                */
} yaqs_noise_kind;

typedef struct {
    int64_t dl, d, dr;
    const double *data;      /* (dl*d*dr) ComplexF64, column-major */
} yaqs_mps_tensor;

typedef struct {
    int64_t dl, dout, din, dr;
    const double *data;      /* (dl*dout*din*dr) ComplexF64, column-major */
} yaqs_mpo_tensor;

/* One LocalNoiseProcess (src/Noise.jl:15-20).  `matrix` holds the 2x2
 * (first 8 doubles) or 4x4 (32 doubles) jump operator, column-major.
 * `pauli_name` is the name-based Pauli test of src/Dissipation.jl:14-21 and
 * src/StochasticProcess.jl:337-353 (names pauli_* / crosstalk_*).
 *
 * YAQS_NOISE_LONG_RANGE: one MPONoiseProcess (src/Noise.jl:22-33) on sites
 * site0 < site1 - 1, whose MPO is build_mpo_phys's O = α I + β σ_site0 τ_site1
 * (src/Noise.jl:77-128): matrix[0..7] = σ, matrix[8..15] = τ (2x2,
 * column-major), matrix[16..17] = α, matrix[18..19] = β.  Dissipation
 * scales site1 by exp(-dt/2 γ) (handle_longrange_scalar!,
 * src/Dissipation.jl:161-204).  `pauli_name` = 1 marks the candidate the
 * reference tags "pauli_long_range" (is_pauli_proc with factors,
 * src/StochasticProcess.jl:168-181): weight dt γ ‖ψ‖², and its jump applies
 * no operator (only the final normalize!, src/StochasticProcess.jl:317-330);
 * 0 the general MPO jump: weight dt γ ‖normalize!(truncate!(Oψ))‖², jump
 * apply_mpo_jump! (src/StochasticProcess.jl:46-82) = O applied, truncate!
 * (src/MPS.jl:320-441), normalize!.  In digital jobs it acts after a gate on
 * exactly its two sites (create_local_noise_model, src/DigitalTJM.jl:172-188). */
typedef struct {
    int32_t kind;
    int32_t nsites;          /* 1 or 2 (2 = adjacent pair)          */
    int64_t site0, site1;    /* 1-based; site1 ignored if nsites==1 */
    double strength;
    int32_t pauli_name;
    int32_t _pad;
    double matrix[32];
} yaqs_noise_process;

/* One Observable (src/SimulationConfigs.jl:20-34): 2x2 op on one site or
 * 4x4 op on an adjacent pair. */
typedef struct {
    int32_t nsites;
    int32_t _pad;
    int64_t site0, site1;
    double op[32];
} yaqs_observable;

/* Analog TJM job: the arguments of Simulator.run(ψ0, H::MPO, cfg, noise)
 * (src/Simulator.jl:127-141, src/Simulator.jl:22-69). */
typedef struct {
    int64_t L;
    const yaqs_mps_tensor *mps;        /* L tensors (normalized by the caller, src/Simulator.jl:128) */
    const yaqs_mpo_tensor *mpo;        /* L tensors                                                 */
    int32_t has_noise_model;           /* 0 = `nothing`                                             */
    int32_t tdvp_sites;                /* 0 or 2: two_site_tdvp! (the reference's analog path);
                                          1: single_site_tdvp! (src/Algorithms.jl:449-452, 471-532)
                                          in its place, bond dimensions fixed by the input MPS      */
    int64_t n_noise;
    const yaqs_noise_process *noise;
    int64_t n_obs;
    const yaqs_observable *obs;
    double dt;
    int64_t num_times;                 /* length(collect(0.0:dt:T)), src/SimulationConfigs.jl:89    */
    int64_t num_traj;
    int64_t max_bond_dim;
    double truncation_threshold;
    int32_t order;                     /* 1 = analog_tjm_1, else analog_tjm_2 (src/Simulator.jl:24) */
    int32_t sample_timesteps;
    int32_t krylov_dim;                /* 25 in the reference (src/Algorithms.jl:650)               */
    int32_t precision;                 /* yaqs_precision                                            */
    double krylov_tol;                 /* 1e-10 in the reference (src/Algorithms.jl:185)            */
    uint64_t seed;                     /* per-job seed of the counter-based RNG                     */
    uint64_t traj_offset;              /* global index of trajectory 0 (sharding)                   */
} yaqs_analog_job;

/* Digital TJM gates (src/DigitalTJM.jl:40-66).  A 1-qubit gate carries a 2x2
 * matrix, a 2-qubit gate a 4x4 matrix; YAQS_GATE_SAMPLE is the
 * Barrier("SAMPLE_OBSERVABLES") of src/DigitalTJM.jl:147-170. */
typedef enum {
    YAQS_GATE_1Q = 1,
    YAQS_GATE_2Q = 2,
    YAQS_GATE_SAMPLE = 3
} yaqs_gate_kind;

/* How a 2-qubit gate is applied (apply_window!, src/DigitalTJM.jl:307-358).
 *
 * YAQS_GATE_EXACT: TJMOptions local_method = :TEBD.  `matrix` is the gate;
 *   apply_local_gate_exact! (src/DigitalTJM.jl:251-305): Θ' = G Θ, SVD,
 *   drop singular values from the tail while the discarded weight stays
 *   <= threshold, at least 1 and at most max_bond_dim kept.
 * YAQS_GATE_WINDOW_TDVP: local_method = :TDVP (the reference default,
 *   src/DigitalTJM.jl:76).  `matrix` is the window Hamiltonian
 *   h = coeff * g_a (x) g_b of construct_window_mpo (src/DigitalTJM.jl:192-236,
 *   g_a on the lower site, in the same index order as gate matrices); the
 *   window of an adjacent pair is the pair itself, so the circuit sweep is
 *   its edge step alone (src/Algorithms.jl:606-627, 762-800): Θ' =
 *   exp(-i h) Θ (the Krylov space of h on Θ closes after at most four
 *   vectors, so the engine applies exp(-i h) directly), then
 *   split_mps_tensor_svd (src/Algorithms.jl:409-445): keep max(k, 2) where
 *   k is the first tail index at which the discarded weight reaches the
 *   threshold, at most max_bond_dim.  On a non-adjacent pair a < b the
 *   window is a..b: the window MPO carries the Kronecker factors of h on a
 *   and b and identities between (bond dimension 1), and the circuit sweep
 *   runs forward 2-site steps by +1 with the 1-site back-evolution by -1,
 *   then the edge step by +1 (src/Algorithms.jl:631-672), with identity
 *   environments at the window's outer bonds; h must be a Kronecker product. */
typedef enum {
    YAQS_GATE_EXACT = 0,
    YAQS_GATE_WINDOW_TDVP = 1
} yaqs_gate_method;

typedef struct {
    int32_t kind;
    int32_t method;          /* yaqs_gate_method; 2Q gates only, 0 otherwise */
    int64_t site0, site1;    /* 1-based; non-adjacent 2Q gates: SWAP network or long window */
    double matrix[32];
} yaqs_gate;

/* Digital TJM job: run_digital_tjm (src/DigitalTJM.jl:363-450) with the
 * local method chosen per 2-qubit gate (yaqs_gate_method).  A non-adjacent
 * 2-qubit gate with YAQS_GATE_EXACT runs as the TEBD SWAP network of
 * apply_window! (src/DigitalTJM.jl:314-330: SWAPs move the higher site next
 * to the lower one, the gate acts there, the SWAPs unwind, each step an
 * apply_local_gate_exact!), then the gate's local noise acts on its two
 * sites; with YAQS_GATE_WINDOW_TDVP it runs the TDVP window a..b. */
typedef struct {
    int64_t L;
    const yaqs_mps_tensor *mps;
    int64_t n_gates;
    const yaqs_gate *gates;
    int32_t has_noise_model;
    int32_t _pad0;
    int64_t n_noise;
    const yaqs_noise_process *noise;
    int64_t n_obs;
    const yaqs_observable *obs;
    int64_t num_traj;
    int64_t max_bond_dim;
    double truncation_threshold;
    int32_t precision;
    int32_t _pad1;
    uint64_t seed;
    uint64_t traj_offset;
} yaqs_digital_job;

typedef struct {
    double wall_s;              /* submit -> done                             */
    double queue_s;             /* submit -> first device work                */
    double flop_model;          /* algorithmic FLOP (reference contraction order) */
    int64_t matvecs;            /* H_eff applies executed (all trajectories)  */
    int64_t jumps;              /* quantum jumps applied                      */
    int64_t max_bond;           /* largest bond reached                       */
    int64_t num_samples;        /* columns of the output (times or samples)   */
    int64_t fp64_retries;       /* mixed-precision trajectories rerun in FP64
                                   after NaN/Inf before a split              */
} yaqs_job_stats;

typedef struct {
    int64_t hbm_budget_bytes;   /* 0 = 90% of free HBM at create             */
    int32_t max_batch;          /* trajectories per lockstep batch (0 = auto) */
    int32_t verbose;
} yaqs_ctx_opts;

typedef struct yaqs_ctx yaqs_ctx;

int yaqs_abi_version(void);
const char *yaqs_last_error(void);

/* Context: binds one HIP device, starts its worker threads. */
int yaqs_create(yaqs_ctx **out, int device, const yaqs_ctx_opts *opts);
void yaqs_destroy(yaqs_ctx *ctx);

/* Replaces the Threads.@threads trajectory loop of
 * Simulator._run_analog (src/Simulator.jl:37-65) running analog_tjm_2 /
 * analog_tjm_1 (src/AnalogTJM.jl:104-176).  Inputs are copied before return. */
int yaqs_submit_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, uint64_t *job_id);

/* Replaces the trajectory loop of Simulator._run_digital
 * (src/Simulator.jl:71-125) running run_digital_tjm (src/DigitalTJM.jl:363-450). */
int yaqs_submit_digital(yaqs_ctx *ctx, const yaqs_digital_job *job, uint64_t *job_id);

/* Blocks until the job is done.  traj_out: [n_obs][num_samples][num_traj]
 * (per observable a column-major (num_traj, num_samples) matrix, i.e. the
 * real parts of obs.trajectories, src/Simulator.jl:45-47); mean_out:
 * [n_obs][num_samples] (obs.results, src/SimulationConfigs.jl:168-182);
 * bond_out (digital only, may be NULL): [num_traj][num_samples] max bond
 * per sample (src/DigitalTJM.jl:385-392).  Any output pointer may be NULL. */
int yaqs_wait(yaqs_ctx *ctx, uint64_t job_id, double *traj_out, double *mean_out,
              int64_t *bond_out, yaqs_job_stats *stats);

/* Non-blocking yaqs_wait: YAQS_EAGAIN while the job runs. */
int yaqs_poll(yaqs_ctx *ctx, uint64_t job_id, double *traj_out, double *mean_out,
              int64_t *bond_out, yaqs_job_stats *stats);

/* submit + wait: what the Julia shim calls for Simulator.run(...; backend=:hip). */
int yaqs_run_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, double *traj_out,
                    double *mean_out, yaqs_job_stats *stats);
int yaqs_run_digital(yaqs_ctx *ctx, const yaqs_digital_job *job, double *traj_out,
                     double *mean_out, int64_t *bond_out, yaqs_job_stats *stats);

/* Number of output columns a job produces (times for analog, samples for
 * digital), so callers can size traj_out/mean_out. */
int64_t yaqs_analog_num_samples(const yaqs_analog_job *job);
int64_t yaqs_digital_num_samples(const yaqs_digital_job *job);

/* Projected device footprint of a job in bytes (the admission model). */
int64_t yaqs_job_footprint_bytes(const yaqs_analog_job *job);

/* Check of that model (no counterpart in the reference; for capacity
 * planning and tests): builds the engine of one wave of `batch` trajectories
 * of `job` on `device`, reads the free device memory (hipMemGetInfo) before
 * and after, releases it, and returns the model's bytes for that wave
 * (*model) and the memory the engine took (*measured).  Meaningful only
 * while nothing else allocates on the device. */
int yaqs_engine_device_bytes(int device, const yaqs_analog_job *job, int64_t batch, int64_t *model,
                             int64_t *measured);

/* Measurement entry point (no counterpart in the reference, whose timing is
 * stdout-only, experiments/benchmark_julia.jl:66-69): runs an analog job
 * synchronously, treats its first `warmup_steps` time steps as untimed
 * warm-up and reports the stream-synchronised wall time of the remaining
 * steps, plus HIP-event timing of every matrix-core GEMM launched inside
 * that window.  The waves are the ones the context would run: a job that
 * fits one wave is split into up to three concurrent waves of at least 128
 * trajectories (environment variable YAQS_WAVES, 1..8; 1 = one wave), run
 * side by side on their own streams with windows that start together (the
 * window is the longest); waves limited by the HBM budget run one after the
 * other on one stream. */
typedef struct {
    double timed_s;          /* wall time of the timed steps                      */
    int64_t timed_steps;     /* time steps inside the window, summed over waves   */
    double gemm_ms;          /* summed device time of the MFMA GEMM launches      */
    int64_t gemm_launches;
    double gemm_flop;        /* 8 FLOP per complex MAC, active trajectories only  */
    double flop_model;       /* reference-order H_eff FLOP in the window          */
    int64_t matvecs;         /* H_eff applies in the window                       */
    int64_t max_bond;
    int64_t concurrent_waves; /* waves run side by side on separate streams       */
} yaqs_bench_report;

int yaqs_bench_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, int64_t warmup_steps, double *traj_out,
                      double *mean_out, yaqs_bench_report *report);

/* measure_shots(ψ, shots) (src/MPS.jl:742-749, one single_shot_measure per
 * shot, :683-738): `shots` bitstrings sampled site by site from the MPS
 * (`L` tensors in the layout of yaqs_mps_tensor, physical dimension 2,
 * L <= 63).  bits_out[t] packs shot t's outcomes with site i (1-based) in
 * bit i-1, as single_shot_measure returns them; aggregating them into
 * counts is the caller's (measure_shots' Dict).  Shot t draws the
 * counter-RNG uniforms (seed, t, i-1) for site i (the reference draws
 * Julia's global rand()).  The MPS need not be canonical or normalised: the
 * library right-canonicalises a device copy and draws site 1 from its
 * normalised weights (the reference moves the centre from its orth_center
 * field to site 1, which presumes canonical tensors around that field, and
 * draws from the raw weights, which presumes <ψ|ψ> = 1).  Synchronous;
 * runs on the context's device. */
int yaqs_measure_shots(yaqs_ctx *ctx, int64_t L, const yaqs_mps_tensor *mps, int64_t shots, uint64_t seed,
                       int64_t *bits_out);

/* Allocation-free variant of yaqs_run_analog.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same results as yaqs_run_analog (same kernels; the launch
 * tuner only takes choices it has already cached).  It calls no hipMalloc,
 * hipFree or other allocating routine: every device buffer of the job is
 * carved from one per-device scratch array of yaqs_noalloc_scratch_bytes()
 * (256 MiB) that each GPU the library runs on holds until the process ends;
 * its device address and a stream are set up by yaqs_create.  Runs on the
 * calling thread and returns when done; trajectories run in waves that fit
 * the scratch.  A job of which one trajectory does not fit is rejected with
 * YAQS_ENOMEM_DEVICE before anything is launched.  Calls on one GPU share its
 * scratch, which keeps the last call's contents: calls from one process are
 * serialised, calls from several processes (or contexts in several
 * libraries) on one GPU race. */
int yaqs_run_analog_noalloc(yaqs_ctx *ctx, const yaqs_analog_job *job, double *traj_out, double *mean_out,
                            yaqs_job_stats *stats);
int64_t yaqs_noalloc_scratch_bytes(void);

/* Counter-based RNG of the trajectory seeding contract: draw k of global
 * trajectory t under `seed` is a pure function of (seed, t, k).  The
 * reference draws Random.default_rng() at src/StochasticProcess.jl:221-244;
 * it has no per-trajectory seeding, so this contract is the library's own. */
double yaqs_rng_uniform(uint64_t seed, uint64_t traj, uint64_t k);

#ifdef __cplusplus
}
#endif

#endif /* YAQS_HIP_H */
