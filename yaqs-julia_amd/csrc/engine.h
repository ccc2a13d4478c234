// Batched TJM engine: job description (deep copy of a C-ABI job) and the
// entry points that run all trajectories of a job on one device.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace yq {

struct ProcIn {
    int nsites;
    int s0, s1;       // 0-based sites (s1 == s0 for 1-site processes)
    double gamma;
    bool pauli_name;  // name-based Pauli test (src/Dissipation.jl:14-21)
    c64 m[16];        // reference order: column-major, 4x4 on p1 + 2 p2
    // long-range process (s0 < s1 - 1): O = alpha I + beta sig_s0 tau_s1,
    // sig/tau in reference order; pauli_name then marks the candidate whose
    // jump applies no operator ("pauli_long_range")
    bool lr = false;
    c64 sig[4], tau[4], alpha, beta;
};

struct ObsIn {
    int nsites;
    int s0, s1;
    c64 op[16];  // reference order
};

struct GateIn {
    int kind;  // 1: 1q, 2: 2q, 3: sample barrier
    int s0, s1;
    int rule = 2;  // truncation rule of the 2q split: 2 (TEBD) or 0 (window TDVP)
    bool win_long = false;  // long-range TDVP window: m holds h = coeff g_a (x) g_b
    c64 m[16];  // reference order
};

struct JobIn {
    bool digital = false;
    int64_t L = 0;
    // MPS in engine layout: site i row-major [a][s][b]
    std::vector<std::vector<c64>> mps;
    std::vector<int> mps_dl, mps_dr;
    // MPO in engine layout: site i row-major [m][so][si][n]
    std::vector<std::vector<c64>> mpo;
    std::vector<int> mpo_dl, mpo_dr;
    bool has_noise = false;
    std::vector<ProcIn> noise;
    std::vector<ObsIn> obs;
    std::vector<GateIn> gates;
    double dt = 0.1;
    int64_t num_times = 1, num_traj = 1, max_bond = 4096;
    double thr = 1e-9;
    int order = 2, sample_timesteps = 1, kdim = 25, precision = 0;
    int tdvp_sites = 2;  // 1: single_site_tdvp! in place of two_site_tdvp!
    double ktol = 1e-10;
    uint64_t seed = 0, traj_offset = 0;
    int64_t bench_warmup = -1;  // >= 0: time the steps after this many (analog)
    // Concurrent waves of one measurement share the start of the timed window.
    std::shared_ptr<struct WindowSync> wsync;
};

// Start barrier of the timed window for waves running concurrently on
// several streams: every wave syncs its stream, arrives, and all take the
// time of the last arrival as the common start.
struct WindowSync {
    explicit WindowSync(int n) : n(n) {}
    double arrive();
    void release();  // a failed wave lets the others go on
    std::mutex mu;
    std::condition_variable cv;
    int n, arrived = 0;
    double t0 = 0.0;
};

struct JobResult {
    int64_t ncol = 0;
    std::vector<double> traj;     // [obs][col][traj]
    std::vector<double> mean;     // [obs][col]
    std::vector<int64_t> bonds;   // [traj][col] (digital)
    double flop = 0;
    int64_t matvecs = 0, jumps = 0, max_bond = 0;
    // job-level indices of trajectories whose two-site tensor held NaN/Inf
    // before a split (their results are not valid)
    std::vector<int64_t> failed;
    int64_t retried = 0;  // mixed-precision trajectories rerun in FP64
    // measurement window (bench_warmup >= 0)
    double timed_s = 0, gemm_ms = 0, gemm_flop = 0, win_flop = 0;
    int64_t timed_steps = 0, gemm_launches = 0, win_matvecs = 0;
};

// Host objects one stream reuses across the engines it runs: the coherent
// pinned words the Lanczos finalize kernel stores its active counts into
// (and their device address) and two events.  A worker creates them once,
// so waves neither allocate pinned memory nor create events.
struct StreamRes {
    int *h_pinned = nullptr;
    int *d_pinned = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    void create();   // hipHostMalloc + hipEventCreate (throws HipError)
    void destroy();  // after the stream that used them is synchronised
};

// Device memory of engines built on this thread comes from `base` (bump
// allocation, 256-byte aligned) instead of hipMalloc while an ArenaScope is
// alive; running out throws ArenaFull before anything is launched into it.
struct ArenaFull : std::runtime_error {
    explicit ArenaFull(const std::string &s) : std::runtime_error(s) {}
};
class ArenaScope {
  public:
    ArenaScope(void *base, size_t bytes);
    ~ArenaScope();
    size_t used() const;
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;
};

// Throws std::invalid_argument for malformed jobs, HipError for device errors.
void validate_job(const JobIn &job);
int64_t num_columns(const JobIn &job);
int64_t footprint_bytes(const JobIn &job, int64_t batch);
// Runs every trajectory of `job` in waves of at most `batch` trajectories.
void run_job(const JobIn &job, int64_t batch, hipStream_t stream, JobResult &out, std::atomic<bool> *cancel = nullptr);

// One engine of `batch` trajectory slots on `stream`; run() executes the
// trajectories [t0, t0 + nb) of the job, writing their observables into the
// job's device output d_out ([obs][col][num_traj], FP64) and, for digital
// jobs, their bond dimensions into h_bonds ([num_traj][col], host).  Waves of
// one job on different engines write disjoint slices.
class WaveEngine {
  public:
    virtual ~WaveEngine() = default;
    virtual void run(int64_t t0, int nb, double *d_out, int64_t *h_bonds, JobResult &res) = 0;
};
// res: the stream's reusable host objects, or nullptr (the engine creates
// and destroys its own).
std::unique_ptr<WaveEngine> make_engine(const JobIn &job, int batch, hipStream_t stream, StreamRes *res = nullptr);

double rng_uniform_host(uint64_t seed, uint64_t traj, uint64_t k);

}  // namespace yq
