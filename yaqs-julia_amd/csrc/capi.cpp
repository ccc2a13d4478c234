// C-ABI of the TJM engine (include/yaqs_hip.h): contexts, the job queue with
// admission by projected HBM footprint, and worker threads that run admitted
// jobs concurrently on their own HIP streams.
#include "../../include/yaqs_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "noalloc.h"
#include "shots.h"
#include "tune.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

yq::c64 mkc(double r, double i) {
    yq::c64 z;
    z.re = r;
    z.im = i;
    return z;
}

// exp(-i h) of a 4x4 matrix in place (column-major), by scaling and squaring
// of the Taylor series: h / 2^s has 1-norm <= 1/2, the series runs until the
// term's 1-norm falls below 1e-18 of the sum, then s squarings.
void expm_minus_i4(yq::c64 *m) {
    using cd = std::complex<double>;
    cd A[16];
    double nrm = 0;
    for (int c = 0; c < 4; ++c) {
        double col = 0;
        for (int r = 0; r < 4; ++r) col += std::abs(cd(m[r + 4 * c].re, m[r + 4 * c].im));
        nrm = std::max(nrm, col);
    }
    if (!std::isfinite(nrm)) throw std::invalid_argument("window Hamiltonian is not finite");
    int sq = 0;
    while (nrm > 0.5 && sq < 60) { nrm *= 0.5; ++sq; }
    const double f = std::ldexp(1.0, -sq);
    for (int k = 0; k < 16; ++k) A[k] = cd(0, -1) * cd(m[k].re, m[k].im) * f;
    cd E[16], T[16], U[16];
    for (int k = 0; k < 16; ++k) E[k] = T[k] = (k % 5 == 0) ? cd(1, 0) : cd(0, 0);
    auto mul = [](const cd *X, const cd *Y, cd *Z) {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) {
                cd acc = 0;
                for (int k = 0; k < 4; ++k) acc += X[r + 4 * k] * Y[k + 4 * c];
                Z[r + 4 * c] = acc;
            }
    };
    for (int n = 1; n <= 40; ++n) {
        mul(T, A, U);
        double tn = 0;
        for (int k = 0; k < 16; ++k) { T[k] = U[k] / double(n); E[k] += T[k]; tn += std::abs(T[k]); }
        if (tn < 1e-18) break;
    }
    for (int q = 0; q < sq; ++q) {
        mul(E, E, U);
        for (int k = 0; k < 16; ++k) E[k] = U[k];
    }
    for (int k = 0; k < 16; ++k) m[k] = mkc(E[k].real(), E[k].imag());
}

// Julia column-major (Dl, d, Dr) -> engine row-major [a][s][b]
void load_mps(yq::JobIn &J, int64_t L, const yaqs_mps_tensor *mps) {
    if (!mps) throw std::invalid_argument("mps is NULL");
    J.mps.resize(L);
    J.mps_dl.resize(L);
    J.mps_dr.resize(L);
    for (int64_t i = 0; i < L; ++i) {
        const yaqs_mps_tensor &t = mps[i];
        if (t.dl < 1 || t.dr < 1 || t.d != 2 || !t.data) throw std::invalid_argument("bad MPS tensor (physical dim must be 2)");
        J.mps_dl[i] = (int)t.dl;
        J.mps_dr[i] = (int)t.dr;
        J.mps[i].resize((size_t)t.dl * 2 * t.dr);
        for (int64_t a = 0; a < t.dl; ++a)
            for (int64_t s = 0; s < 2; ++s)
                for (int64_t b = 0; b < t.dr; ++b) {
                    const size_t src = (size_t)(a + t.dl * (s + 2 * b));
                    J.mps[i][(size_t)(a * 2 + s) * t.dr + b] = mkc(t.data[2 * src], t.data[2 * src + 1]);
                }
    }
}

void load_noise(yq::JobIn &J, int32_t has, int64_t n, const yaqs_noise_process *p, int64_t L) {
    J.has_noise = has != 0;
    if (!has) return;
    if (n > 0 && !p) throw std::invalid_argument("noise is NULL");
    for (int64_t i = 0; i < n; ++i) {
        yq::ProcIn q;
        q.nsites = p[i].nsites;
        if (q.nsites != 1 && q.nsites != 2) throw std::invalid_argument("noise process must act on 1 or 2 sites");
        q.s0 = (int)p[i].site0 - 1;
        q.s1 = q.nsites == 2 ? (int)p[i].site1 - 1 : q.s0;
        if (q.s0 < 0 || q.s0 >= L || q.s1 < 0 || q.s1 >= L) throw std::invalid_argument("noise site out of range");
        q.gamma = p[i].strength;
        q.pauli_name = p[i].pauli_name != 0;
        const int d = q.nsites == 1 ? 2 : 4;
        for (int k = 0; k < 16; ++k) q.m[k] = mkc(0, 0);
        if (p[i].kind == YAQS_NOISE_LONG_RANGE) {
            if (q.nsites != 2 || q.s1 < q.s0 + 2)
                throw std::invalid_argument("a long-range noise process needs sites site0 < site1 - 1");
            q.lr = true;
            for (int k = 0; k < 4; ++k) {
                q.sig[k] = mkc(p[i].matrix[2 * k], p[i].matrix[2 * k + 1]);
                q.tau[k] = mkc(p[i].matrix[8 + 2 * k], p[i].matrix[8 + 2 * k + 1]);
            }
            q.alpha = mkc(p[i].matrix[16], p[i].matrix[17]);
            q.beta = mkc(p[i].matrix[18], p[i].matrix[19]);
            J.noise.push_back(q);
            continue;
        }
        for (int k = 0; k < d * d; ++k) q.m[k] = mkc(p[i].matrix[2 * k], p[i].matrix[2 * k + 1]);
        J.noise.push_back(q);
    }
}

void load_obs(yq::JobIn &J, int64_t n, const yaqs_observable *o) {
    if (n > 0 && !o) throw std::invalid_argument("obs is NULL");
    for (int64_t i = 0; i < n; ++i) {
        yq::ObsIn q;
        q.nsites = o[i].nsites;
        if (q.nsites != 1 && q.nsites != 2) throw std::invalid_argument("observables act on 1 or 2 sites");
        q.s0 = (int)o[i].site0 - 1;
        q.s1 = q.nsites == 2 ? (int)o[i].site1 - 1 : q.s0;
        const int d = q.nsites == 1 ? 2 : 4;
        for (int k = 0; k < 16; ++k) q.op[k] = mkc(0, 0);
        for (int k = 0; k < d * d; ++k) q.op[k] = mkc(o[i].op[2 * k], o[i].op[2 * k + 1]);
        J.obs.push_back(q);
    }
}

yq::JobIn convert_analog(const yaqs_analog_job *job) {
    if (!job) throw std::invalid_argument("job is NULL");
    yq::JobIn J;
    J.digital = false;
    J.L = job->L;
    if (J.L < 1) throw std::invalid_argument("L must be >= 1");
    load_mps(J, J.L, job->mps);
    if (!job->mpo) throw std::invalid_argument("mpo is NULL");
    J.mpo.resize(J.L);
    J.mpo_dl.resize(J.L);
    J.mpo_dr.resize(J.L);
    for (int64_t i = 0; i < J.L; ++i) {
        const yaqs_mpo_tensor &w = job->mpo[i];
        if (w.dout != 2 || w.din != 2 || w.dl < 1 || w.dr < 1 || !w.data) throw std::invalid_argument("bad MPO tensor");
        J.mpo_dl[i] = (int)w.dl;
        J.mpo_dr[i] = (int)w.dr;
        J.mpo[i].resize((size_t)w.dl * 4 * w.dr);
        for (int64_t m = 0; m < w.dl; ++m)
            for (int64_t so = 0; so < 2; ++so)
                for (int64_t si = 0; si < 2; ++si)
                    for (int64_t n = 0; n < w.dr; ++n) {
                        const size_t src = (size_t)(m + w.dl * (so + 2 * (si + 2 * n)));
                        J.mpo[i][(size_t)((m * 2 + so) * 2 + si) * w.dr + n] = mkc(w.data[2 * src], w.data[2 * src + 1]);
                    }
    }
    load_noise(J, job->has_noise_model, job->n_noise, job->noise, J.L);
    load_obs(J, job->n_obs, job->obs);
    J.dt = job->dt;
    J.num_times = job->num_times;
    J.num_traj = job->num_traj;
    J.max_bond = job->max_bond_dim;
    J.thr = job->truncation_threshold;
    J.order = job->order;
    J.tdvp_sites = job->tdvp_sites == 1 ? 1 : 2;
    if (job->tdvp_sites != 0 && job->tdvp_sites != 1 && job->tdvp_sites != 2)
        throw std::invalid_argument("tdvp_sites must be 0, 1 or 2");
    J.sample_timesteps = job->sample_timesteps;
    J.kdim = job->krylov_dim > 0 ? job->krylov_dim : 25;
    J.ktol = job->krylov_tol > 0 ? job->krylov_tol : 1e-10;
    J.precision = job->precision;
    J.seed = job->seed;
    J.traj_offset = job->traj_offset;
    if (J.precision != YAQS_PREC_FP64 && J.precision != YAQS_PREC_MIXED) throw std::invalid_argument("bad precision");
    return J;
}

yq::JobIn convert_digital(const yaqs_digital_job *job) {
    if (!job) throw std::invalid_argument("job is NULL");
    yq::JobIn J;
    J.digital = true;
    J.L = job->L;
    if (J.L < 1) throw std::invalid_argument("L must be >= 1");
    load_mps(J, J.L, job->mps);
    if (job->n_gates > 0 && !job->gates) throw std::invalid_argument("gates is NULL");
    for (int64_t g = 0; g < job->n_gates; ++g) {
        const yaqs_gate &G = job->gates[g];
        yq::GateIn q;
        q.kind = G.kind;
        if (q.kind < 1 || q.kind > 3) throw std::invalid_argument("bad gate kind");
        q.s0 = (int)G.site0 - 1;
        q.s1 = q.kind == 2 ? (int)G.site1 - 1 : q.s0;
        const int d = q.kind == 1 ? 2 : 4;
        for (int k = 0; k < 16; ++k) q.m[k] = mkc(0, 0);
        if (q.kind != 3)
            for (int k = 0; k < d * d; ++k) q.m[k] = mkc(G.matrix[2 * k], G.matrix[2 * k + 1]);
        if (G.method != YAQS_GATE_EXACT) {
            if (G.method != YAQS_GATE_WINDOW_TDVP || q.kind != 2)
                throw std::invalid_argument("bad gate method (window TDVP applies to 2-qubit gates only)");
            q.rule = 0;
            if (std::abs(q.s1 - q.s0) == 1) expm_minus_i4(q.m);  // the edge step alone
            else q.win_long = true;                               // the circuit sweep of the window
        }
        J.gates.push_back(q);
    }
    load_noise(J, job->has_noise_model, job->n_noise, job->noise, J.L);
    load_obs(J, job->n_obs, job->obs);
    J.num_traj = job->num_traj;
    J.max_bond = job->max_bond_dim;
    J.thr = job->truncation_threshold;
    J.precision = job->precision;
    J.seed = job->seed;
    J.traj_offset = job->traj_offset;
    if (J.precision != YAQS_PREC_FP64 && J.precision != YAQS_PREC_MIXED) throw std::invalid_argument("bad precision");
    return J;
}

struct Job {
    uint64_t id;
    yq::JobIn in;
    yq::JobResult out;
    int status = -1;  // -1 queued/running, else yaqs_status
    std::string err;
    double t_submit = 0, t_start = 0, t_end = 0;
    // wave schedule, fixed when the job is admitted
    int64_t batch = 0;  // trajectories per wave (0: not admitted yet)
    int64_t nwaves = 0, next_wave = 0, done_waves = 0, running = 0;
    double *d_out = nullptr;  // [obs][col][traj] on the device, shared by the waves
    bool failed = false;
    int status_code_pending = YAQS_OK;
};

}  // namespace

struct yaqs_ctx {
    int device = 0;
    int64_t budget = 0;   // HBM bytes the context may use for jobs
    int64_t in_use = 0;
    int max_batch = 0;
    int verbose = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;  // jobs not finished yet, in submission order
    std::map<uint64_t, std::shared_ptr<Job>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> workers;
    // allocation-free entry point: device address of the device's scratch
    // array, and a stream with its pinned words and events, all set up by
    // yaqs_create
    void *na_scratch = nullptr;
    hipStream_t na_stream = nullptr;
    yq::StreamRes na_res;
};

namespace {

// Concurrent waves for a job whose trajectories fit one wave: at least 128
// trajectories a wave, YAQS_WAVES (1..8) of them, three by default.  The
// latency-bound phases of one wave (Lanczos tails, host round trips, the
// eigensolver) overlap the matrix-core phases of the others (C2, 1000
// trajectories: 14.7k traj-site-steps/s as three waves, 14.5k as two, 14.3k
// as one or four; the digital brickwork, 512 trajectories: 49.8k
// traj-site-layers/s as three waves, 47.3k as two, 42.6k as one).
int concurrent_waves(int64_t ntraj) {
    static const int k = [] {
        const char *e = std::getenv("YAQS_WAVES");
        return e ? std::max(1, std::min(8, atoi(e))) : 3;
    }();
    return (int)std::max<int64_t>(1, std::min<int64_t>(k, ntraj / 128));
}

// Batch size and bytes for a job under the remaining budget; 0 if it cannot
// start now.  conc: the waves meant to run side by side (1 unless the job
// fits one wave and is split by concurrent_waves).
int64_t choose_batch(yaqs_ctx *c, const yq::JobIn &in, int64_t avail, int64_t *bytes, int *conc = nullptr) {
    const int64_t per = std::max<int64_t>(1, yq::footprint_bytes(in, 1));
    int64_t b = std::min<int64_t>(in.num_traj, c->max_batch > 0 ? c->max_batch : 4096);
    b = std::min<int64_t>(b, avail / per);
    if (b < 1) return 0;
    // balance waves: same number of waves, fewer padded slots
    int64_t waves = (in.num_traj + b - 1) / b;
    int k = 1;
    if (waves == 1 && c->max_batch == 0) {
        k = std::min<int>(concurrent_waves(in.num_traj), std::max<int>(1, (int)c->workers.size()));
        waves = k;
    }
    b = (in.num_traj + waves - 1) / waves;
    if (conc) *conc = k;
    *bytes = yq::footprint_bytes(in, b);
    return b;
}

int64_t output_elems(const yq::JobIn &in) {
    return std::max<int64_t>(1, (int64_t)in.obs.size() * yq::num_columns(in) * in.num_traj);
}

void finish_locked(yaqs_ctx *c, const std::shared_ptr<Job> &job, int status, const std::string &err) {
    job->status = status;
    job->err = err;
    job->t_end = now_s();
    for (auto it = c->queue.begin(); it != c->queue.end(); ++it)
        if (*it == job) {
            c->queue.erase(it);
            break;
        }
    c->cv_done.notify_all();
}

struct Wave {
    std::shared_ptr<Job> job;
    int64_t t0 = 0;
    int nb = 0;
    int64_t bytes = 0;
};

// Continuous batching (called with the lock held).  Among admitted jobs with
// pending waves whose footprint fits, the one with the fewest running waves
// (earliest submitted among equals) gets the worker, so a job that arrives
// behind a long one starts on the next free worker; the oldest job not yet
// admitted competes as a job with no running wave, and is admitted when its
// first wave fits (admission stays in submission order).
bool pick_wave(yaqs_ctx *c, Wave &w) {
    std::shared_ptr<Job> best;
    int64_t best_nb = 0, best_bytes = 0;
    const int64_t avail = c->budget - c->in_use;
    for (size_t qi = 0; qi < c->queue.size();) {
        std::shared_ptr<Job> j = c->queue[qi];
        if (j->failed) {
            ++qi;
            continue;
        }
        if (j->batch == 0) {
            // oldest job not admitted yet
            if (best && best->running == 0) break;
            int64_t bytes = 0;
            const int64_t b = choose_batch(c, j->in, avail, &bytes);
            if (b < 1) {
                if (c->in_use == 0) {  // cannot fit even alone
                    finish_locked(c, j, YAQS_ENOMEM_DEVICE, "job footprint exceeds the HBM budget");
                    continue;  // the queue shifted: same qi
                }
                break;
            }
            try {
                const int64_t n = output_elems(j->in);
                if (hipMalloc(&j->d_out, n * sizeof(double)) != hipSuccess) throw std::bad_alloc();
                // the workers' streams do not synchronise with the null stream:
                // the zeroing completes before any wave can write the output
                if (hipMemset(j->d_out, 0, n * sizeof(double)) != hipSuccess ||
                    hipStreamSynchronize(nullptr) != hipSuccess)
                    throw std::bad_alloc();
                j->out.ncol = yq::num_columns(j->in);
                j->out.traj.assign((size_t)n, 0.0);
                j->out.mean.assign((size_t)std::max<int64_t>(1, (int64_t)j->in.obs.size() * j->out.ncol), 0.0);
                if (j->in.digital) j->out.bonds.assign((size_t)j->in.num_traj * j->out.ncol, 0);
            } catch (const std::exception &) {
                if (j->d_out) (void)hipFree(j->d_out);
                j->d_out = nullptr;
                finish_locked(c, j, YAQS_ENOMEM_DEVICE, "cannot allocate the job output on the device");
                continue;
            }
            j->batch = b;
            j->nwaves = (j->in.num_traj + b - 1) / b;
            j->t_start = now_s();
            best = j;
            best_nb = (int)std::min<int64_t>(b, j->in.num_traj);
            best_bytes = bytes;
            break;
        }
        if (j->next_wave < j->nwaves && (!best || j->running < best->running)) {
            const int64_t t0 = j->next_wave * j->batch;
            const int64_t nb = std::min<int64_t>(j->batch, j->in.num_traj - t0);
            const int64_t bytes = yq::footprint_bytes(j->in, nb);
            if (bytes <= avail) {
                best = j;
                best_nb = nb;
                best_bytes = bytes;
            }
        }
        ++qi;
    }
    if (!best) return false;
    w.job = best;
    w.t0 = best->next_wave * best->batch;
    w.nb = (int)best_nb;
    w.bytes = best_bytes;
    best->next_wave++;
    best->running++;
    c->in_use += best_bytes;
    return true;
}

void worker_main(yaqs_ctx *c) {
    if (hipSetDevice(c->device) != hipSuccess) return;
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return;
    // the worker's pinned words and events, reused by every wave it runs
    yq::StreamRes sres;
    try {
        sres.create();
    } catch (const std::exception &) {
        (void)hipStreamDestroy(stream);
        return;
    }
    for (;;) {
        Wave w;
        {
            std::unique_lock<std::mutex> lk(c->mu);
            for (;;) {
                if (c->stop) break;
                if (pick_wave(c, w)) break;
                c->cv_work.wait(lk);
            }
        }
        if (!w.job) break;
        Job &job = *w.job;
        int st = YAQS_OK;
        std::string err;
        yq::JobResult wr;
        try {
            auto eng = yq::make_engine(job.in, w.nb, stream, &sres);
            eng->run(w.t0, w.nb, job.d_out, job.in.digital ? job.out.bonds.data() : nullptr, wr);
            YQ_HIP(hipStreamSynchronize(stream));
        } catch (const yq::HipError &e) {
            st = YAQS_EDEVICE;
            err = e.what();
        } catch (const std::invalid_argument &e) {
            st = YAQS_EINVAL;
            err = e.what();
        } catch (const std::bad_alloc &e) {
            st = YAQS_ENOMEM_DEVICE;
            err = e.what();
        } catch (const std::exception &e) {
            st = YAQS_EDEVICE;
            err = e.what();
        }
        bool last = false;
        {
            std::lock_guard<std::mutex> lk(c->mu);
            c->in_use -= w.bytes;
            job.running--;
            job.done_waves++;
            job.out.flop += wr.flop;
            job.out.matvecs += wr.matvecs;
            job.out.jumps += wr.jumps;
            job.out.max_bond = std::max(job.out.max_bond, wr.max_bond);
            job.out.failed.insert(job.out.failed.end(), wr.failed.begin(), wr.failed.end());
            if (st != YAQS_OK && !job.failed) {
                job.failed = true;
                job.status_code_pending = st;
                job.err = err;
            }
            last = job.running == 0 && (job.failed || job.done_waves == job.nwaves);
        }
        if (last) {
            // only this worker can reach here for the job: no wave is running or pending
            int fst = job.failed ? job.status_code_pending : YAQS_OK;
            std::string ferr = job.err;
            if (!job.failed && !job.out.failed.empty()) {
                // NaN/Inf before a split (src/Algorithms.jl:414-417): a
                // mixed-precision trajectory is run again in FP64 on its
                // own; what still fails (or failed in FP64) ends the job
                // with YAQS_ENUMERIC
                std::vector<int64_t> bad = job.out.failed;
                std::sort(bad.begin(), bad.end());
                bad.erase(std::unique(bad.begin(), bad.end()), bad.end());
                const bool mixed = job.in.precision == YAQS_PREC_MIXED;
                std::vector<int64_t> still;
                if (mixed) {
                    yq::JobIn f = job.in;
                    f.precision = YAQS_PREC_FP64;
                    f.wsync = nullptr;
                    f.bench_warmup = -1;
                    try {
                        for (int64_t t : bad) {
                            yq::JobResult rr;
                            auto eng = yq::make_engine(f, 1, stream, &sres);
                            eng->run(t, 1, job.d_out, job.in.digital ? job.out.bonds.data() : nullptr, rr);
                            YQ_HIP(hipStreamSynchronize(stream));
                            if (!rr.failed.empty()) still.push_back(t);
                            job.out.jumps += rr.jumps;
                            job.out.matvecs += rr.matvecs;
                            job.out.flop += rr.flop;
                        }
                    } catch (const std::exception &e) {
                        fst = YAQS_EDEVICE;
                        ferr = std::string("FP64 retry of a trajectory with NaN/Inf failed: ") + e.what();
                    }
                } else {
                    still = bad;
                }
                job.out.retried = mixed ? (int64_t)bad.size() : 0;
                if (fst == YAQS_OK && !still.empty()) {
                    fst = YAQS_ENUMERIC;
                    ferr = "NaN/Inf in the two-site tensor before a split in " + std::to_string(still.size()) +
                           " trajectories (first: " + std::to_string(still.front()) + ")" +
                           (mixed ? ", also in the FP64 retry" : "");
                }
            }
            if (!job.failed && fst == YAQS_OK) {
                if (hipMemcpy(job.out.traj.data(), job.d_out, job.out.traj.size() * sizeof(double),
                              hipMemcpyDeviceToHost) != hipSuccess) {
                    fst = YAQS_EDEVICE;
                    ferr = "copy of the job output failed";
                } else {
                    const int64_t nobs = (int64_t)job.in.obs.size(), ncol = job.out.ncol, ntraj = job.in.num_traj;
                    for (int64_t q = 0; q < nobs; ++q)
                        for (int64_t col = 0; col < ncol; ++col) {
                            double sum = 0;
                            for (int64_t t = 0; t < ntraj; ++t) sum += job.out.traj[((size_t)q * ncol + col) * ntraj + t];
                            job.out.mean[(size_t)q * ncol + col] = sum / (double)ntraj;
                        }
                }
            }
            (void)hipFree(job.d_out);
            job.d_out = nullptr;
            std::lock_guard<std::mutex> lk(c->mu);
            finish_locked(c, w.job, fst, ferr);
        }
        c->cv_work.notify_all();
    }
    (void)hipStreamSynchronize(stream);
    sres.destroy();
    (void)hipStreamDestroy(stream);
}

int submit(yaqs_ctx *ctx, yq::JobIn &&in, uint64_t *job_id) {
    yq::validate_job(in);
    auto job = std::make_shared<Job>();
    job->in = std::move(in);
    job->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        job->id = ctx->next_id++;
        ctx->jobs[job->id] = job;
        ctx->queue.push_back(job);
    }
    ctx->cv_work.notify_all();
    if (job_id) *job_id = job->id;
    return YAQS_OK;
}

int collect(yaqs_ctx *ctx, uint64_t id, double *traj_out, double *mean_out, int64_t *bond_out, yaqs_job_stats *stats,
            bool block) {
    std::shared_ptr<Job> job;
    {
        std::unique_lock<std::mutex> lk(ctx->mu);
        auto it = ctx->jobs.find(id);
        if (it == ctx->jobs.end()) return fail(YAQS_ENOTFOUND, "unknown job id");
        job = it->second;
        if (job->status < 0) {
            if (!block) return YAQS_EAGAIN;
            ctx->cv_done.wait(lk, [&] { return job->status >= 0; });
        }
        ctx->jobs.erase(id);
    }
    if (job->status != YAQS_OK) return fail(job->status, job->err);
    const yq::JobResult &r = job->out;
    const int64_t nobs = (int64_t)job->in.obs.size();
    if (traj_out && nobs) std::copy(r.traj.begin(), r.traj.begin() + nobs * r.ncol * job->in.num_traj, traj_out);
    if (mean_out && nobs) std::copy(r.mean.begin(), r.mean.begin() + nobs * r.ncol, mean_out);
    if (bond_out && !r.bonds.empty()) std::copy(r.bonds.begin(), r.bonds.end(), bond_out);
    if (stats) {
        stats->wall_s = job->t_end - job->t_submit;
        stats->queue_s = job->t_start - job->t_submit;
        stats->flop_model = r.flop;
        stats->matvecs = r.matvecs;
        stats->jumps = r.jumps;
        stats->max_bond = r.max_bond;
        stats->num_samples = r.ncol;
        stats->fp64_retries = r.retried;
    }
    return YAQS_OK;
}

}  // namespace

extern "C" {

int yaqs_abi_version(void) { return YAQS_ABI_VERSION; }

const char *yaqs_last_error(void) { return g_last_error.c_str(); }

int yaqs_create(yaqs_ctx **out, int device, const yaqs_ctx_opts *opts) {
    if (!out) return fail(YAQS_EINVAL, "out is NULL");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(YAQS_EDEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(YAQS_EINVAL, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(YAQS_EDEVICE, "hipSetDevice failed");
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return fail(YAQS_EDEVICE, "hipMemGetInfo failed");
    auto *c = new yaqs_ctx();
    c->device = device;
    if (yq::noalloc_init(device) != 0) {
        delete c;
        return fail(YAQS_EDEVICE, "cannot look up the device scratch of the allocation-free entry point");
    }
    c->na_scratch = yq::noalloc_scratch(device);
    if (hipStreamCreateWithFlags(&c->na_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(YAQS_EDEVICE, "hipStreamCreate failed");
    }
    try {
        c->na_res.create();
    } catch (const std::exception &e) {
        (void)hipStreamDestroy(c->na_stream);
        delete c;
        return fail(YAQS_EDEVICE, e.what());
    }
    c->budget = (opts && opts->hbm_budget_bytes > 0) ? opts->hbm_budget_bytes : (int64_t)(0.9 * (double)freeb);
    c->max_batch = opts ? opts->max_batch : 0;
    c->verbose = opts ? opts->verbose : 0;
    // worker threads, each with its own stream (YAQS_WORKERS, 1..8): four,
    // one per hardware queue of the process, measured best on the serving
    // stream (p50 0.73 s / p99 2.0 s against 1.05 s / 4.9 s with two; more
    // streams than queues were slower)
    int nworkers = 4;
    if (const char *e = getenv("YAQS_WORKERS")) nworkers = std::max(1, std::min(8, atoi(e)));
    for (int i = 0; i < nworkers; ++i) c->workers.emplace_back(worker_main, c);
    *out = c;
    return YAQS_OK;
}

void yaqs_destroy(yaqs_ctx *ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->stop = true;
    }
    ctx->cv_work.notify_all();
    for (auto &t : ctx->workers) t.join();
    if (ctx->na_stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->na_stream);
        ctx->na_res.destroy();
        (void)hipStreamDestroy(ctx->na_stream);
    }
    delete ctx;
}

int yaqs_submit_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, uint64_t *job_id) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    try {
        return submit(ctx, convert_analog(job), job_id);
    } catch (const std::exception &e) {
        return fail(YAQS_EINVAL, e.what());
    }
}

int yaqs_submit_digital(yaqs_ctx *ctx, const yaqs_digital_job *job, uint64_t *job_id) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    try {
        return submit(ctx, convert_digital(job), job_id);
    } catch (const std::exception &e) {
        return fail(YAQS_EINVAL, e.what());
    }
}

int yaqs_wait(yaqs_ctx *ctx, uint64_t job_id, double *traj_out, double *mean_out, int64_t *bond_out,
              yaqs_job_stats *stats) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    return collect(ctx, job_id, traj_out, mean_out, bond_out, stats, true);
}

int yaqs_poll(yaqs_ctx *ctx, uint64_t job_id, double *traj_out, double *mean_out, int64_t *bond_out,
              yaqs_job_stats *stats) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    return collect(ctx, job_id, traj_out, mean_out, bond_out, stats, false);
}

int yaqs_run_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, double *traj_out, double *mean_out,
                    yaqs_job_stats *stats) {
    uint64_t id = 0;
    int rc = yaqs_submit_analog(ctx, job, &id);
    if (rc != YAQS_OK) return rc;
    return yaqs_wait(ctx, id, traj_out, mean_out, nullptr, stats);
}

int yaqs_run_digital(yaqs_ctx *ctx, const yaqs_digital_job *job, double *traj_out, double *mean_out,
                     int64_t *bond_out, yaqs_job_stats *stats) {
    uint64_t id = 0;
    int rc = yaqs_submit_digital(ctx, job, &id);
    if (rc != YAQS_OK) return rc;
    return yaqs_wait(ctx, id, traj_out, mean_out, bond_out, stats);
}

int yaqs_measure_shots(yaqs_ctx *ctx, int64_t L, const yaqs_mps_tensor *mps, int64_t shots, uint64_t seed,
                       int64_t *bits_out) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    if (shots > 0 && !bits_out) return fail(YAQS_EINVAL, "bits_out is NULL");
    hipStream_t s = nullptr;
    try {
        if (L < 1 || L > 63) throw std::invalid_argument("L must be 1..63");
        yq::JobIn J;
        load_mps(J, L, mps);
        YQ_HIP(hipSetDevice(ctx->device));
        YQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        yq::measure_shots((int)L, J.mps, J.mps_dl, J.mps_dr, shots, seed, bits_out, s);
        YQ_HIP(hipStreamDestroy(s));
        return YAQS_OK;
    } catch (const std::invalid_argument &e) {
        if (s) (void)hipStreamDestroy(s);
        return fail(YAQS_EINVAL, e.what());
    } catch (const yq::HipError &e) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        return fail(YAQS_EDEVICE, e.what());
    } catch (const std::exception &e) {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        return fail(YAQS_EDEVICE, e.what());
    }
}

int64_t yaqs_analog_num_samples(const yaqs_analog_job *job) {
    if (!job) return -1;
    return job->sample_timesteps ? job->num_times : 1;
}

int64_t yaqs_digital_num_samples(const yaqs_digital_job *job) {
    try {
        return yq::num_columns(convert_digital(job));
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

int64_t yaqs_job_footprint_bytes(const yaqs_analog_job *job) {
    try {
        yq::JobIn in = convert_analog(job);
        return yq::footprint_bytes(in, std::max<int64_t>(1, in.num_traj));
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

int yaqs_engine_device_bytes(int device, const yaqs_analog_job *job, int64_t batch, int64_t *model,
                             int64_t *measured) {
    if (!model || !measured) return fail(YAQS_EINVAL, "model / measured is NULL");
    hipStream_t st = nullptr;
    try {
        yq::JobIn in = convert_analog(job);
        yq::validate_job(in);
        batch = std::max<int64_t>(1, batch);
        if (hipSetDevice(device) != hipSuccess) return fail(YAQS_EDEVICE, "hipSetDevice failed");
        YQ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        size_t f0 = 0, f1 = 0, tot = 0;
        YQ_HIP(hipDeviceSynchronize());
        YQ_HIP(hipMemGetInfo(&f0, &tot));
        {
            std::unique_ptr<yq::WaveEngine> eng = yq::make_engine(in, (int)batch, st, nullptr);
            YQ_HIP(hipDeviceSynchronize());
            YQ_HIP(hipMemGetInfo(&f1, &tot));
        }
        YQ_HIP(hipStreamSynchronize(st));
        YQ_HIP(hipStreamDestroy(st));
        st = nullptr;
        *model = yq::footprint_bytes(in, batch);
        *measured = (int64_t)f0 - (int64_t)f1;
        return YAQS_OK;
    } catch (const std::bad_alloc &e) {
        if (st) (void)hipStreamDestroy(st);
        return fail(YAQS_ENOMEM_DEVICE, e.what());
    } catch (const std::invalid_argument &e) {
        if (st) (void)hipStreamDestroy(st);
        return fail(YAQS_EINVAL, e.what());
    } catch (const std::exception &e) {
        if (st) (void)hipStreamDestroy(st);
        return fail(YAQS_EDEVICE, e.what());
    }
}

double yaqs_rng_uniform(uint64_t seed, uint64_t traj, uint64_t k) { return yq::rng_uniform_host(seed, traj, k); }

int64_t yaqs_noalloc_scratch_bytes(void) { return (int64_t)yq::kNoAllocScratchBytes; }

int yaqs_run_analog_noalloc(yaqs_ctx *ctx, const yaqs_analog_job *job, double *traj_out, double *mean_out,
                            yaqs_job_stats *stats) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    if (!ctx->na_scratch) return fail(YAQS_EDEVICE, "the device scratch was not set up");
    const double t_submit = now_s();
    yq::JobIn in;
    try {
        in = convert_analog(job);
        yq::validate_job(in);
    } catch (const std::exception &e) {
        return fail(YAQS_EINVAL, e.what());
    }
    const int64_t ntraj = in.num_traj, nobs = (int64_t)in.obs.size(), ncol = yq::num_columns(in);
    const int64_t nout = std::max<int64_t>(2, nobs * ncol * ntraj);
    const int64_t out_bytes = ((nout * (int64_t)sizeof(double)) + 255) & ~(int64_t)255;
    const int64_t cap = (int64_t)yq::kNoAllocScratchBytes;
    // Slots per wave from the footprint model, with room for the buffers'
    // 256-byte alignment and the job's own tables (MPS/MPO uploads, sparse
    // MPO steps); the arena itself stops an underestimate with an error
    // before anything is launched into memory it does not hold.
    int64_t tables = 0;
    for (const auto &m : in.mps) tables += (int64_t)m.size() * (int64_t)sizeof(yq::c64) + 256;
    for (const auto &w : in.mpo) tables += 6 * ((int64_t)w.size() * (int64_t)sizeof(yq::c64) + 3 * 256);
    const int64_t slack = ((int64_t)4 << 20) + tables + (12 * in.L + 256) * 256;
    const int64_t per = yq::footprint_bytes(in, 1);
    const int64_t room = cap - out_bytes - slack;
    const int64_t B = room > 0 ? std::min<int64_t>(ntraj, room / std::max<int64_t>(1, per)) : 0;
    if (B < 1)
        return fail(YAQS_ENOMEM_DEVICE,
                    "job exceeds the allocation-free entry point's scratch: one trajectory needs " +
                        std::to_string(per + out_bytes + slack) + " bytes, the scratch holds " + std::to_string(cap) +
                        " bytes (use yaqs_run_analog)");
    std::lock_guard<std::mutex> lk(yq::noalloc_lock(ctx->device));
    hipStream_t s = ctx->na_stream;
    double *d_out = static_cast<double *>(ctx->na_scratch);
    char *arena = static_cast<char *>(ctx->na_scratch) + out_bytes;
    yq::JobResult res;
    try {
        YQ_HIP(hipSetDevice(ctx->device));
        // the scratch keeps the previous call's contents: clear the output
        yq::fill_zero<double>(reinterpret_cast<yq::c64 *>(d_out), 0, (nout + 1) / 2, 1, s);
        yq::tune_freeze(true);
        for (int64_t t0 = 0; t0 < ntraj; t0 += B) {
            const int nb = (int)std::min<int64_t>(B, ntraj - t0);
            yq::ArenaScope scope(arena, (size_t)(cap - out_bytes));
            auto eng = yq::make_engine(in, (int)B, s, &ctx->na_res);
            eng->run(t0, nb, d_out, nullptr, res);
            YQ_HIP(hipStreamSynchronize(s));
        }
        yq::tune_freeze(false);
        if (!res.failed.empty())
            return fail(YAQS_ENUMERIC, "NaN/Inf in the two-site tensor before a split in " +
                                           std::to_string(res.failed.size()) + " trajectories (first: " +
                                           std::to_string(res.failed.front()) + ")");
        if (traj_out && nobs)
            YQ_HIP(hipMemcpyAsync(traj_out, d_out, (size_t)nobs * ncol * ntraj * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
    } catch (const yq::ArenaFull &e) {
        yq::tune_freeze(false);
        (void)hipStreamSynchronize(s);
        return fail(YAQS_ENOMEM_DEVICE, e.what());
    } catch (const yq::HipError &e) {
        yq::tune_freeze(false);
        (void)hipStreamSynchronize(s);
        return fail(YAQS_EDEVICE, e.what());
    } catch (const std::invalid_argument &e) {
        yq::tune_freeze(false);
        (void)hipStreamSynchronize(s);
        return fail(YAQS_EINVAL, e.what());
    } catch (const std::exception &e) {
        yq::tune_freeze(false);
        (void)hipStreamSynchronize(s);
        return fail(YAQS_EDEVICE, e.what());
    }
    if (mean_out && nobs) {
        for (int64_t q = 0; q < nobs; ++q)
            for (int64_t c = 0; c < ncol; ++c) mean_out[q * ncol + c] = 0.0;
        if (traj_out) {
            for (int64_t q = 0; q < nobs; ++q)
                for (int64_t c = 0; c < ncol; ++c) {
                    double sum = 0;
                    for (int64_t t = 0; t < ntraj; ++t) sum += traj_out[(q * ncol + c) * ntraj + t];
                    mean_out[q * ncol + c] = sum / (double)ntraj;
                }
        } else {
            std::vector<double> h((size_t)nobs * ncol * ntraj);
            if (hipMemcpy(h.data(), d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(YAQS_EDEVICE, "copy of the job output failed");
            for (int64_t q = 0; q < nobs; ++q)
                for (int64_t c = 0; c < ncol; ++c) {
                    double sum = 0;
                    for (int64_t t = 0; t < ntraj; ++t) sum += h[(size_t)(q * ncol + c) * ntraj + t];
                    mean_out[q * ncol + c] = sum / (double)ntraj;
                }
        }
    }
    if (stats) {
        stats->wall_s = now_s() - t_submit;
        stats->queue_s = 0;
        stats->flop_model = res.flop;
        stats->matvecs = res.matvecs;
        stats->jumps = res.jumps;
        stats->max_bond = res.max_bond;
        stats->num_samples = ncol;
        stats->fp64_retries = 0;
    }
    return YAQS_OK;
}

int yaqs_bench_analog(yaqs_ctx *ctx, const yaqs_analog_job *job, int64_t warmup_steps, double *traj_out,
                      double *mean_out, yaqs_bench_report *report) {
    if (!ctx) return fail(YAQS_EINVAL, "ctx is NULL");
    try {
        yq::JobIn in = convert_analog(job);
        if (warmup_steps < 0 || warmup_steps + 2 > in.num_times)
            throw std::invalid_argument("warmup_steps must leave at least one timed step");
        in.bench_warmup = warmup_steps;
        yq::validate_job(in);
        int64_t bytes = 0;
        int conc = 1;
        int64_t b = choose_batch(ctx, in, ctx->budget, &bytes, &conc);
        if (b < 1) return fail(YAQS_ENOMEM_DEVICE, "job footprint exceeds the HBM budget");
        if (hipSetDevice(ctx->device) != hipSuccess) return fail(YAQS_EDEVICE, "hipSetDevice failed");
        // The waves run as the context would run them: side by side on
        // their own streams when choose_batch split the job for concurrency
        // (concurrent_waves), else one after the other on one stream.
        const int64_t ntraj = in.num_traj;
        const bool split = conc > 1;
        const int nconc = conc;
        in.wsync = split ? std::make_shared<yq::WindowSync>(nconc) : nullptr;
        const int64_t nout = std::max<int64_t>(1, (int64_t)in.obs.size() * yq::num_columns(in) * ntraj);
        double *d_out = nullptr;
        YQ_HIP(hipMalloc(&d_out, nout * sizeof(double)));
        YQ_HIP(hipMemset(d_out, 0, nout * sizeof(double)));
        YQ_HIP(hipStreamSynchronize(nullptr));  // ordered before the non-blocking worker streams
        std::vector<yq::JobResult> res(nconc);
        std::vector<std::string> errs(nconc);
        auto lane = [&](int w) {
            hipStream_t stream = nullptr;
            yq::StreamRes sres;
            try {
                YQ_HIP(hipSetDevice(ctx->device));
                YQ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
                sres.create();
                for (int64_t t0 = (int64_t)w * b; t0 < ntraj; t0 += (int64_t)nconc * b) {
                    const int nb = (int)std::min<int64_t>(b, ntraj - t0);
                    auto eng = yq::make_engine(in, (int)b, stream, &sres);
                    eng->run(t0, nb, d_out, nullptr, res[w]);
                }
                YQ_HIP(hipStreamSynchronize(stream));
            } catch (const std::exception &e) {
                errs[w] = e.what();
                if (in.wsync) in.wsync->release();
            }
            if (stream) (void)hipStreamSynchronize(stream);
            sres.destroy();
            if (stream) (void)hipStreamDestroy(stream);
        };
        if (nconc == 1) {
            lane(0);
        } else {
            std::vector<std::thread> th;
            for (int w = 0; w < nconc; ++w) th.emplace_back(lane, w);
            for (auto &t : th) t.join();
        }
        for (auto &e : errs)
            if (!e.empty()) {
                (void)hipFree(d_out);
                return fail(YAQS_EDEVICE, e);
            }
        std::vector<double> traj((size_t)nout);
        YQ_HIP(hipMemcpy(traj.data(), d_out, nout * sizeof(double), hipMemcpyDeviceToHost));
        (void)hipFree(d_out);
        const int64_t nobs = (int64_t)in.obs.size(), ncol = yq::num_columns(in);
        if (traj_out && nobs) std::copy(traj.begin(), traj.begin() + nobs * ncol * ntraj, traj_out);
        if (mean_out && nobs)
            for (int64_t q = 0; q < nobs; ++q)
                for (int64_t c = 0; c < ncol; ++c) {
                    double s = 0;
                    for (int64_t t = 0; t < ntraj; ++t) s += traj[((size_t)q * ncol + c) * ntraj + t];
                    mean_out[q * ncol + c] = s / (double)ntraj;
                }
        if (report) {
            *report = yaqs_bench_report{};
            for (const auto &r : res) {
                // concurrent waves share the window start: the window is the longest of them
                report->timed_s = split ? std::max(report->timed_s, r.timed_s) : report->timed_s + r.timed_s;
                report->timed_steps += r.timed_steps;
                report->gemm_ms += r.gemm_ms;
                report->gemm_launches += r.gemm_launches;
                report->gemm_flop += r.gemm_flop;
                report->flop_model += r.win_flop;
                report->matvecs += r.win_matvecs;
                report->max_bond = std::max<int64_t>(report->max_bond, r.max_bond);
            }
            report->concurrent_waves = nconc;
        }
        return YAQS_OK;
    } catch (const yq::HipError &e) {
        return fail(YAQS_EDEVICE, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(YAQS_EINVAL, e.what());
    } catch (const std::exception &e) {
        return fail(YAQS_EDEVICE, e.what());
    }
}

}  // extern "C"
