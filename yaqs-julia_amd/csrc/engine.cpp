// Host orchestration of the batched TJM engine.
//
// All trajectories of a wave advance in lockstep through the same sequence of
// batched kernels; per-trajectory differences (jumps, Krylov convergence,
// truncation ranks) live in device-side masks and scalars.  Bond dimensions
// are zero-padded to the wave's maximum at each bond (`wd`), which leaves
// every contraction exact.  Conventions (engine layout, all row-major):
//   site tensor   A_i[a][s][b]           (wd[i] x 2 x wd[i+1])
//   MPO           W_i[m][s'][s][n]       (Dm[i] x 2 x 2 x Dm[i+1])
//   left env      EL[b][bra][mpo][ket]   (wd[b] x Dm[b] x wd[b])
//   right env     ER[b][mpo][ket][bra]   (Dm[b] x wd[b] x wd[b])
//   two-site Θ    [a][2 s1 + s2][b]
#include "engine.h"

#include <algorithm>
#include <functional>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <type_traits>
#include <stdexcept>

#include "kernels.h"
#include "tune.h"

namespace yq {

namespace {

typedef std::complex<double> cd;

double host_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------ host utils ---
// Bump arena of the calling thread (ArenaScope), or none.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
};
thread_local Arena *g_arena = nullptr;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool owned = false;  // hipMalloc'd (else carved from an arena)
    DevBuf() {}
    explicit DevBuf(size_t b) { alloc(b); }
    void alloc(size_t b) {
        release();
        bytes = b ? b : 16;
        if (g_arena) {
            const size_t off = (g_arena->used + 255) & ~(size_t)255;
            if (off + bytes > g_arena->size)
                throw ArenaFull("scratch arena exhausted: " + std::to_string(off + bytes) + " bytes needed, " +
                                std::to_string(g_arena->size) + " available");
            p = g_arena->base + off;
            g_arena->used = off + bytes;
            owned = false;
            return;
        }
        YQ_HIP(hipMalloc(&p, bytes));
        owned = true;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        owned = false;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned) {
        o.p = nullptr;
        o.bytes = 0;
        o.owned = false;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            owned = o.owned;
            o.p = nullptr;
            o.bytes = 0;
            o.owned = false;
        }
        return *this;
    }
    template <typename X>
    X *as() const { return static_cast<X *>(p); }
};

template <typename X>
void upload(DevBuf &b, const std::vector<X> &v, hipStream_t s) {
    b.alloc(std::max<size_t>(1, v.size()) * sizeof(X));
    if (!v.empty()) YQ_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, s));
}

c64 mk(double r, double i) {
    c64 z;
    z.re = r;
    z.im = i;
    return z;
}
cd toc(const c64 &z) { return cd(z.re, z.im); }
c64 fromc(const cd &z) { return mk(z.real(), z.imag()); }

// 2x2 reference-order (column-major) -> engine row-major M[r][c]
void op2_engine(const c64 *ref, c64 *out) {
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) out[r * 2 + c] = ref[r + 2 * c];
}
// 4x4 reference-order (column-major, index p1 + 2 p2) -> engine row-major on 2 s1 + s2
void op4_engine(const c64 *ref, c64 *out) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            const int R = 2 * (r & 1) + (r >> 1), C = 2 * (c & 1) + (c >> 1);
            out[R * 4 + C] = ref[r + 4 * c];
        }
}

// exp(-scale * L^†L) for a 2x2 L (engine row-major), closed form for the
// Hermitian PSD 2x2 (the role of exp(::Matrix), src/Dissipation.jl:63-64).
void damp2(const c64 *Le, double scale, c64 *out) {
    cd L[4] = {toc(Le[0]), toc(Le[1]), toc(Le[2]), toc(Le[3])};
    cd M[4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) M[i * 2 + j] = std::conj(L[0 * 2 + i]) * L[0 * 2 + j] + std::conj(L[1 * 2 + i]) * L[1 * 2 + j];
    const double a = M[0].real(), d = M[3].real();
    const cd b = M[1];
    const double mid = 0.5 * (a + d), rad = std::sqrt(0.25 * (a - d) * (a - d) + std::norm(b));
    const double l1 = mid + rad, l2 = mid - rad;
    const double e1 = std::exp(-scale * l1), e2 = std::exp(-scale * l2);
    // f(M) = e2 I + (e1 - e2)/(l1 - l2) (M - l2 I)   (rad > 0), else e1 I
    cd R[4];
    if (rad < 1e-300) {
        R[0] = e1; R[1] = 0; R[2] = 0; R[3] = e1;
    } else {
        const double f = (e1 - e2) / (l1 - l2);
        R[0] = e2 + f * (M[0] - l2);
        R[1] = f * M[1];
        R[2] = f * M[2];
        R[3] = e2 + f * (M[3] - l2);
    }
    for (int i = 0; i < 4; ++i) out[i] = fromc(R[i]);
}

// exp(-scale * L^†L) for a 4x4 L (engine row-major) by scaling and squaring
// of the Taylor series (L^†L is Hermitian PSD; scale*||L^†L|| / 2^s <= 1/2,
// terms until they fall below 1e-18 of the sum): the 4x4 `exp` of
// src/Dissipation.jl:94-95.
void damp4(const c64 *Le, double scale, c64 *out) {
    cd M[16];
    double nrm = 0;
    for (int r = 0; r < 4; ++r) {
        double row = 0;
        for (int c = 0; c < 4; ++c) {
            cd acc = 0;
            for (int k = 0; k < 4; ++k) acc += std::conj(toc(Le[k * 4 + r])) * toc(Le[k * 4 + c]);
            M[r * 4 + c] = -scale * acc;
            row += std::abs(M[r * 4 + c]);
        }
        nrm = std::max(nrm, row);
    }
    int sq = 0;
    while (nrm > 0.5 && sq < 60) { nrm *= 0.5; ++sq; }
    const double f = std::ldexp(1.0, -sq);
    for (auto &z : M) z *= f;
    cd E[16], Tm[16], U[16];
    for (int k = 0; k < 16; ++k) E[k] = Tm[k] = (k % 5 == 0) ? cd(1, 0) : cd(0, 0);
    auto mul = [](const cd *X, const cd *Y, cd *Z) {
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                cd acc = 0;
                for (int k = 0; k < 4; ++k) acc += X[r * 4 + k] * Y[k * 4 + c];
                Z[r * 4 + c] = acc;
            }
    };
    for (int n = 1; n <= 40; ++n) {
        mul(Tm, M, U);
        double tn = 0, en = 0;
        for (int k = 0; k < 16; ++k) {
            Tm[k] = U[k] / double(n);
            E[k] += Tm[k];
            tn += std::abs(Tm[k]);
            en += std::abs(E[k]);
        }
        if (tn <= 1e-18 * en) break;
    }
    for (int q = 0; q < sq; ++q) {
        mul(E, E, U);
        for (int k = 0; k < 16; ++k) E[k] = U[k];
    }
    for (int k = 0; k < 16; ++k) out[k] = fromc(E[k]);
}

void matmul2(const c64 *A, const c64 *B, c64 *C) {  // C = A B (2x2 row-major)
    c64 r[4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r[i * 2 + j] = fromc(toc(A[i * 2]) * toc(B[j]) + toc(A[i * 2 + 1]) * toc(B[2 + j]));
    std::memcpy(C, r, sizeof(r));
}

bool is_unitary2(const c64 *m) {  // engine or reference order: U^†U = I is order-invariant
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            cd acc = 0;
            for (int k = 0; k < 2; ++k) acc += std::conj(toc(m[k * 2 + i])) * toc(m[k * 2 + j]);
            if (std::abs(acc - (i == j ? 1.0 : 0.0)) > 1e-10) return false;
        }
    return true;
}

// Factor a 4x4 (engine order, index 2 s1 + s2) as kron(A, B): M[(2a+b)][(2c+e)] = A[a][c] B[b][e].
bool kron_factor(const c64 *M, c64 *A, c64 *B) {
    int bi = 0;
    double best = -1;
    for (int i = 0; i < 16; ++i) {
        double v = std::abs(toc(M[i]));
        if (v > best) { best = v; bi = i; }
    }
    if (best <= 0) return false;
    const int R = bi / 4, C = bi % 4;
    const int a0 = R >> 1, b0 = R & 1, c0 = C >> 1, e0 = C & 1;
    // B[b][e] = M[(2a0+b)][(2c0+e)], A[a][c] = M[(2a+b0)][(2c+e0)] / B[b0][e0]
    for (int b = 0; b < 2; ++b)
        for (int e = 0; e < 2; ++e) B[b * 2 + e] = M[(2 * a0 + b) * 4 + (2 * c0 + e)];
    const cd piv = toc(B[b0 * 2 + e0]);
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) A[a * 2 + c] = fromc(toc(M[(2 * a + b0) * 4 + (2 * c + e0)]) / piv);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c)
                for (int e = 0; e < 2; ++e)
                    if (std::abs(toc(A[a * 2 + c]) * toc(B[b * 2 + e]) - toc(M[(2 * a + b) * 4 + (2 * c + e)])) > 1e-12)
                        return false;
    return true;
}

// CSR of the effective-Hamiltonian MPO step Y[p][j][q] = Σ_i W[j][i] X[p][i][q].
struct Csr {
    std::vector<int> rowptr, col;
    std::vector<c64> val;
    int I = 0, J = 0;
};
struct CsrDev {
    DevBuf rowptr, col, val;
    int I = 0, J = 0, nnz = 0;
};
void upload_csr(CsrDev &d, const Csr &c, hipStream_t s) {
    upload(d.rowptr, c.rowptr, s);
    upload(d.col, c.col, s);
    upload(d.val, c.val, s);
    d.I = c.I;
    d.J = c.J;
    d.nnz = (int)c.col.size();
}
// Left form (matvec and left env): i = m d + s, j = s' Dr + n.
Csr csr_left(const std::vector<c64> &W, int Dl, int d, int Dr) {
    Csr c;
    c.I = Dl * d;
    c.J = d * Dr;
    c.rowptr.push_back(0);
    for (int sp = 0; sp < d; ++sp)
        for (int n = 0; n < Dr; ++n) {
            for (int m = 0; m < Dl; ++m)
                for (int s = 0; s < d; ++s) {
                    const c64 w = W[((m * d + sp) * d + s) * Dr + n];
                    if (w.re != 0.0 || w.im != 0.0) {
                        c.col.push_back(m * d + s);
                        c.val.push_back(w);
                    }
                }
            c.rowptr.push_back((int)c.col.size());
        }
    return c;
}
// Right form (right env): i = s Dr + n, j = n' d + s' with n' the left MPO index.
Csr csr_right(const std::vector<c64> &W, int Dl, int d, int Dr) {
    Csr c;
    c.I = d * Dr;
    c.J = Dl * d;
    c.rowptr.push_back(0);
    for (int np = 0; np < Dl; ++np)
        for (int sp = 0; sp < d; ++sp) {
            for (int s = 0; s < d; ++s)
                for (int n = 0; n < Dr; ++n) {
                    const c64 w = W[((np * d + sp) * d + s) * Dr + n];
                    if (w.re != 0.0 || w.im != 0.0) {
                        c.col.push_back(s * Dr + n);
                        c.val.push_back(w);
                    }
                }
            c.rowptr.push_back((int)c.col.size());
        }
    return c;
}
// Two-site MPO W2[m][S'][S][n] with S = 2 s1 + s2 (src/Algorithms.jl:639-643).
std::vector<c64> merge_mpo(const std::vector<c64> &W1, const std::vector<c64> &W2, int Dl, int Dk, int Dr) {
    std::vector<c64> G((size_t)Dl * 16 * Dr, mk(0, 0));
    for (int m = 0; m < Dl; ++m)
        for (int k = 0; k < Dk; ++k)
            for (int n = 0; n < Dr; ++n)
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const cd w1 = toc(W1[((m * 2 + a) * 2 + b) * Dk + k]);
                        if (w1 == 0.0) continue;
                        for (int c = 0; c < 2; ++c)
                            for (int e = 0; e < 2; ++e) {
                                const cd w2 = toc(W2[((k * 2 + c) * 2 + e) * Dr + n]);
                                const int So = 2 * a + c, Si = 2 * b + e;
                                c64 &g = G[((m * 4 + So) * 4 + Si) * Dr + n];
                                g = fromc(toc(g) + w1 * w2);
                            }
                    }
    return G;
}

GemmDesc gd(int M, int N, int K, int opA, int opB, int64_t lda, int64_t ldb, int64_t ldc, int64_t sA, int64_t sB,
            int64_t sC, int outer, const int *active = nullptr) {
    GemmDesc d;
    d.M = M; d.N = N; d.K = K; d.opA = opA; d.opB = opB;
    d.lda = lda; d.ldb = ldb; d.ldc = ldc;
    d.sA_outer = sA; d.sB_outer = sB; d.sC_outer = sC;
    d.sA_inner = d.sB_inner = d.sC_inner = 0;
    d.inner = 1; d.outer = outer; d.accumulate = 0; d.active = active;
    return d;
}

// HIP-event timing of matrix-core GEMM launches inside a measurement window.
// A launch's FLOP = per-unit FLOP x units (active trajectories); launches
// queued before their active count is known carry units -1 until resolve().
struct GemmProf {
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> fl;
    std::vector<int> un;
    size_t used = 0;
    double ms = 0, flop = 0, pend = 0;
    int64_t n = 0;
    hipStream_t s = nullptr;
    void start(hipStream_t st) {
        s = st;
        if (ev.empty()) {
            ev.resize(4096);
            for (auto &e : ev) YQ_HIP(hipEventCreate(&e));
            fl.resize(2048);
            un.resize(2048);
        }
        used = 0;
        ms = flop = pend = 0;
        n = 0;
        on = true;
    }
    void record(double per_unit, int units) {
        fl[used / 2] = per_unit;
        un[used / 2] = units;
        used += 2;
    }
    void resolve(int units) {
        for (size_t i = 0; i < used; i += 2)
            if (un[i / 2] < 0) un[i / 2] = units;
        flop += pend * units;
        pend = 0;
    }
    void flush() {
        if (!used) return;
        YQ_HIP(hipEventSynchronize(ev[used - 1]));
        for (size_t i = 0; i < used; i += 2) {
            float x = 0;
            YQ_HIP(hipEventElapsedTime(&x, ev[i], ev[i + 1]));
            ms += x;
            if (un[i / 2] < 0) pend += fl[i / 2];
            else flop += fl[i / 2] * un[i / 2];
            ++n;
        }
        used = 0;
    }
    void stop() {
        flush();
        on = false;
    }
    ~GemmProf() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
};
thread_local GemmProf *g_prof = nullptr;

// Test hook (yaqs_debug_poison_theta): NaN stored into the two-site tensor
// of one trajectory before its given split, for engines of one precision.
std::atomic<int64_t> g_poison_traj{-1};
std::atomic<int64_t> g_poison_split{-1};
std::atomic<int> g_poison_prec{-1};
// Trajectories active in the current masked launch (Lanczos), -1 = all, -2 =
// not known yet (a Lanczos iteration queued ahead of its count): the
// profiler counts algorithmic FLOP of active trajectories only.
thread_local int g_nact = -1;
inline int launch_units(const GemmDesc &d) {
    if (!d.active || g_nact == -1) return d.outer;
    return g_nact >= 0 ? g_nact : -1;
}

// Tuner keys: GEMM dimensions in units of 16 (the MFMA block and the k
// tile), MPO-step dimensions P and Q in units of 8, batch counts by power
// of two, and whether the direct-to-LDS variant applies.
inline int b8(int64_t x) { return (int)((x + 7) / 8); }
inline int b16(int64_t x) { return (int)((x + 15) / 16); }
inline int pow2b(int64_t x) {
    int b = 0;
    while ((int64_t)1 << b < x) ++b;
    return b;
}

template <typename T>
void gemmT(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C, hipStream_t s) {
    // tile shape: every shape gives the same bits, so it is picked by speed
    // (a launch queued ahead of its active count takes the cached choice
    // only: a measurement would wait on the stream)
    int tile = 0;
    if (!d.accumulate && d.opB == OP_N && tune_enabled()) {
        const TuneKey k = {0,        (int)sizeof(T),    d.opA, d.opB, b16(d.M), b16(d.N), b16(d.K), d.inner,
                           pow2b(d.outer), gemm_dl_fits<T>(d, A, B) ? 1 : 0};
        tile = g_nact == -2 ? cached_variant(k)
                            : tuned_variant(
                                  k, GEMM_TILES,
                                  [&](int v) {
                                      if (!gemm_mfma<T>(d, A, B, C, s, v)) gemm<T, T, T, T>(d, A, B, C, s);
                                  },
                                  s);
    }
    GemmProf *p = g_prof;
    const bool prof = p && p->on && !tune_measuring();
    if (prof) {
        if (p->used + 2 > p->ev.size()) p->flush();
        YQ_HIP(hipEventRecord(p->ev[p->used], s));
    }
    const bool mf = gemm_mfma<T>(d, A, B, C, s, tile);
    if (!mf) gemm<T, T, T, T>(d, A, B, C, s);
    if (prof) {
        if (mf) {
            YQ_HIP(hipEventRecord(p->ev[p->used + 1], s));
            p->record(8.0 * d.M * (double)d.N * d.K * d.inner, launch_units(d));
        }
    }
}

// T1 = A B followed by the MPO step into Y: the GEMM with the MPO step
// fused into its epilogue, or `unfused` (the GEMM into T1, then wstep),
// whichever the tuner measured faster for the shape (same bits either way).
template <typename T>
void gemmW(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, int P, int Dl, int dph, int Dr, int Q, cplx<T> *Y,
           int64_t sY, const CsrDev &W, const std::function<void()> &unfused, hipStream_t s) {
    // variants: 0 fused, 1 unfused, 2 fused with k tiles loaded straight
    // into LDS (where that kernel applies)
    auto fused = [&](bool dl) {
        return gemm_w_mfma<T>(d, A, B, P, Dl, dph, Dr, Q, Y, sY, W.rowptr.as<int>(), W.col.as<int>(),
                              W.val.as<c64>(), W.nnz, s, dl);
    };
    int var = gemm_w_covers(d, P, Dl, dph, Dr, Q, W.nnz) ? 0 : 1;
    if (var == 0 && tune_enabled()) {
        const bool dl = gemm_w_dl_fits<T>(d, A, B, Dl);
        const int nv = dl ? 3 : 2;
        const TuneKey k = {dl ? 3 : 1, (int)sizeof(T), Dl, dph, Dr, b8(P), b8(Q), b16(d.K), W.nnz, pow2b(d.outer)};
        var = g_nact == -2 ? cached_variant(k)
                           : tuned_variant(
                                 k, nv,
                                 [&](int v) {
                                     if (v == 1) unfused();
                                     else (void)fused(v == 2);
                                 },
                                 s);
    }
    if (var == 1) {
        unfused();
        return;
    }
    GemmProf *p = g_prof;
    const bool prof = p && p->on && !tune_measuring();
    if (prof) {
        if (p->used + 2 > p->ev.size()) p->flush();
        YQ_HIP(hipEventRecord(p->ev[p->used], s));
    }
    (void)fused(var == 2);
    if (prof) {
        YQ_HIP(hipEventRecord(p->ev[p->used + 1], s));
        // GEMM plus the sparse MPO step (one complex MAC per W non-zero per (p, q))
        p->record(8.0 * d.M * (double)d.N * d.K * d.inner + 8.0 * W.nnz * (double)P * Q, launch_units(d));
    }
}

// --------------------------------------------------------------- engine ----
struct State {
    std::vector<DevBuf> A;  // per site, B slots
    std::vector<int> wd;    // working dims per bond
    DevBuf dims;            // [B][L+1] true dims
};

struct PairCand {  // jump candidates of the local noise model of a gate pair
    DevBuf site, op, site2, op2, LdL, gamma;
    int n = 0;
    // long-range processes on a long-range gate's pair: candidate kinds (0,
    // or 2 for a general MPO jump), the general ones as (candidate, index
    // into lr_cands), and the product of their dissipation scalars (dt = 1)
    DevBuf kind, scal;
    std::vector<std::pair<int, int>> lr;
    bool has_scal = false;
};

template <typename T>
class Engine {
  public:
    typedef cplx<T> CT;

    Engine(const JobIn &job, int B, hipStream_t s, StreamRes *res = nullptr) : J(job), B(B), s(s) {
        if (res) {
            h_pinned = res->h_pinned;
            d_pinned = res->d_pinned;
            lz_ev[0] = res->ev[0];
            lz_ev[1] = res->ev[1];
        } else {
            own.create();
            h_pinned = own.h_pinned;
            d_pinned = own.d_pinned;
            lz_ev[0] = own.ev[0];
            lz_ev[1] = own.ev[1];
        }
        try {
            setup();
        } catch (...) {
            (void)hipStreamSynchronize(s);
            own.destroy();
            throw;
        }
    }
    ~Engine() {
        // nothing of this engine may still be in flight when its buffers,
        // pinned words and events go (an exception can end a wave between
        // launches)
        (void)hipStreamSynchronize(s);
        own.destroy();
    }
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    void run_analog(int64_t t0, int nb, double *d_out, JobResult &res);
    // appends the trajectories the NaN/Inf guard flagged in this wave
    void collect_failed(JobResult &res);
    void run_digital(int64_t t0, int nb, double *d_out, int64_t *h_bonds, JobResult &res);

    static int64_t footprint(const JobIn &job, int B);

  private:

    const JobIn &J;
    int B;      // slots allocated
    int nb = 0; // trajectories in the current wave
    hipStream_t s;
    int L = 0, kdim = 25;
    std::vector<int> cap, Dm;
    std::vector<int64_t> slotA, slotE;
    int64_t nmax = 0, slotS = 0, gdim = 0, slotG = 0, capmax = 0, slotX = 0;
    State PHI, PSI;
    std::vector<DevBuf> EL, ER;
    std::vector<CsrDev> csrL, csrR, csr2;
    DevBuf theta, V, T1, T2, G, U, evals, perm, keep, Cf, absorb, Zbuf, Pbuf, Hwork, Hfb;
    DevBuf cgsS, cgsF;  // CGS2 projection scratch and Frobenius norms
    int64_t slotZ = 0, slotP = 0, slotH = 0, slotCgs = 0;
    DevBuf Ebuf0, Ebuf1, Xbuf, rho1, rho2;
    DevBuf beta0, galpha, alpha, beta, coef, nit, active, part, partn, nactive;
    DevBuf norm2v, fac, jstate, chosen, opidx, mask, rngk, wbuf, count;
    DevBuf diss_half, diss_full, diss_one;  // per-site 2x2 dissipation operators
    // Adjacent 2-site non-Pauli processes (analog): the dissipation runs the
    // reference's sweep (merge, 4x4 exp(-dt/2 γ L†L), relative-threshold
    // split, centre shift per site) and jumps on a pair are merged-pair
    // jumps; d4_at[i]: indices (of 4x4 operators) of the processes whose
    // upper site is i, in process order
    bool has2 = false;
    std::vector<std::vector<int>> d4_at;
    DevBuf diss4_half, diss4_full;  // 16 c64 per 2-site non-Pauli process
    bool scr_ready = false;         // SCR holds ψ~ of the last sweep (see stochastic)
    // merged-pair jump candidates: cand_pair[c] = index of the candidate's
    // 4x4 operator in pair_ops (-1: not a pair jump), cand_kind[c] = 1 when
    // its weight is Tr(L†L ρ2) (cand_LdL4), 0 for Tr(L†L ρ1)
    DevBuf cand_pair, cand_kind, cand_LdL4, pair_ops, opidx4;
    std::vector<char> cand2_at;        // pairs (i, i+1) whose ρ2 the weights need
    std::vector<char> pair_cand_site;  // pairs with merged-pair jump candidates
    std::vector<int> h_cand_site, h_cand_pair;
    DevBuf rho2b, mask2;
    DevBuf invL_half, invL_full;            // inverse of the last site's operator
    bool last_site_damped = false;
    std::vector<int> diss_site_active;
    State SCR;       // scratch state for the reference's jump weights (see stochastic)
    DevBuf rho1b;
    DevBuf cand_site, cand_op, cand_site2, cand_op2, cand_LdL, cand_gamma, jump_ops;
    int ncand = 0;
    // general long-range candidates (apply_mpo_jump!): candidate index and
    // process; lr_ops: their sig, tau, sig^†sig, tau^†tau; lrnu[t][cand]:
    // ν of the weight dt γ ν; lr_val: two correlators per trajectory
    struct LrCand {
        int cand = 0, proc = 0;
    };
    std::vector<LrCand> lr_cands;
    DevBuf lr_ops, lrnu, lr_val;
    DevBuf obs_site, obs_n, obs_op;
    bool any_obs2 = false;
    std::vector<char> obs2_at;
    StreamRes own;            // the engine's own host objects when none were passed
    int *h_pinned = nullptr;
    int *d_pinned = nullptr;  // device address of h_pinned
    DevBuf fin_done;          // lz_finalize's last-workgroup ticket
    DevBuf nfail;             // [B] NaN/Inf guard flags of the wave
    // FP64 splits: kept singular values below 1e-6.5 of the largest are
    // resolved by a deflated second Gram (see split)
    bool defl = false;
    DevBuf Dq, Dcnt, Dflag, keepR, dflag_count;
    int64_t slotDq = 0;
    int64_t nsplit = 0;       // splits since the wave started (test hook)
    hipEvent_t lz_ev[2] = {nullptr, nullptr};  // Lanczos active counts land in h_pinned[1 + (j & 1)]
    // The count the finalize kernel stores over the -1 left in
    // h_pinned[slot]: the host spins on the pinned word (the next launches
    // go out as soon as the count lands, without the event's wake-up), then
    // falls back to the event after some tens of milliseconds.
    int wait_pinned(int slot, hipEvent_t ev) {
        for (int i = 0; i < 200000; ++i) {
            const int v = __atomic_load_n(h_pinned + slot, __ATOMIC_ACQUIRE);
            if (v >= 0) return v;
            __builtin_ia32_pause();
        }
        YQ_HIP(hipEventSynchronize(ev));
        return __atomic_load_n(h_pinned + slot, __ATOMIC_ACQUIRE);
    }
    uint64_t traj0 = 0;
    int64_t nchunk_max = 0;
    std::map<int, PairCand> pair_cand;  // digital candidate tables per gate pair
    std::map<int, PairCand> lr_cand;    // per long-range gate pair (a, b), key a * L + b
    DevBuf dsite_ops;                   // digital 1-site dissipation (dt = 1) per site: [q * 4]
    std::vector<int> dsite_active;
    DevBuf gate_ops;  // digital: gate matrices (engine order), 16 c64 per gate
    std::vector<int64_t> gate_off;
    int64_t swap_off = -1;  // SWAP matrix in gate_ops
    std::map<int, std::vector<CsrDev>> win_csr;  // per long-range TDVP gate: left, right, pair tables
    DevBuf dissd_ops;  // digital dissipation (dt = 1) per gate pair and site: [(2 a + r) * 4]
    std::vector<int> dgt_active;
    // gate pairs whose two dissipators are multiples of the identity (Pauli
    // processes only): dscal_ops[a * 4] = their product, applied to the centre
    DevBuf dscal_ops;
    std::vector<int> dgt_scalar;

    void setup();
    void alloc_state(State &S);
    void init_state(State &S);
    void copy_state(const State &S, State &D);

    CT *A(State &S, int i) { return S.A[i].template as<CT>(); }
    CT *Vq(int q) { return V.as<CT>() + (int64_t)q * nmax; }
    int64_t sV() const { return (int64_t)(kdim + 1) * nmax; }
    int sync_count(int *dcount);

    // primitives
    void matvec(State &S, bool two, int i, const CT *x, int64_t sx, CT *y, int64_t sy, const int *act);
    void matvec0(State &S, int b, const CT *x, int64_t sx, CT *y, int64_t sy, const int *act);
    void expm(State &S, int sites, int i, CT *x, int64_t sx, double dt, JobResult &res);
    // `sites` counts sites (2, 1, 0 = bond): a bool there is a mistake
    void expm(State &S, bool sites, int i, CT *x, int64_t sx, double dt, JobResult &res) = delete;
    void merge(State &S, int i);
    void split(State &S, int i, bool right, int rule, double thr, const int *msk = nullptr, int max_keep = 0);
    void deflate_tail(State &S, int i, bool right, int kw, int mb);
    void env_left(State &S, int i);
    void env_right(State &S, int i);
    void tdvp(State &S, JobResult &res);
    void tdvp1(State &S, JobResult &res);
    void lq_sweep(State &S, int from, const int *msk);
    void qr_step(State &S, int i, const int *msk);
    void dissipate(State &S, bool half, int field = 0);
    void dissipate_sweep(State &S, bool half, int field);
    void lq_step(State &S, int i, const int *msk, bool keep_factor = true);
    void regrow(State &S, int b, int w);
    void shrink_to_dims(State &S);
    void transfer(State &S, const int *msk, c64 *rho_out, c64 *rho2_out = nullptr,
                  const std::vector<char> *at2 = nullptr);
    void measure(State &S, int col, int ncol, double *d_out);
    void stochastic(State &S, double dt, const DevBuf *invL, JobResult &res);
    void absorb_left(State &S, int i, const int *msk);
    void correlate(State &S, int a, int b, const c64 *opa, const c64 *opb, const int *msk, c64 *out);
    void mpo_jump(State &S, int li, const int *msk);
};

// A general long-range jump (apply_mpo_jump!) doubles the bonds between its
// two sites before truncate! cuts them back: those bonds get twice the
// capacity (once, whatever the number of processes over them).
static void lr_double_caps(const JobIn &job, std::vector<int> &cap) {
    std::vector<char> dbl(cap.size(), 0);
    for (auto &p : job.noise)
        if (p.lr && !p.pauli_name)
            for (int k = std::min(p.s0, p.s1) + 1; k <= std::max(p.s0, p.s1); ++k) dbl[k] = 1;
    for (size_t k = 0; k < cap.size(); ++k)
        if (dbl[k]) cap[k] *= 2;
}

template <typename T>
int64_t Engine<T>::footprint(const JobIn &job, int B) {
    // Device bytes of an engine with B trajectory slots (mirrors setup()).
    const int L = (int)job.L;
    std::vector<int> cap(L + 1, 1);
    for (int b = 1; b < L; ++b) {
        int e = std::min(b, L - b);
        int64_t c = e >= 30 ? (int64_t)1 << 30 : (int64_t)1 << e;
        cap[b] = (int)std::min<int64_t>(c, job.max_bond);
        cap[b] = std::max(cap[b], std::max(job.mps_dr[b - 1], job.mps_dl[b]));
    }
    lr_double_caps(job, cap);
    int Dmax = 1;
    if (!job.digital)
        for (int i = 0; i < L; ++i) Dmax = std::max(Dmax, std::max(job.mpo_dl[i], job.mpo_dr[i]));
    int64_t cm = 1, sites = 0, envs = 0, nmax = 0, smax = 0, gd = 2;
    for (int b = 0; b <= L; ++b) cm = std::max<int64_t>(cm, cap[b]);
    for (int i = 0; i < L; ++i) sites += (int64_t)cap[i] * 2 * cap[i + 1];
    for (int b = 0; b <= L; ++b) envs += (int64_t)cap[b] * cap[b] * Dmax;
    nmax = 2 * cm * cm;
    smax = 2 * cm * cm * Dmax;
    for (int i = 0; i + 1 < L; ++i) {
        nmax = std::max<int64_t>(nmax, (int64_t)cap[i] * 4 * cap[i + 2]);
        smax = std::max<int64_t>(smax, (int64_t)cap[i] * Dmax * 4 * cap[i + 2]);
        gd = std::max<int64_t>(gd, 2 * std::max(cap[i], cap[i + 2]));
    }
    const bool noise = job.has_noise;
    int64_t elems = sites * (job.digital ? 1 : (noise ? 3 : 2));  // PHI (+ PSI, + scratch)
    bool windows = false;  // digital: long-range TDVP windows use the environments and Lanczos basis too
    for (auto &g : job.gates) windows = windows || g.win_long;
    if (!job.digital || windows) elems += 2 * envs + (int64_t)(job.kdim + 1) * nmax + 2 * smax;
    elems += nmax + cm * cm + sites;  // theta, Cf, absorb (bounded by the state size)
    elems += cm * 32;                 // CGS2 projection scratch
    int64_t bytes = elems * (int64_t)sizeof(cplx<T>);
    bytes += 2 * gd * gd * (int64_t)sizeof(c64);                       // Gram + eigenvectors
    bytes += (gd * gd + 2 * gd + 4) * (int64_t)sizeof(double);         // eigensolver scratch
    if (gd > 128) bytes += gd * (gd + 1) / 2 * (int64_t)sizeof(c64);
    bytes += heev_work_bytes((int)gd, (int)std::min<int64_t>(gd, job.max_bond)) + (int64_t)sizeof(int);  // eigensolver work
    bytes += (2 * cm * cm + std::max<int64_t>(nmax, 2 * cm * cm)) * (int64_t)sizeof(c64);  // transfer
    bytes += (int64_t)L * 20 * sizeof(c64) * 2 + (job.kdim + 2) * 64 * 16 + 256 + 8;
    if (std::is_same<T, double>::value && !job.digital) bytes += gd * cm * (int64_t)sizeof(cplx<T>) + 16;  // Dq
    return bytes * B;
}

template <typename T>
void Engine<T>::setup() {
    L = (int)J.L;
    kdim = J.kdim;
    cap.assign(L + 1, 1);
    for (int b = 1; b < L; ++b) {
        int e = std::min(b, L - b);
        int64_t c = e >= 30 ? (int64_t)1 << 30 : (int64_t)1 << e;
        cap[b] = (int)std::min<int64_t>(c, J.max_bond);
        cap[b] = std::max(cap[b], std::max(J.mps_dr[b - 1], J.mps_dl[b]));
    }
    cap[0] = J.mps_dl[0];
    cap[L] = J.mps_dr[L - 1];
    lr_double_caps(J, cap);
    Dm.assign(L + 1, 1);
    Dm[0] = J.mpo.empty() ? 1 : J.mpo_dl[0];
    for (int i = 0; i < L; ++i) Dm[i + 1] = J.mpo.empty() ? 1 : J.mpo_dr[i];
    int Dmax = *std::max_element(Dm.begin(), Dm.end());
    slotA.resize(L);
    for (int i = 0; i < L; ++i) slotA[i] = (int64_t)cap[i] * 2 * cap[i + 1];
    slotE.resize(L + 1);
    for (int b = 0; b <= L; ++b) slotE[b] = (int64_t)cap[b] * cap[b] * Dm[b];
    capmax = *std::max_element(cap.begin(), cap.end());
    nmax = 2 * capmax * capmax;
    slotS = (int64_t)capmax * Dmax * 2 * capmax;
    gdim = 2;
    for (int i = 0; i + 1 < L; ++i) {
        nmax = std::max<int64_t>(nmax, (int64_t)cap[i] * 4 * cap[i + 2]);
        slotS = std::max<int64_t>(slotS, (int64_t)cap[i] * Dmax * 4 * cap[i + 2]);
        gdim = std::max<int64_t>(gdim, 2 * std::max(cap[i], cap[i + 2]));
    }
    slotG = gdim * gdim;
    slotX = std::max<int64_t>(nmax, 2 * capmax * capmax);
    alloc_state(PHI);
    if (!J.digital) alloc_state(PSI);
    EL.resize(L + 1);
    ER.resize(L + 1);
    bool need_win = false;  // digital: long-range TDVP windows run the TDVP steps with window MPOs
    for (auto &G2 : J.gates) need_win = need_win || G2.win_long;
    if (!J.digital || need_win) {
        for (int b = 0; b <= L; ++b) {
            EL[b].alloc(slotE[b] * B * sizeof(CT));
            ER[b].alloc(slotE[b] * B * sizeof(CT));
        }
        csrL.resize(L);
        csrR.resize(L);
        csr2.resize(std::max(0, L - 1));
        if (!J.digital) {
            for (int i = 0; i < L; ++i) {
                upload_csr(csrL[i], csr_left(J.mpo[i], Dm[i], 2, Dm[i + 1]), s);
                upload_csr(csrR[i], csr_right(J.mpo[i], Dm[i], 2, Dm[i + 1]), s);
            }
            for (int i = 0; i + 1 < L; ++i)
                upload_csr(csr2[i], csr_left(merge_mpo(J.mpo[i], J.mpo[i + 1], Dm[i], Dm[i + 1], Dm[i + 2]), Dm[i],
                                             4, Dm[i + 2]), s);
        } else {
            // window MPO of each long-range TDVP gate (construct_window_mpo,
            // src/DigitalTJM.jl:192-236): the Kronecker factors of h on a and b
            // (unique up to a scalar that cancels in every effective operator),
            // identities between, bonds of dimension 1
            const std::vector<c64> WI = {mk(1, 0), mk(0, 0), mk(0, 0), mk(1, 0)};
            for (size_t g = 0; g < J.gates.size(); ++g) {
                const GateIn &G2 = J.gates[g];
                if (!G2.win_long) continue;
                const int a = std::min(G2.s0, G2.s1), b = std::max(G2.s0, G2.s1), n = b - a + 1;
                c64 e4[16], Af[4], Bf[4];
                op4_engine(G2.m, e4);
                if (!kron_factor(e4, Af, Bf))
                    throw std::invalid_argument("the window Hamiltonian of a long-range gate must be g_a (x) g_b");
                const std::vector<c64> WA(Af, Af + 4), WB(Bf, Bf + 4);
                auto W = [&](int q) -> const std::vector<c64> & { return q == 0 ? WA : (q == n - 1 ? WB : WI); };
                std::vector<CsrDev> v((size_t)3 * n - 1);
                for (int q = 0; q < n; ++q) {
                    upload_csr(v[q], csr_left(W(q), 1, 2, 1), s);
                    upload_csr(v[n + q], csr_right(W(q), 1, 2, 1), s);
                }
                for (int q = 0; q + 1 < n; ++q)
                    upload_csr(v[2 * n + q], csr_left(merge_mpo(W(q), W(q + 1), 1, 1, 1), 1, 4, 1), s);
                win_csr[(int)g] = std::move(v);
            }
        }
        V.alloc((size_t)B * sV() * sizeof(CT));
        T1.alloc((size_t)B * slotS * sizeof(CT));
        T2.alloc((size_t)B * slotS * sizeof(CT));
    }
    theta.alloc((size_t)B * nmax * sizeof(CT));
    G.alloc((size_t)B * slotG * sizeof(c64));
    U.alloc((size_t)B * slotG * sizeof(c64));
    slotZ = gdim * gdim + 2 * gdim + 4;
    Zbuf.alloc((size_t)B * slotZ * sizeof(double));
    if (gdim > 128) {
        slotP = gdim * (gdim + 1) / 2;
        Pbuf.alloc((size_t)B * slotP * sizeof(double2));
    }
    slotH = heev_work_bytes((int)gdim, (int)std::min<int64_t>(gdim, J.max_bond));
    if (slotH) {
        Hwork.alloc((size_t)B * slotH);
        Hfb.alloc((size_t)B * sizeof(int));
    }
    evals.alloc((size_t)B * (gdim + 2) * sizeof(double));
    perm.alloc((size_t)B * (gdim + 2) * sizeof(int));
    keep.alloc((size_t)B * sizeof(int));
    Cf.alloc((size_t)B * capmax * capmax * sizeof(CT));
    slotCgs = (int64_t)capmax * 32;  // nvec x panel width (<= 32)
    cgsS.alloc((size_t)B * slotCgs * sizeof(CT));
    cgsF.alloc((size_t)B * sizeof(double));
    int64_t amax = 0;
    for (int i = 0; i < L; ++i) amax = std::max(amax, slotA[i]);
    absorb.alloc((size_t)B * amax * sizeof(CT));
    Ebuf0.alloc((size_t)B * capmax * capmax * sizeof(c64));
    Ebuf1.alloc((size_t)B * capmax * capmax * sizeof(c64));
    Xbuf.alloc((size_t)B * slotX * sizeof(c64));
    rho1.alloc((size_t)B * L * 4 * sizeof(c64));
    rho2.alloc((size_t)B * L * 16 * sizeof(c64));
    nchunk_max = (nmax + LZ_CHUNK_ELEMS - 1) / LZ_CHUNK_ELEMS;
    beta0.alloc(B * sizeof(double));
    galpha.alloc(B * sizeof(double));
    alpha.alloc((size_t)B * kdim * sizeof(double));
    beta.alloc((size_t)B * (kdim + 1) * sizeof(double));
    coef.alloc((size_t)B * kdim * sizeof(c64));
    nit.alloc(B * sizeof(int));
    active.alloc(B * sizeof(int));
    part.alloc((size_t)B * (kdim + 1) * nchunk_max * sizeof(c64));
    partn.alloc((size_t)B * nchunk_max * sizeof(double));
    nactive.alloc(kdim * sizeof(int));  // one counter per Lanczos iteration
    norm2v.alloc(B * sizeof(double));
    fac.alloc(B * sizeof(double));
    jstate.alloc(B * sizeof(int));
    chosen.alloc(B * sizeof(int));
    opidx.alloc(B * sizeof(int));
    mask.alloc(B * sizeof(int));
    rngk.alloc(B * sizeof(uint64_t));
    count.alloc(sizeof(int));
    fin_done.alloc(sizeof(unsigned long long));
    fill_int(fin_done.as<int>(), 0, 2, s);
    nfail.alloc((size_t)B * sizeof(int));
    static const bool defl_off = [] {
        const char *e = std::getenv("YAQS_SPLIT_DEFLATE");
        return e && e[0] == '0';
    }();
    defl = std::is_same<T, double>::value && !J.digital && !defl_off;
    if (defl) {
        slotDq = gdim * capmax;
        Dq.alloc((size_t)B * slotDq * sizeof(CT));
        Dcnt.alloc((size_t)B * sizeof(int));
        Dflag.alloc((size_t)B * sizeof(int));
        keepR.alloc((size_t)B * sizeof(int));
        dflag_count.alloc(sizeof(int));
    }

    // ---- noise: dissipation operators per site and jump candidates
    auto site_ops = [&](double dt, std::vector<c64> &ops, std::vector<int> *act) {
        ops.assign((size_t)L * 4, mk(0, 0));
        if (act) act->assign(L, 0);
        for (int i = 0; i < L; ++i) { ops[i * 4 + 0] = mk(1, 0); ops[i * 4 + 3] = mk(1, 0); }
        bool allzero = true;
        for (auto &p : J.noise) allzero = allzero && p.gamma == 0.0;
        if (!J.has_noise || allzero) return;
        for (int i = 0; i < L; ++i) {
            for (auto &p : J.noise) {
                if (p.nsites == 1 && p.s0 == i) {
                    c64 M[4];
                    if (p.pauli_name) {
                        const double f = std::exp(-0.5 * dt * p.gamma);
                        M[0] = mk(f, 0); M[1] = mk(0, 0); M[2] = mk(0, 0); M[3] = mk(f, 0);
                    } else {
                        c64 Le[4];
                        op2_engine(p.m, Le);
                        damp2(Le, 0.5 * dt * p.gamma, M);
                    }
                    matmul2(M, &ops[i * 4], &ops[i * 4]);
                    if (act) (*act)[i] = 1;
                }
            }
            // 2-site Pauli processes and long-range processes (handle_longrange_scalar!,
            // src/Dissipation.jl:161-204) scale site max(s0, s1)
            for (auto &p : J.noise) {
                if (p.nsites == 2 && std::max(p.s0, p.s1) == i && (p.pauli_name || p.lr)) {
                    const double f = std::exp(-0.5 * dt * p.gamma);
                    for (int k = 0; k < 4; ++k) ops[i * 4 + k] = mk(ops[i * 4 + k].re * f, ops[i * 4 + k].im * f);
                    if (act) (*act)[i] = 1;
                }
            }
        }
    };
    std::vector<c64> hops;
    auto inv_last = [&](const std::vector<c64> &ops, DevBuf &dst) {
        const c64 *M = &ops[(size_t)(L - 1) * 4];
        const cd a = toc(M[0]), b = toc(M[1]), c = toc(M[2]), d = toc(M[3]);
        const cd det = a * d - b * c;
        std::vector<c64> inv = {fromc(d / det), fromc(-b / det), fromc(-c / det), fromc(a / det)};
        upload(dst, inv, s);
    };
    site_ops(J.dt / 2, hops, &diss_site_active);
    upload(diss_half, hops, s);
    inv_last(hops, invL_half);
    last_site_damped = L >= 2 && diss_site_active[L - 1];
    site_ops(J.dt, hops, nullptr);
    upload(diss_full, hops, s);
    inv_last(hops, invL_full);
    // adjacent 2-site non-Pauli processes: 4x4 dissipators (dt/2 and dt)
    d4_at.assign(L, {});
    {
        bool allzero = true;
        for (auto &p : J.noise) allzero = allzero && p.gamma == 0.0;
        std::vector<c64> h4, f4;
        int n4 = 0;
        if (!J.digital && J.has_noise && !allzero)
            for (auto &p : J.noise) {
                if (p.nsites != 2 || p.pauli_name || p.lr) continue;
                c64 e4[16], M[16];
                op4_engine(p.m, e4);
                damp4(e4, 0.5 * (J.dt / 2) * p.gamma, M);
                h4.insert(h4.end(), M, M + 16);
                damp4(e4, 0.5 * J.dt * p.gamma, M);
                f4.insert(f4.end(), M, M + 16);
                d4_at[std::max(p.s0, p.s1)].push_back(n4++);
                has2 = true;
            }
        upload(diss4_half, h4, s);
        upload(diss4_full, f4, s);
    }
    if (!J.digital && J.has_noise && (last_site_damped || has2)) {
        alloc_state(SCR);
        rho1b.alloc((size_t)B * L * 4 * sizeof(c64));
        rho2b.alloc((size_t)B * L * 16 * sizeof(c64));
    }
    site_ops(1.0, hops, nullptr);
    upload(diss_one, hops, s);

    // jump operator table: per process (and second factor for 2-site Pauli)
    std::vector<c64> jops;
    std::vector<int> proc_op(J.noise.size(), -1), proc_op2(J.noise.size(), -1), proc_site2(J.noise.size(), -1);
    for (size_t q = 0; q < J.noise.size(); ++q) {
        const ProcIn &p = J.noise[q];
        if (p.nsites == 1) {
            c64 e[4];
            op2_engine(p.m, e);
            proc_op[q] = (int)(jops.size() / 4);
            jops.insert(jops.end(), e, e + 4);
        } else if (!J.digital) {
            // analog: a merged-pair jump (src/StochasticProcess.jl:266-309)
        } else if (p.lr) {
            // long-range: no operator ("pauli_long_range") or an MPO jump (mpo_jump)
        } else {
            c64 e4[16], Af[4], Bf[4];
            op4_engine(p.m, e4);
            if (!kron_factor(e4, Af, Bf)) throw std::invalid_argument("non-product 2-site jump operator");
            // engine index 2 s1 + s2: A acts on site min, B on site max
            proc_op[q] = (int)(jops.size() / 4);
            jops.insert(jops.end(), Af, Af + 4);
            proc_op2[q] = (int)(jops.size() / 4);
            jops.insert(jops.end(), Bf, Bf + 4);
            proc_site2[q] = std::max(p.s0, p.s1);
        }
    }
    upload(jump_ops, jops, s);
    // analog candidates in the reference's order (src/StochasticProcess.jl:97-167)
    {
        std::vector<int> cs, co, cs2, co2, cpair, ckind;
        std::vector<c64> ldl, ldl4, pops;
        std::vector<double> cg;
        auto add = [&](size_t q, int site, bool two) {
            const ProcIn &p = J.noise[q];
            cs.push_back(site);
            co.push_back(proc_op[q]);
            cs2.push_back(two ? proc_site2[q] : -1);
            co2.push_back(two ? proc_op2[q] : -1);
            // analog pair candidates: the 4x4 jump operator, and L†L for the
            // merged-pair weight of processes that is_pauli_proc
            // (src/StochasticProcess.jl:343-353) does not call Pauli: names
            // with pauli/crosstalk (not projector) or a unitary matrix weigh
            // by the centre norm instead (src/StochasticProcess.jl:143-147)
            int pidx = -1, kind = 0;
            c64 M4[16];
            for (auto &z : M4) z = mk(0, 0);
            if (two && !J.digital) {
                c64 e4[16];
                op4_engine(p.m, e4);
                pidx = (int)(pops.size() / 16);
                pops.insert(pops.end(), e4, e4 + 16);
                bool unitary = true;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) {
                        cd acc = 0;
                        for (int k = 0; k < 4; ++k) acc += std::conj(toc(e4[k * 4 + r])) * toc(e4[k * 4 + c]);
                        M4[r * 4 + c] = fromc(acc);
                        unitary = unitary && std::abs(acc - (r == c ? 1.0 : 0.0)) <= 1e-10;
                    }
                if (!p.pauli_name && !unitary) kind = 1;
            }
            cpair.push_back(pidx);
            ckind.push_back(kind);
            ldl4.insert(ldl4.end(), M4, M4 + 16);
            c64 M[4];
            if (!two) {
                c64 e[4];
                op2_engine(p.m, e);
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) M[i * 2 + j] = fromc(std::conj(toc(e[0 * 2 + i])) * toc(e[0 * 2 + j]) +
                                                                     std::conj(toc(e[1 * 2 + i])) * toc(e[1 * 2 + j]));
            } else {  // Pauli product: weight = dt γ ‖ψ‖² = dt γ Tr ρ_site
                M[0] = mk(1, 0); M[1] = mk(0, 0); M[2] = mk(0, 0); M[3] = mk(1, 0);
            }
            ldl.insert(ldl.end(), M, M + 4);
            cg.push_back(p.gamma);
        };
        for (int site = 0; site < L; ++site) {
            for (size_t q = 0; q < J.noise.size(); ++q)
                if (J.noise[q].nsites == 1 && J.noise[q].s0 == site) add(q, site, false);
            for (size_t q = 0; q < J.noise.size(); ++q) {
                const ProcIn &p = J.noise[q];
                if (p.nsites == 2 && !p.lr && std::min(p.s0, p.s1) == site && p.s0 == site && site + 1 < L)
                    add(q, site, true);
            }
        }
        // long-range candidates after the site loop, in process order
        // (src/StochasticProcess.jl:168-198): "pauli_long_range" weighs
        // dt γ ‖ψ‖² (Tr ρ of the last site, whose ρ is the true state's) and
        // applies no operator; a general one weighs dt γ ν with ν = 1 unless
        // O annihilates the state (lr_nu, per jump step), kind 2
        for (size_t q = 0; q < J.noise.size(); ++q) {
            const ProcIn &p = J.noise[q];
            if (!p.lr) continue;
            cs.push_back(L - 1);
            co.push_back(-1);
            cs2.push_back(-1);
            co2.push_back(-1);
            cpair.push_back(-1);
            ckind.push_back(p.pauli_name ? 0 : 2);
            for (int k = 0; k < 16; ++k) ldl4.push_back(mk(0, 0));
            const c64 I2[4] = {mk(1, 0), mk(0, 0), mk(0, 0), mk(1, 0)};
            ldl.insert(ldl.end(), I2, I2 + 4);
            cg.push_back(p.gamma);
            if (!p.pauli_name) {
                LrCand lc;
                lc.cand = (int)cs.size() - 1;
                lc.proc = (int)q;
                lr_cands.push_back(lc);
            }
        }
        ncand = (int)cs.size();
        cand2_at.assign(L, 0);
        pair_cand_site.assign(L, 0);
        for (int c = 0; c < ncand; ++c) {
            if (ckind[c] == 1) cand2_at[cs[c]] = 1;
            if (cpair[c] >= 0) pair_cand_site[cs[c]] = 1;
        }
        h_cand_site = cs;
        h_cand_pair = cpair;
        mask2.alloc((size_t)B * sizeof(int));
        upload(cand_pair, cpair, s);
        upload(cand_kind, ckind, s);
        upload(cand_LdL4, ldl4, s);
        upload(pair_ops, pops, s);
        opidx4.alloc((size_t)B * sizeof(int));
        upload(cand_site, cs, s);
        upload(cand_op, co, s);
        upload(cand_site2, cs2, s);
        upload(cand_op2, co2, s);
        upload(cand_LdL, ldl, s);
        upload(cand_gamma, cg, s);
        wbuf.alloc((size_t)B * std::max(1, ncand) * sizeof(double));
        if (!lr_cands.empty()) {
            // per general long-range candidate: sig, tau, sig^†sig, tau^†tau (engine order)
            std::vector<c64> lops;
            for (auto &lc : lr_cands) {
                const ProcIn &p = J.noise[lc.proc];
                c64 e[4], f[4];
                op2_engine(p.sig, e);
                op2_engine(p.tau, f);
                lops.insert(lops.end(), e, e + 4);
                lops.insert(lops.end(), f, f + 4);
                for (const c64 *o : {e, f}) {  // o^† o, row-major
                    c64 g[4];
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            g[i * 2 + j] = fromc(std::conj(toc(o[0 * 2 + i])) * toc(o[0 * 2 + j]) +
                                                 std::conj(toc(o[1 * 2 + i])) * toc(o[1 * 2 + j]));
                    lops.insert(lops.end(), g, g + 4);
                }
            }
            upload(lr_ops, lops, s);
            lrnu.alloc((size_t)B * ncand * sizeof(double));
            lr_val.alloc((size_t)2 * B * sizeof(c64));
        }
        if (J.digital) {
            // per adjacent pair (s1, s1+1): processes on s1, pair processes, processes on s2
            for (int s1 = 0; s1 + 1 < L; ++s1) {
                const int s2 = s1 + 1;
                std::vector<int> ps, po, ps2, po2;
                std::vector<c64> pl;
                std::vector<double> pg;
                cs.clear(); co.clear(); cs2.clear(); co2.clear(); ldl.clear(); cg.clear();
                for (int site : {s1, s2}) {
                    for (size_t q = 0; q < J.noise.size(); ++q)
                        if (J.noise[q].nsites == 1 && J.noise[q].s0 == site) add(q, site, false);
                    for (size_t q = 0; q < J.noise.size(); ++q) {
                        const ProcIn &p = J.noise[q];
                        if (site == s1 && p.nsites == 2 && p.s0 == s1 && p.s1 == s2) add(q, site, true);
                    }
                }
                PairCand &pc = pair_cand[s1];
                pc.n = (int)cs.size();
                upload(pc.site, cs, s);
                upload(pc.op, co, s);
                upload(pc.site2, cs2, s);
                upload(pc.op2, co2, s);
                upload(pc.LdL, ldl, s);
                upload(pc.gamma, cg, s);
            }
            // local dissipation operators of a gate on (a, a+1) (create_local_noise_model,
            // src/DigitalTJM.jl:172-188, then apply_dissipation with dt = 1)
            std::vector<c64> dops((size_t)std::max(1, 2 * L) * 4, mk(0, 0));
            dgt_active.assign((size_t)2 * L, 0);
            bool allzero = true;
            for (auto &p : J.noise) allzero = allzero && p.gamma == 0.0;
            for (int a = 0; a + 1 < L; ++a)
                for (int r = 0; r < 2; ++r) {
                    const int q = a + r;
                    c64 *M = &dops[(size_t)(2 * a + r) * 4];
                    M[0] = mk(1, 0);
                    M[3] = mk(1, 0);
                    if (!J.has_noise || allzero) continue;
                    for (auto &p : J.noise) {
                        if (p.nsites == 1 && p.s0 == q) {
                            c64 D[4];
                            if (p.pauli_name) {
                                const double f = std::exp(-0.5 * p.gamma);
                                D[0] = mk(f, 0); D[1] = mk(0, 0); D[2] = mk(0, 0); D[3] = mk(f, 0);
                            } else {
                                c64 Le[4];
                                op2_engine(p.m, Le);
                                damp2(Le, 0.5 * p.gamma, D);
                            }
                            matmul2(D, M, M);
                            dgt_active[2 * a + r] = 1;
                        }
                        if (r == 1 && p.nsites == 2 && p.pauli_name && std::min(p.s0, p.s1) == a &&
                            std::max(p.s0, p.s1) == a + 1) {
                            const double f = std::exp(-0.5 * p.gamma);
                            for (int k = 0; k < 4; ++k) M[k] = mk(M[k].re * f, M[k].im * f);
                            dgt_active[2 * a + r] = 1;
                        }
                    }
                }
            upload(dissd_ops, dops, s);
            std::vector<c64> sops((size_t)std::max(1, L) * 4, mk(0, 0));
            dgt_scalar.assign((size_t)L, 0);
            for (int a = 0; a + 1 < L; ++a) {
                bool sc = true;
                double f = 1.0;
                for (int r = 0; r < 2; ++r) {
                    const c64 *M = &dops[(size_t)(2 * a + r) * 4];
                    sc = sc && M[1].re == 0 && M[1].im == 0 && M[2].re == 0 && M[2].im == 0 && M[0].im == 0 &&
                         M[3].im == 0 && M[0].re == M[3].re;
                    f *= M[0].re;
                }
                dgt_scalar[a] = sc ? 1 : 0;
                sops[(size_t)a * 4] = mk(f, 0);
                sops[(size_t)a * 4 + 3] = mk(f, 0);
            }
            upload(dscal_ops, sops, s);
            // long-range gates (SWAP networks): candidates of the processes on a, then
            // on b (create_local_noise_model, src/DigitalTJM.jl:172-188), and each
            // site's 1-site dissipator (the r = 0 part of the pair operators above)
            std::vector<c64> dso((size_t)L * 4, mk(0, 0));
            dsite_active.assign((size_t)L, 0);
            for (int q = 0; q < L; ++q) {
                c64 *M = &dso[(size_t)q * 4];
                M[0] = mk(1, 0);
                M[3] = mk(1, 0);
                if (!J.has_noise || allzero) continue;
                for (auto &p : J.noise) {
                    if (p.nsites != 1 || p.s0 != q) continue;
                    c64 D[4];
                    if (p.pauli_name) {
                        const double f = std::exp(-0.5 * p.gamma);
                        D[0] = mk(f, 0); D[1] = mk(0, 0); D[2] = mk(0, 0); D[3] = mk(f, 0);
                    } else {
                        c64 Le[4];
                        op2_engine(p.m, Le);
                        damp2(Le, 0.5 * p.gamma, D);
                    }
                    matmul2(D, M, M);
                    dsite_active[q] = 1;
                }
            }
            upload(dsite_ops, dso, s);
            for (const GateIn &G2 : J.gates) {
                if (G2.kind != 2) continue;
                const int la = std::min(G2.s0, G2.s1), lb = std::max(G2.s0, G2.s1);
                if (lb <= la + 1 || lr_cand.count(la * L + lb)) continue;
                cs.clear(); co.clear(); cs2.clear(); co2.clear(); ldl.clear(); cg.clear();
                for (int site : {la, lb})
                    for (size_t q = 0; q < J.noise.size(); ++q)
                        if (J.noise[q].nsites == 1 && J.noise[q].s0 == site) add(q, site, false);
                PairCand &pc = lr_cand[la * L + lb];
                // long-range processes on exactly this pair (create_local_noise_model,
                // src/DigitalTJM.jl:172-188), after the 1-site ones as in the reference's
                // candidate list: weight on ρ_lb (the true state's norm) for
                // "pauli_long_range", dt γ ν for a general MPO jump
                std::vector<int> kinds(cs.size(), 0);
                double fsc = 1.0;
                for (size_t q = 0; q < J.noise.size(); ++q) {
                    const ProcIn &p = J.noise[q];
                    if (!p.lr || std::min(p.s0, p.s1) != la || std::max(p.s0, p.s1) != lb) continue;
                    if (!allzero) fsc *= std::exp(-0.5 * p.gamma);
                    cs.push_back(lb);
                    co.push_back(-1);
                    cs2.push_back(-1);
                    co2.push_back(-1);
                    const c64 I2[4] = {mk(1, 0), mk(0, 0), mk(0, 0), mk(1, 0)};
                    ldl.insert(ldl.end(), I2, I2 + 4);
                    cg.push_back(p.gamma);
                    kinds.push_back(p.pauli_name ? 0 : 2);
                    if (!p.pauli_name)
                        for (size_t li = 0; li < lr_cands.size(); ++li)
                            if (lr_cands[li].proc == (int)q) pc.lr.push_back({(int)cs.size() - 1, (int)li});
                    pc.has_scal = true;
                }
                if (pc.has_scal) {
                    const std::vector<c64> sc = {mk(fsc, 0), mk(0, 0), mk(0, 0), mk(fsc, 0)};
                    upload(pc.scal, sc, s);
                }
                upload(pc.kind, kinds, s);
                pc.n = (int)cs.size();
                upload(pc.site, cs, s);
                upload(pc.op, co, s);
                upload(pc.site2, cs2, s);
                upload(pc.op2, co2, s);
                upload(pc.LdL, ldl, s);
                upload(pc.gamma, cg, s);
            }
            std::vector<c64> gops;
            gate_off.assign(J.gates.size(), -1);
            for (size_t g = 0; g < J.gates.size(); ++g) {
                const GateIn &G2 = J.gates[g];
                if (G2.kind == 1) {
                    c64 e[4];
                    op2_engine(G2.m, e);
                    gate_off[g] = (int64_t)gops.size();
                    gops.insert(gops.end(), e, e + 4);
                } else if (G2.kind == 2) {
                    c64 e[16];
                    op4_engine(G2.m, e);
                    gate_off[g] = (int64_t)gops.size();
                    gops.insert(gops.end(), e, e + 16);
                }
            }
            {  // the SWAP of the long-range TEBD network (src/GateLibrary.jl SWAPGate)
                c64 sw[16], e[16];
                for (int k = 0; k < 16; ++k) sw[k] = mk(0, 0);
                sw[0] = sw[6] = sw[9] = sw[15] = mk(1, 0);
                op4_engine(sw, e);
                swap_off = (int64_t)gops.size();
                gops.insert(gops.end(), e, e + 16);
            }
            upload(gate_ops, gops, s);
        }
    }
    // observables
    {
        std::vector<int> os, on;
        std::vector<c64> oo;
        obs2_at.assign(L, 0);
        for (auto &o : J.obs) {
            os.push_back(std::min(o.s0, o.s1));
            on.push_back(o.nsites);
            c64 e[16];
            for (auto &z : e) z = mk(0, 0);
            if (o.nsites == 1) op2_engine(o.op, e);
            else {
                op4_engine(o.op, e);
                obs2_at[std::min(o.s0, o.s1)] = 1;
                any_obs2 = true;
            }
            oo.insert(oo.end(), e, e + 16);
        }
        upload(obs_site, os, s);
        upload(obs_n, on, s);
        upload(obs_op, oo, s);
    }
    YQ_HIP(hipStreamSynchronize(s));
}

template <typename T>
void Engine<T>::alloc_state(State &S) {
    S.A.resize(L);
    for (int i = 0; i < L; ++i) S.A[i].alloc((size_t)B * slotA[i] * sizeof(CT));
    S.dims.alloc((size_t)B * (L + 1) * sizeof(int));
    S.wd.assign(L + 1, 1);
}

template <typename T>
void Engine<T>::init_state(State &S) {
    // broadcast the (normalized) initial MPS into every slot
    for (int i = 0; i < L; ++i) {
        const int n = J.mps_dl[i] * 2 * J.mps_dr[i];
        DevBuf tmp;
        upload(tmp, J.mps[i], s);
        copy2d<double, T>(tmp.as<c64>(), 0, n, A(S, i), slotA[i], n, 1, n, nullptr, nb, s);
        YQ_HIP(hipStreamSynchronize(s));
    }
    S.wd[0] = J.mps_dl[0];
    for (int i = 0; i < L; ++i) S.wd[i + 1] = J.mps_dr[i];
    std::vector<int> hd((size_t)B * (L + 1));
    for (int t = 0; t < B; ++t)
        for (int b = 0; b <= L; ++b) hd[(size_t)t * (L + 1) + b] = S.wd[b];
    YQ_HIP(hipMemcpyAsync(S.dims.p, hd.data(), hd.size() * sizeof(int), hipMemcpyHostToDevice, s));
    YQ_HIP(hipStreamSynchronize(s));
}

template <typename T>
void Engine<T>::copy_state(const State &S, State &D) {
    for (int i = 0; i < L; ++i)
        YQ_HIP(hipMemcpyAsync(D.A[i].p, S.A[i].p, (size_t)nb * slotA[i] * sizeof(CT), hipMemcpyDeviceToDevice, s));
    YQ_HIP(hipMemcpyAsync(D.dims.p, S.dims.p, (size_t)B * (L + 1) * sizeof(int), hipMemcpyDeviceToDevice, s));
    D.wd = S.wd;
}

template <typename T>
int Engine<T>::sync_count(int *dcount) {
    YQ_HIP(hipMemcpyAsync(h_pinned, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
    YQ_HIP(hipStreamSynchronize(s));
    return h_pinned[0];
}

// y = H_eff x (src/Algorithms.jl:330-361): two GEMMs around the MPO step.
template <typename T>
void Engine<T>::matvec(State &S, bool two, int i, const CT *x, int64_t sx, CT *y, int64_t sy, const int *act) {
    const int d = two ? 4 : 2;
    const int bl = i, br = two ? i + 2 : i + 1;
    const int cl = S.wd[bl], cr = S.wd[br], Dl = Dm[bl], Dr = Dm[br];
    const CsrDev &W = two ? csr2[i] : csrL[i];
    CT *t1 = T1.as<CT>(), *t2 = T2.as<CT>();
    const GemmDesc g1 = gd(cl * Dl, d * cr, cl, OP_N, OP_N, cl, d * cr, d * cr, slotE[bl], sx, slotS, nb, act);
    gemmW<T>(g1, EL[bl].as<CT>(), x, cl, Dl, d, Dr, cr, t2, slotS, W,
             [&]() {
                 gemmT<T>(g1, EL[bl].as<CT>(), x, t1, s);
                 wstep<T>(t1, t2, cl, Dl * d, d * Dr, cr, slotS, slotS, W.rowptr.as<int>(), W.col.as<int>(),
                          W.val.as<c64>(), act, nb, s);
             },
             s);
    gemmT<T>(gd(cl * d, cr, Dr * cr, OP_N, OP_N, Dr * cr, cr, cr, slotS, slotE[br], sy, nb, act), t2,
             ER[br].as<CT>(), y, s);
}

// The bond (0-site) effective Hamiltonian of the 1-site sweep (_bond_op,
// src/Algorithms.jl:222-226): y[a'][b'] = sum EL[b][a'][w][a] x[a][b] ER[b][w][b][b'],
// as EL·x then ·ER; x and y have leading dimension wd[b] rounded up to even
// and a zero padding column.
template <typename T>
void Engine<T>::matvec0(State &S, int b, const CT *x, int64_t sx, CT *y, int64_t sy, const int *act) {
    const int c = S.wd[b], cp = (c + 1) & ~1, D = Dm[b];
    CT *t1 = T1.as<CT>();
    gemmT<T>(gd(c * D, c, c, OP_N, OP_N, c, cp, c, slotE[b], sx, slotS, nb, act), EL[b].as<CT>(), x, t1, s);
    gemmT<T>(gd(c, c, D * c, OP_N, OP_N, D * c, c, cp, slotS, slotE[b], sy, nb, act), t1, ER[b].as<CT>(), y, s);
    if (cp != c) zero_col<T>(y, sy, c, cp, c, act, nb, s);
}

// x <- exp(-i dt H_eff) x by batched Lanczos (src/Algorithms.jl:144-187).
// sites = 2: Θ of sites i, i+1; 1: site i; 0: the bond matrix of bond i
// (rows wd[i], leading dimension wd[i] rounded up to even, the padding
// column zero).
template <typename T>
void Engine<T>::expm(State &S, int sites, int i, CT *x, int64_t sx, double dt, JobResult &res) {
    const bool two = sites == 2;
    const int d = two ? 4 : 2;
    const int cl = S.wd[i], cr = S.wd[two ? i + 2 : i + 1];
    const int64_t n = sites == 0 ? (int64_t)cl * ((cl + 1) & ~1) : (int64_t)cl * d * cr;
    const double fl = sites == 0 ? 8.0 * 2.0 * Dm[i] * (double)cl * cl * cl
                      : two ? 8.0 * (4.0 * Dm[i] * cl * cr * (double)(cl + cr) + 16.0 * Dm[i] * Dm[i] * (double)cl * cr)
                            : 8.0 * (2.0 * Dm[i] * cl * cr * (double)(cl + cr) + 4.0 * Dm[i] * Dm[i] * (double)cl * cr);
    lz_init<T>(x, sx, V.as<CT>(), sV(), n, beta0.as<double>(), active.as<int>(), nit.as<int>(), nactive.as<int>(),
               kdim, nb, s);
    // One Lanczos iteration for the still-active trajectories (every kernel is
    // masked by the device-side active flags); its active count is copied to
    // the host slot of its parity.
    auto iteration = [&](int j, int units) {
        g_nact = units;
        if (sites == 0) matvec0(S, i, Vq(j), sV(), Vq(j + 1), sV(), active.as<int>());
        else matvec(S, two, i, Vq(j), sV(), Vq(j + 1), sV(), active.as<int>());
        g_nact = -1;
        lz_dots<T>(V.as<CT>(), sV(), nmax, n, j, part.as<c64>(), kdim + 1, active.as<int>(), nb, s);
        lz_update<T>(V.as<CT>(), sV(), nmax, n, j, part.as<c64>(), kdim + 1, partn.as<double>(), galpha.as<double>(),
                     active.as<int>(), beta.as<double>(), kdim, nb, s);
        // the kernel's count overwrites the -1 left in the host slot (see wait_pinned)
        __atomic_store_n(h_pinned + 1 + (j & 1), -1, __ATOMIC_RELEASE);
        lz_finalize(j, kdim, dt, J.ktol, partn.as<double>(), n, galpha.as<double>(), alpha.as<double>(),
                    beta.as<double>(), beta0.as<double>(), coef.as<c64>(), nit.as<int>(), active.as<int>(),
                    nactive.as<int>(), nb, fin_done.as<unsigned long long>(), d_pinned + 1 + (j & 1), s);
        YQ_HIP(hipEventRecord(lz_ev[j & 1], s));
    };
    // With `ahead`, iteration j+1 is queued before the host waits for the
    // count of iteration j, so the device does not idle on that round trip;
    // when iteration j left no trajectory active, j+1 ran fully masked.
    // YAQS_LZ_AHEAD=1 queues iteration j+1 before waiting for the count of
    // iteration j (measured neutral at C2 and slightly slower on the serving
    // stream, where the masked extra iterations compete for the device, so
    // off by default)
    static const bool ahead = [] {
        const char *e = std::getenv("YAQS_LZ_AHEAD");
        return e && e[0] == '1';
    }();
    int nact = nb;
    iteration(0, nact);
    res.matvecs += nact;
    res.flop += fl * nact;
    for (int j = 0; j < kdim; ++j) {
        const bool more = j + 1 < kdim;
        if (more && ahead) iteration(j + 1, -2);  // active count resolved below
        nact = wait_pinned(1 + (j & 1), lz_ev[j & 1]);
        if (!more) break;
        if (!ahead) {
            if (nact == 0) break;
            iteration(j + 1, nact);
        } else if (g_prof) {
            g_prof->resolve(nact);
        }
        res.matvecs += nact;
        res.flop += fl * nact;
        if (nact == 0) {
            // the masked iteration j+1 queued ahead still copies its count
            // into the other host slot: let it land before the slots are
            // reused by the next exponential
            if (ahead) YQ_HIP(hipEventSynchronize(lz_ev[(j + 1) & 1]));
            break;
        }
    }
    lz_combine<T>(x, sx, V.as<CT>(), sV(), nmax, n, beta0.as<double>(), coef.as<c64>(), nit.as<int>(), beta.as<double>(),
                  kdim, nullptr,
                  nb, s);
}

// Θ = A_i · A_{i+1} -> theta
template <typename T>
void Engine<T>::merge(State &S, int i) {
    const int cl = S.wd[i], cm = S.wd[i + 1], cr = S.wd[i + 2];
    gemmT<T>(gd(2 * cl, 2 * cr, cm, OP_N, OP_N, cm, 2 * cr, 2 * cr, slotA[i], slotA[i + 1], nmax, nb), A(S, i),
             A(S, i + 1), theta.as<CT>(), s);
}

// SVD split of theta (split_mps_tensor_svd, src/Algorithms.jl:409-445) via
// the FP64 Gram matrix and a Jacobi eigensolver.  right = centre moves right
// (A_i = U, A_{i+1} = U^† Θ); else A_{i+1} = V^†, A_i = Θ V.
template <typename T>
void Engine<T>::split(State &S, int i, bool right, int rule, double thr, const int *msk, int max_keep) {
    const int cl = S.wd[i], cr = S.wd[i + 2];
    const int w_old = S.wd[i + 1];
    const int m = 2 * cl, n = 2 * cr;
    const int ns = std::min(m, n);
    const int gn = right ? m : n;
    if (g_poison_traj.load() >= 0 && g_poison_prec.load() == J.precision && g_poison_split.load() == nsplit) {
        const int64_t t = g_poison_traj.load() - (int64_t)traj0;
        if (t >= 0 && t < nb) poison_slot<T>(theta.as<CT>(), nmax, (int)t, s);
    }
    ++nsplit;
    GemmDesc gg = right ? gd(m, m, n, OP_N, OP_C, n, n, m, nmax, nmax, slotG, nb, msk)
                        : gd(n, n, m, OP_C, OP_N, n, n, n, nmax, nmax, slotG, nb, msk);
    gg.lower = heev_gram_lower_only(gn) ? 1 : 0;  // the solver reads the lower triangle only
    bool gram_mfma;
    if constexpr (std::is_same<T, float>::value)
        gram_mfma = gemm_mfma_gram(gg, theta.as<CT>(), theta.as<CT>(), G.as<c64>(), s);
    else
        gram_mfma = gemm_mfma<double>(gg, theta.as<CT>(), theta.as<CT>(), G.as<c64>(), s, 0);
    if (!gram_mfma) gemm<T, T, double, double>(gg, theta.as<CT>(), theta.as<CT>(), G.as<c64>(), s);
    guard_gram(G.as<c64>(), slotG, gn, nfail.as<int>(), msk, nb, s);
    if (msk) fill_int(keep.as<int>(), 0, nb, s);  // masked-out slots report rank 0
    const int mb = max_keep > 0 ? max_keep : (int)std::min<int64_t>(J.max_bond, 1 << 30);
    double2 *Pg = Pbuf.p ? Pbuf.as<double2>() : nullptr;
    int *fb = Hfb.p ? Hfb.as<int>() : nullptr;
    const bool multi = heev(G.as<c64>(), slotG, U.as<c64>(), slotG, Zbuf.as<double>(), slotZ, Pg, slotP, Hwork.p, slotH,
                            fb, gn, ns, evals.as<double>(), perm.as<int>(), gdim + 2, keep.as<int>(), rule, thr, mb,
                            msk, nb, s, false);
    // new working dimension = max kept rank over the wave; the kept ranks and
    // the inverse-iteration breakdown flags come back in one readback, and the
    // single-workgroup solver runs only when a matrix needs it
    std::vector<int> hk(nb), hf(multi ? nb : 0);
    int nflag = 0;
    // Kept eigenvalues below kTailCut of the largest (singular values below
    // 3e-7 of the largest, e.g. the second one the T-A rule keeps in a nearly
    // product state) lie at the rounding level of the Gram matrix, so its
    // eigenvectors there are not the singular vectors; those trajectories get
    // a second, deflated Gram below (FP64 only: complex-FP32 tensors do not
    // resolve such singular values either way)
    constexpr double kTailCut = 1e-13;
    auto flag_tails = [&]() {
        fill_int(dflag_count.as<int>(), 0, 1, s);
        tail_flags(evals.as<double>(), gdim + 2, keep.as<int>(), kTailCut, Dcnt.as<int>(), Dflag.as<int>(),
                   dflag_count.as<int>(), nb, s);
        YQ_HIP(hipMemcpyAsync(&nflag, dflag_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
    };
    if (defl) flag_tails();
    YQ_HIP(hipMemcpyAsync(hk.data(), keep.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    if (multi) YQ_HIP(hipMemcpyAsync(hf.data(), fb, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    YQ_HIP(hipStreamSynchronize(s));
    if (multi && std::any_of(hf.begin(), hf.end(), [](int f) { return f != 0; })) {
        heev_fixup(G.as<c64>(), slotG, U.as<c64>(), slotG, Zbuf.as<double>(), slotZ, Pg, slotP, fb, gn, ns,
                   evals.as<double>(), perm.as<int>(), gdim + 2, keep.as<int>(), rule, thr, mb, msk, nb, s);
        if (defl) flag_tails();
        YQ_HIP(hipMemcpyAsync(hk.data(), keep.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
    }
    int kw = 1;
    for (int v : hk) kw = std::max(kw, v);
    if (kw > cap[i + 1]) throw std::runtime_error("bond dimension exceeds capacity");
    if (msk) {
        // slots outside the mask keep their tensors: the working dimension
        // only grows, and their tensors are re-laid out for it first
        if (kw > w_old) regrow(S, i + 1, kw);
        kw = std::max(kw, w_old);
    }
    if (right) {
        write_iso<T>(U.as<c64>(), slotG, perm.as<int>(), gdim + 2, keep.as<int>(), m, kw, 0, A(S, i), slotA[i], msk,
                     nb, s);
        if (nflag > 0) deflate_tail(S, i, true, kw, mb);
        gemmT<T>(gd(kw, n, m, OP_C, OP_N, kw, n, n, slotA[i], nmax, slotA[i + 1], nb, msk), A(S, i), theta.as<CT>(),
                 A(S, i + 1), s);
    } else {
        write_iso<T>(U.as<c64>(), slotG, perm.as<int>(), gdim + 2, keep.as<int>(), n, kw, 1, A(S, i + 1), slotA[i + 1],
                     msk, nb, s);
        if (nflag > 0) deflate_tail(S, i, false, kw, mb);
        gemm<T, T, T, T>(gd(m, kw, n, OP_N, OP_C, n, n, kw, nmax, slotA[i + 1], slotA[i], nb, msk), theta.as<CT>(),
                         A(S, i + 1), A(S, i), s);
    }
    S.wd[i + 1] = kw;
    set_dims(S.dims.as<int>(), L + 1, i + 1, keep.as<int>(), 0, msk, nb, s);
}

// The kept tail of a split whose singular values the Gram route cannot
// resolve (flagged trajectories only, FP64): with Q_d the resolved kept
// columns (right = true: left vectors U_d; else right vectors V_d), the
// residual R = Θ - U_d U_d^† Θ (or Θ - Θ V_d V_d^†) holds the rest of Θ at
// its own scale, so the leading eigenvectors of R R^† (R^† R) are the tail's
// singular vectors to working accuracy; they replace the tail columns
// (rows) of the isometry just written.
template <typename T>
void Engine<T>::deflate_tail(State &S, int i, bool right, int kw, int mb) {
    const int cl = S.wd[i], cr = S.wd[i + 2];
    const int m = 2 * cl, n = 2 * cr, ns = std::min(m, n), gn = right ? m : n;
    const int *fl = Dflag.as<int>();
    CT *th = theta.as<CT>(), *q = Dq.as<CT>(), *t1 = T1.as<CT>(), *t2 = T2.as<CT>(), *r = V.as<CT>();
    write_iso<T>(U.as<c64>(), slotG, perm.as<int>(), gdim + 2, Dcnt.as<int>(), gn, kw, right ? 0 : 1, q, slotDq, fl,
                 nb, s);
    if (right) {
        gemm<T, T, T, T>(gd(kw, n, m, OP_C, OP_N, kw, n, n, slotDq, nmax, slotS, nb, fl), q, th, t1, s);  // U_d^† Θ
        gemm<T, T, T, T>(gd(m, n, kw, OP_N, OP_N, kw, n, n, slotDq, slotS, slotS, nb, fl), q, t1, t2, s);
    } else {
        gemm<T, T, T, T>(gd(m, kw, n, OP_N, OP_C, n, n, kw, nmax, slotDq, slotS, nb, fl), th, q, t1, s);  // Θ V_d
        gemm<T, T, T, T>(gd(m, n, kw, OP_N, OP_N, kw, n, n, slotS, slotDq, slotS, nb, fl), t1, q, t2, s);
    }
    sub<T>(th, nmax, t2, slotS, r, sV(), (int64_t)m * n, fl, nb, s);
    if (right)
        gemm<T, T, double, double>(gd(m, m, n, OP_N, OP_C, n, n, m, sV(), sV(), slotG, nb, fl), r, r, G.as<c64>(), s);
    else
        gemm<T, T, double, double>(gd(n, n, m, OP_C, OP_N, n, n, n, sV(), sV(), slotG, nb, fl), r, r, G.as<c64>(), s);
    double2 *Pg = Pbuf.p ? Pbuf.as<double2>() : nullptr;
    (void)heev(G.as<c64>(), slotG, U.as<c64>(), slotG, Zbuf.as<double>(), slotZ, Pg, slotP, Hwork.p, slotH,
               Hfb.p ? Hfb.as<int>() : nullptr, gn, ns, evals.as<double>(), perm.as<int>(), gdim + 2, keepR.as<int>(),
               0, 0.0, mb, fl, nb, s, true);
    if (right)
        tail_fix<T>(A(S, i), slotA[i], m, kw, 0, Dcnt.as<int>(), keep.as<int>(), U.as<c64>(), slotG, perm.as<int>(),
                    gdim + 2, fl, nb, s);
    else
        tail_fix<T>(A(S, i + 1), slotA[i + 1], n, kw, 1, Dcnt.as<int>(), keep.as<int>(), U.as<c64>(), slotG,
                    perm.as<int>(), gdim + 2, fl, nb, s);
}

// E_left[i+1] from E_left[i] and A_i (update_left_environment, src/Algorithms.jl:199-204)
template <typename T>
void Engine<T>::env_left(State &S, int i) {
    const int cl = S.wd[i], cr = S.wd[i + 1], Dl = Dm[i], Dr = Dm[i + 1];
    CT *t1 = T1.as<CT>(), *t2 = T2.as<CT>();
    const GemmDesc g1 = gd(cl * Dl, 2 * cr, cl, OP_N, OP_N, cl, 2 * cr, 2 * cr, slotE[i], slotA[i], slotS, nb);
    gemmW<T>(g1, EL[i].as<CT>(), A(S, i), cl, Dl, 2, Dr, cr, t2, slotS, csrL[i],
             [&]() {
                 gemmT<T>(g1, EL[i].as<CT>(), A(S, i), t1, s);
                 wstep<T>(t1, t2, cl, Dl * 2, 2 * Dr, cr, slotS, slotS, csrL[i].rowptr.as<int>(),
                          csrL[i].col.as<int>(), csrL[i].val.as<c64>(), nullptr, nb, s);
             },
             s);
    gemmT<T>(gd(cr, Dr * cr, 2 * cl, OP_C, OP_N, cr, Dr * cr, Dr * cr, slotA[i], slotS, slotE[i + 1], nb), A(S, i), t2,
             EL[i + 1].as<CT>(), s);
}

// E_right[i] from E_right[i+1] and A_i (update_right_environment, src/Algorithms.jl:206-211)
template <typename T>
void Engine<T>::env_right(State &S, int i) {
    const int cl = S.wd[i], cr = S.wd[i + 1], Dl = Dm[i], Dr = Dm[i + 1];
    CT *t1 = T1.as<CT>(), *t2 = T2.as<CT>();
    GemmDesc g1 = gd(2 * cl, cr, cr, OP_N, OP_N, cr, cr, (int64_t)Dr * cr, slotA[i], slotE[i + 1], slotS, nb);
    g1.inner = Dr;
    g1.sB_inner = (int64_t)cr * cr;
    g1.sC_inner = cr;
    gemmT<T>(g1, A(S, i), ER[i + 1].as<CT>(), t1, s);
    wstep<T>(t1, t2, cl, 2 * Dr, Dl * 2, cr, slotS, slotS, csrR[i].rowptr.as<int>(), csrR[i].col.as<int>(),
             csrR[i].val.as<c64>(), nullptr, nb, s);
    GemmDesc g3 = gd(cl, cl, 2 * cr, OP_N, OP_C, (int64_t)Dl * 2 * cr, 2 * cr, cl, slotS, slotA[i], slotE[i], nb);
    g3.inner = Dl;
    g3.sA_inner = 2 * cr;
    g3.sC_inner = (int64_t)cl * cl;
    gemmT<T>(g3, t2, A(S, i), ER[i].as<CT>(), s);
}

// One symmetric 2-site TDVP sweep (src/Algorithms.jl:580-603, 631-760).
template <typename T>
void Engine<T>::tdvp(State &S, JobResult &res) {
    if (J.tdvp_sites == 1) {
        tdvp1(S, res);
        return;
    }
    if (L < 2) return;
    const double dt = J.dt;
    identity_env<T>(ER[L].as<CT>(), slotE[L], S.wd[L], Dm[L], 1, nb, s);
    for (int i = L - 1; i >= 1; --i) env_right(S, i);
    identity_env<T>(EL[0].as<CT>(), slotE[0], S.wd[0], Dm[0], 0, nb, s);
    for (int i = 0; i + 2 < L; ++i) {  // forward (src/Algorithms.jl:631-672)
        merge(S, i);
        expm(S, 2, i, theta.as<CT>(), nmax, dt / 2, res);
        split(S, i, true, 0, J.thr);
        env_left(S, i);
        expm(S, 1, i + 1, A(S, i + 1), slotA[i + 1], -dt / 2, res);
    }
    {  // edge (src/Algorithms.jl:674-711)
        const int i = L - 2;
        merge(S, i);
        expm(S, 2, i, theta.as<CT>(), nmax, dt, res);
        split(S, i, false, 0, J.thr);
        env_right(S, i + 1);
    }
    for (int i = L - 3; i >= 0; --i) {  // backward (src/Algorithms.jl:713-760)
        expm(S, 1, i + 1, A(S, i + 1), slotA[i + 1], -dt / 2, res);
        merge(S, i);
        expm(S, 2, i, theta.as<CT>(), nmax, dt / 2, res);
        split(S, i, false, 0, J.thr);
        env_right(S, i + 1);
    }
}

// single_site_tdvp! (_tdvp_sweep_hamiltonian_1site!, src/Algorithms.jl:471-532)
// on a state with its centre at site 0: forward, each site evolves by dt/2,
// then (but the last) its QR moves the centre right, the R factor evolves
// backward by dt/2 under the bond Hamiltonian and is absorbed into the next
// site; backward the same with LQ, from site L-1 (which so takes dt in two
// halves) to site 0.  Bond dimensions do not change; the gauge steps drop
// dead directions as everywhere in the engine.  The bond matrices live in
// the theta slots (free in this sweep).
template <typename T>
void Engine<T>::tdvp1(State &S, JobResult &res) {
    const double dt = J.dt;
    const int64_t cap2 = (int64_t)capmax * capmax;
    CT *cb = theta.as<CT>();
    identity_env<T>(ER[L].as<CT>(), slotE[L], S.wd[L], Dm[L], 1, nb, s);
    for (int i = L - 1; i >= 1; --i) env_right(S, i);
    identity_env<T>(EL[0].as<CT>(), slotE[0], S.wd[0], Dm[0], 0, nb, s);
    for (int i = 0; i < L; ++i) {  // forward (src/Algorithms.jl:480-505)
        expm(S, 1, i, A(S, i), slotA[i], dt / 2, res);
        if (i + 1 == L) break;
        const int c0 = S.wd[i], c1 = S.wd[i + 1], c2 = S.wd[i + 2], cp = (c1 + 1) & ~1;
        // A_i = Q R (CGS2 on columns), R[j][k] = Cf[k][j]
        cgs2<T>(A(S, i), slotA[i], c1, 2 * c0, 1, c1, Cf.as<CT>(), cap2, cgsS.as<CT>(), slotCgs, cgsF.as<double>(),
                nullptr, nb, s);
        copy_bond<T>(Cf.as<CT>(), cap2, c1, true, cb, nmax, cp, c1, nullptr, nb, s);
        env_left(S, i);
        expm(S, 0, i + 1, cb, nmax, -dt / 2, res);
        gemmT<T>(gd(c1, 2 * c2, c1, OP_N, OP_N, cp, 2 * c2, 2 * c2, nmax, slotA[i + 1], slotA[i + 1], nb), cb,
                 A(S, i + 1), absorb.as<CT>(), s);
        copy2d<T, T>(absorb.as<CT>(), slotA[i + 1], 2 * c2, A(S, i + 1), slotA[i + 1], 2 * c2, c1, 2 * c2, nullptr, nb,
                     s);
    }
    for (int i = L - 1; i >= 0; --i) {  // backward (src/Algorithms.jl:507-531)
        expm(S, 1, i, A(S, i), slotA[i], dt / 2, res);
        if (i == 0) break;
        const int cm = S.wd[i - 1], c0 = S.wd[i], c1 = S.wd[i + 1], cp = (c0 + 1) & ~1;
        // A_i = L Q (CGS2 on rows), L = Cf
        cgs2<T>(A(S, i), slotA[i], c0, 2 * c1, 2 * c1, 1, Cf.as<CT>(), cap2, cgsS.as<CT>(), slotCgs, cgsF.as<double>(),
                nullptr, nb, s);
        copy_bond<T>(Cf.as<CT>(), cap2, c0, false, cb, nmax, cp, c0, nullptr, nb, s);
        env_right(S, i);
        expm(S, 0, i, cb, nmax, -dt / 2, res);
        gemmT<T>(gd(2 * cm, c0, c0, OP_N, OP_N, c0, cp, c0, slotA[i - 1], nmax, slotA[i - 1], nb), A(S, i - 1), cb,
                 absorb.as<CT>(), s);
        copy2d<T, T>(absorb.as<CT>(), slotA[i - 1], c0, A(S, i - 1), slotA[i - 1], c0, 2 * cm, c0, nullptr, nb, s);
    }
}

// A_{i-1} <- A_{i-1} · Cf (absorb the L factor of an LQ)
template <typename T>
void Engine<T>::absorb_left(State &S, int i, const int *msk) {
    const int c0 = S.wd[i - 1], c1 = S.wd[i];
    gemmT<T>(gd(2 * c0, c1, c1, OP_N, OP_N, c1, c1, c1, slotA[i - 1], (int64_t)capmax * capmax, slotA[i - 1], nb, msk),
             A(S, i - 1), Cf.as<CT>(), absorb.as<CT>(), s);
    copy2d<T, T>(absorb.as<CT>(), slotA[i - 1], 2 * c1, A(S, i - 1), slotA[i - 1], 2 * c1, c0, 2 * c1, msk, nb, s);
}

// Right-canonicalise sites from..1 by LQ (CGS2 on rows), leaving the norm in
// site 0 (normalize! without the final scaling, src/MPS.jl:274-311).
template <typename T>
void Engine<T>::lq_sweep(State &S, int from, const int *msk) {
    for (int i = from; i >= 1; --i) {
        const int c0 = S.wd[i], c1 = S.wd[i + 1];
        cgs2<T>(A(S, i), slotA[i], c0, 2 * c1, 2 * c1, 1, Cf.as<CT>(), (int64_t)capmax * capmax, cgsS.as<CT>(), slotCgs,
                cgsF.as<double>(), msk, nb, s);
        absorb_left(S, i, msk);
    }
}

// Move the centre i -> i+1 by QR (CGS2 on columns) of A_i.
template <typename T>
void Engine<T>::qr_step(State &S, int i, const int *msk) {
    const int c0 = S.wd[i], c1 = S.wd[i + 1], c2 = S.wd[i + 2];
    cgs2<T>(A(S, i), slotA[i], c1, 2 * c0, 1, c1, Cf.as<CT>(), (int64_t)capmax * capmax, cgsS.as<CT>(), slotCgs,
            cgsF.as<double>(), msk, nb, s);
    // A_{i+1} <- R A_{i+1} with R[j][k] = Cf[k][j]
    gemmT<T>(gd(c1, 2 * c2, c1, OP_T, OP_N, c1, 2 * c2, 2 * c2, (int64_t)capmax * capmax, slotA[i + 1], slotA[i + 1],
                nb, msk),
             Cf.as<CT>(), A(S, i + 1), absorb.as<CT>(), s);
    copy2d<T, T>(absorb.as<CT>(), slotA[i + 1], 2 * c2, A(S, i + 1), slotA[i + 1], 2 * c2, c1, 2 * c2, msk, nb, s);
}

// apply_dissipation (src/Dissipation.jl:37-159).  Without adjacent 2-site
// non-Pauli processes every operator is local and exact in any gauge, so the
// per-site 2x2 products are applied in one pass; with them, the reference's
// sweep runs (dissipate_sweep).  field: the reference's centre field on entry
// (0 after a TDVP sweep; the last observable's site right after the t=0
// measurement), where the engine's state has its centre at 0.
template <typename T>
void Engine<T>::dissipate(State &S, bool half, int field) {
    scr_ready = false;
    if (has2) {
        dissipate_sweep(S, half, field);
        return;
    }
    const DevBuf &ops = half ? diss_half : diss_full;
    for (int i = 0; i < L; ++i)
        if (diss_site_active[i])
            local_op<T>(A(S, i), slotA[i], S.wd[i], S.wd[i + 1], ops.as<c64>() + 4 * (int64_t)i, nullptr, nullptr, nb, s);
}

// The reference's dissipation sweep, site i = L..1: the 1-site operators of
// site i, then each 2-site non-Pauli process with upper site i in process
// order (merge (i-1, i), apply exp(-dt/2 γ L†L), split with the relative
// rule T-C: U on i-1, S V† on i), then the centre moves to i-1 (from the
// field by QR when i = L, by LQ of site i after that).  Pauli-named 2-site
// processes are the scalar already folded into site i's operator.  At the
// end ψ~ (site L replaced by the isometry of its LQ) is kept in SCR for the
// jump weights (see stochastic).
template <typename T>
void Engine<T>::dissipate_sweep(State &S, bool half, int field) {
    const DevBuf &ops = half ? diss_half : diss_full;
    const DevBuf &ops4 = half ? diss4_half : diss4_full;
    for (int k = 0; k < field; ++k) qr_step(S, k, nullptr);  // the engine's centre to the field
    for (int i = L - 1; i >= 0; --i) {
        if (diss_site_active[i])
            local_op<T>(A(S, i), slotA[i], S.wd[i], S.wd[i + 1], ops.as<c64>() + 4 * (int64_t)i, nullptr, nullptr, nb, s);
        if (i == 0) break;
        for (int q : d4_at[i]) {
            merge(S, i - 1);
            local_op4<T>(theta.as<CT>(), nmax, S.wd[i - 1], S.wd[i + 1], ops4.as<c64>() + 16 * (int64_t)q, nb, s);
            split(S, i - 1, true, 1, J.thr);
        }
        if (i == L - 1) {
            if (field == L - 1) lq_step(S, i, nullptr);
            else
                for (int k = field; k + 1 < L - 1; ++k) qr_step(S, k, nullptr);
        } else {
            lq_step(S, i, nullptr);
        }
    }
    // ψ~ for the weights: the reference reads the centre norms with site L
    // as it stands (not re-orthogonalised when the field started below L)
    if (field != L - 1 && L >= 2) {
        copy_state(S, SCR);
        lq_step(SCR, L - 1, nullptr, false);
        scr_ready = true;
    }
}

// LQ of site i (CGS2 on its rows): A_i <- Q and, with keep_factor, the L
// factor absorbed into site i-1 (the centre moves to i-1); without, the
// factor is dropped.
template <typename T>
void Engine<T>::lq_step(State &S, int i, const int *msk, bool keep_factor) {
    const int c0 = S.wd[i], c1 = S.wd[i + 1];
    cgs2<T>(A(S, i), slotA[i], c0, 2 * c1, 2 * c1, 1, Cf.as<CT>(), (int64_t)capmax * capmax, cgsS.as<CT>(), slotCgs,
            cgsF.as<double>(), msk, nb, s);
    if (keep_factor) absorb_left(S, i, msk);
}

// Re-lays out every slot's A_{b-1} (columns) and A_b (rows) for a larger
// working dimension w at bond b, with zeros in the new columns / rows.
template <typename T>
void Engine<T>::regrow(State &S, int b, int w) {
    const int wo = S.wd[b];
    if (w <= wo) return;
    const int r0 = 2 * S.wd[b - 1];          // rows of A_{b-1} ([a][s] flattened)
    fill_zero<T>(absorb.as<CT>(), slotA[b - 1], (int64_t)r0 * w, nb, s);
    copy2d<T, T>(A(S, b - 1), slotA[b - 1], wo, absorb.as<CT>(), slotA[b - 1], w, r0, wo, nullptr, nb, s);
    copy2d<T, T>(absorb.as<CT>(), slotA[b - 1], w, A(S, b - 1), slotA[b - 1], w, r0, w, nullptr, nb, s);
    const int64_t row = 2 * (int64_t)S.wd[b + 1];  // A_b is [a][s][c]: one bond row = 2 wd[b+1] elements
    fill_zero<T>(A(S, b) + (int64_t)wo * row, slotA[b], (int64_t)(w - wo) * row, nb, s);
    S.wd[b] = w;
}

// The working dimension of every bond back down to the largest true
// dimension of the wave's trajectories (S.dims): A_{b-1} re-laid out with the
// smaller row length, A_b keeps its leading rows (the dropped columns and
// rows are zero padding in every slot).
template <typename T>
void Engine<T>::shrink_to_dims(State &S) {
    std::vector<int> hd((size_t)nb * (L + 1));
    YQ_HIP(hipMemcpyAsync(hd.data(), S.dims.p, hd.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    YQ_HIP(hipStreamSynchronize(s));
    for (int b = 1; b < L; ++b) {
        int w = 1;
        for (int t = 0; t < nb; ++t) w = std::max(w, hd[(size_t)t * (L + 1) + b]);
        const int wo = S.wd[b];
        if (w >= wo) continue;
        const int r0 = 2 * S.wd[b - 1];
        copy2d<T, T>(A(S, b - 1), slotA[b - 1], wo, absorb.as<CT>(), slotA[b - 1], w, r0, w, nullptr, nb, s);
        copy2d<T, T>(absorb.as<CT>(), slotA[b - 1], w, A(S, b - 1), slotA[b - 1], w, r0, w, nullptr, nb, s);
        S.wd[b] = w;
    }
}

// Left transfer sweep in FP64 giving every site's reduced matrix ρ_s (and
// 4x4 ρ for 2-site observables) of a right-canonical state.
template <typename T>
void Engine<T>::transfer(State &S, const int *msk, c64 *rho_out, c64 *rho2_out, const std::vector<char> *at2) {
    if (!rho2_out) rho2_out = rho2.as<c64>();
    if (!at2) at2 = &obs2_at;
    c64 *E0 = Ebuf0.as<c64>(), *E1 = Ebuf1.as<c64>();
    const int64_t sE = (int64_t)capmax * capmax;
    identity_env<double>(E0, sE, S.wd[0], 1, 0, nb, s);
    for (int i = 0; i < L; ++i) {
        const int cl = S.wd[i], cr = S.wd[i + 1];
        gemm<double, T, double, double>(gd(cl, 2 * cr, cl, OP_N, OP_N, cl, 2 * cr, 2 * cr, sE, slotA[i], slotX, nb, msk),
                                        E0, A(S, i), Xbuf.as<c64>(), s);
        rho<T>(Xbuf.as<c64>(), slotX, A(S, i), slotA[i], cl, 2, cr, rho_out + (int64_t)i * 4, (int64_t)L * 4, msk, nb,
               s);
        if ((*at2)[i] && i + 1 < L) {
            merge(S, i);
            const int c2 = S.wd[i + 2];
            gemm<double, T, double, double>(gd(cl, 4 * c2, cl, OP_N, OP_N, cl, 4 * c2, 4 * c2, sE, nmax, slotX, nb, msk),
                                            E0, theta.as<CT>(), Xbuf.as<c64>(), s);
            rho<T>(Xbuf.as<c64>(), slotX, theta.as<CT>(), nmax, cl, 4, c2, rho2_out + (int64_t)i * 16,
                   (int64_t)L * 16, msk, nb, s);
            gemm<double, T, double, double>(gd(cl, 2 * cr, cl, OP_N, OP_N, cl, 2 * cr, 2 * cr, sE, slotA[i], slotX, nb,
                                               msk),
                                            E0, A(S, i), Xbuf.as<c64>(), s);
        }
        gemm<T, double, double, double>(gd(cr, cr, 2 * cl, OP_C, OP_N, cr, cr, cr, slotA[i], slotX, sE, nb, msk), A(S, i),
                                        Xbuf.as<c64>(), E1, s);
        std::swap(E0, E1);
    }
}

// <ψ| opa_a opb_b |ψ> per masked trajectory (2x2 operators in engine
// order on the ket side), by a left transfer sweep in FP64 over the whole
// chain, which needs no canonical form (the last site closes a bond of
// dimension 1); out[t] for the masked slots.
template <typename T>
void Engine<T>::correlate(State &S, int a, int b, const c64 *opa, const c64 *opb, const int *msk, c64 *out) {
    c64 *E0 = Ebuf0.as<c64>(), *E1 = Ebuf1.as<c64>();
    const int64_t sE = (int64_t)capmax * capmax;
    identity_env<double>(E0, sE, S.wd[0], 1, 0, nb, s);
    for (int i = 0; i < L; ++i) {
        const int cl = S.wd[i], cr = S.wd[i + 1];
        gemm<double, T, double, double>(gd(cl, 2 * cr, cl, OP_N, OP_N, cl, 2 * cr, 2 * cr, sE, slotA[i], slotX, nb, msk),
                                        E0, A(S, i), Xbuf.as<c64>(), s);
        if (i == a) local_op<double>(Xbuf.as<c64>(), slotX, cl, cr, opa, nullptr, msk, nb, s);
        if (i == b) local_op<double>(Xbuf.as<c64>(), slotX, cl, cr, opb, nullptr, msk, nb, s);
        gemm<T, double, double, double>(gd(cr, cr, 2 * cl, OP_C, OP_N, cr, cr, cr, slotA[i], slotX, sE, nb, msk), A(S, i),
                                        Xbuf.as<c64>(), E1, s);
        std::swap(E0, E1);
    }
    copy2d<double, double>(E0, sE, 1, out, 1, 1, 1, 1, msk, nb, s);
}

// apply_mpo_jump! (src/StochasticProcess.jl:46-82) for the masked
// trajectories, long-range candidate li (O = α I + β σ_a τ_b):
// 1. the bonds a+1..b double (regrow) and the site tensors take the blocks of
//    build_mpo_phys's contraction (src/Noise.jl:77-128): site a [A, σA],
//    sites between diag(A, A), site b [α A; β τA];
// 2. truncate! (src/MPS.jl:320-441): its centre shift to site 1, an LQ
//    sweep (CGS2 over every row: a bond keeps its doubled working dimension,
//    dependent rows become zero rows), then its left-to-right sweep with the
//    T-A rule, the job's threshold and bond cap (U on the left site, S V^† on
//    the right), then normalize!.
template <typename T>
void Engine<T>::mpo_jump(State &S, int li, const int *msk) {
    const ProcIn &p = J.noise[lr_cands[li].proc];
    const int a = std::min(p.s0, p.s1), b = std::max(p.s0, p.s1);
    bool fits = true;
    for (int k = a + 1; k <= b; ++k) fits = fits && 2 * S.wd[k] <= cap[k];
    if (!fits) shrink_to_dims(S);  // an earlier jump of this step left doubled working dimensions
    const std::vector<int> w0 = S.wd;
    for (int k = a + 1; k <= b; ++k) {
        if (2 * w0[k] > cap[k]) throw std::runtime_error("long-range jump: doubled bond exceeds its capacity");
        regrow(S, k, 2 * w0[k]);
    }
    const c64 *ops = lr_ops.as<c64>() + (int64_t)li * 16;
    const cd al = toc(p.alpha), be = toc(p.beta);
    mpo_block<T>(A(S, a), slotA[a], S.wd[a + 1], 0, S.wd[a], w0[a + 1], ops, 0, 0, 0, 0, msk, nb, s);
    for (int k = a + 1; k < b; ++k)
        mpo_block<T>(A(S, k), slotA[k], S.wd[k + 1], 1, w0[k], w0[k + 1], ops, 0, 0, 0, 0, msk, nb, s);
    mpo_block<T>(A(S, b), slotA[b], S.wd[b + 1], 2, w0[b], S.wd[b + 1], ops + 4, al.real(), al.imag(), be.real(),
                 be.imag(), msk, nb, s);
    lq_sweep(S, L - 1, msk);
    for (int k = 0; k + 1 < L; ++k) {
        merge(S, k);
        split(S, k, true, 0, J.thr, msk);
    }
    lq_sweep(S, L - 1, msk);
    norm2<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], norm2v.as<double>(), msk, nb, s);
    jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(), fac.as<double>(), nb, s);
    scale<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], fac.as<double>(), msk, nb, s);
}

template <typename T>
void Engine<T>::measure(State &S, int col, int ncol, double *d_out) {
    if (J.obs.empty()) return;
    transfer(S, nullptr, rho1.as<c64>());
    observe(rho1.as<c64>(), (int64_t)L * 4, rho2.as<c64>(), (int64_t)L * 16, (int)J.obs.size(), obs_site.as<int>(),
            obs_n.as<int>(), obs_op.as<c64>(), col, ncol, J.num_traj, (int64_t)traj0 - (int64_t)J.traj_offset, d_out,
            nb, s);
}

// stochastic_process! (src/StochasticProcess.jl:219-330) for a whole wave.
//
// Jump weights follow the reference exactly, including one property of its
// gauge bookkeeping: apply_dissipation (src/Dissipation.jl:37-159) enters
// with the orthogonality-centre field at site 1 after a TDVP sweep, so its
// first shift (to L-1) runs QR from the left and never re-orthogonalises site
// L, which keeps its damping factor D_L.  create_probability_distribution
// then reads ‖L_k A_s‖² at the centre (src/StochasticProcess.jl:100-131),
// which for s < L equals <ψ~|L_k†L_k|ψ~> with ψ~ = D_L^{-1} ψ; only site L
// sees the true state.  `quirk` is false when the field was at L on entry
// (then site L is re-orthogonalised by the first LQ), which happens in
// initialize() right after the t=0 measurement of an observable on site L.
// invL: inverse of the last site's dissipation operator just applied, or
// nullptr when the weights see the true state everywhere.
template <typename T>
void Engine<T>::stochastic(State &S, double dt, const DevBuf *invL, JobResult &res) {
    lq_sweep(S, L - 1, nullptr);
    norm2<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], norm2v.as<double>(), nullptr, nb, s);
    jump_decide(norm2v.as<double>(), J.seed, traj0, rngk.as<uint64_t>(), J.noise.empty() ? 0 : 1, jstate.as<int>(),
                fac.as<double>(), nb, s);
    mask_from_state(jstate.as<int>(), 0, mask.as<int>(), nullptr, nb, s);
    scale<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], fac.as<double>(), mask.as<int>(), nb, s);
    fill_int(count.as<int>(), 0, 1, s);
    mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
    const int njump = sync_count(count.as<int>());
    if (njump == 0) return;
    transfer(S, mask.as<int>(), rho1.as<c64>(), rho2.as<c64>(), &cand2_at);
    if (has2) {
        // ψ~ of the dissipation sweep (site L replaced by its LQ isometry):
        // 1-site weights of sites < L and merged-pair weights of pairs
        // below (L-1, L) read it; the others read the true state
        if (scr_ready) {
            lq_sweep(SCR, L - 2, mask.as<int>());
            transfer(SCR, mask.as<int>(), rho1b.as<c64>(), rho2b.as<c64>(), &cand2_at);
            copy2d<double, double>(rho1b.as<c64>(), (int64_t)L * 4, 4 * (L - 1), rho1.as<c64>(), (int64_t)L * 4,
                                   4 * (L - 1), 1, 4 * (L - 1), mask.as<int>(), nb, s);
            if (L >= 3)
                copy2d<double, double>(rho2b.as<c64>(), (int64_t)L * 16, 16 * (L - 2), rho2.as<c64>(),
                                       (int64_t)L * 16, 16 * (L - 2), 1, 16 * (L - 2), mask.as<int>(), nb, s);
        }
    } else if (invL && last_site_damped) {
        copy_state(S, SCR);
        local_op<T>(A(SCR, L - 1), slotA[L - 1], SCR.wd[L - 1], SCR.wd[L], invL->as<c64>(), nullptr, mask.as<int>(), nb,
                    s);
        lq_sweep(SCR, L - 1, mask.as<int>());
        transfer(SCR, mask.as<int>(), rho1b.as<c64>());
        copy2d<double, double>(rho1b.as<c64>(), (int64_t)L * 4, 4 * (L - 1), rho1.as<c64>(), (int64_t)L * 4,
                               4 * (L - 1), 1, 4 * (L - 1), mask.as<int>(), nb, s);
    }
    // general long-range candidates: ν from two correlators of the true
    // (unnormalised) state, which S holds for the jumping trajectories
    for (size_t li = 0; li < lr_cands.size(); ++li) {
        const ProcIn &p = J.noise[lr_cands[li].proc];
        const int a = std::min(p.s0, p.s1), b = std::max(p.s0, p.s1);
        const c64 *ops = lr_ops.as<c64>() + (int64_t)li * 16;
        correlate(S, a, b, ops, ops + 4, mask.as<int>(), lr_val.as<c64>());
        correlate(S, a, b, ops + 8, ops + 12, mask.as<int>(), lr_val.as<c64>() + nb);
        lr_nu(lr_val.as<c64>(), rho1.as<c64>(), (int64_t)L * 4, L - 1, p.alpha.re, p.alpha.im, p.beta.re, p.beta.im,
              lr_cands[li].cand, ncand, jstate.as<int>(), lrnu.as<double>(), nb, s);
    }
    jump_choose(rho1.as<c64>(), (int64_t)L * 4, L, ncand, cand_site.as<int>(), cand_LdL.as<c64>(),
                cand_gamma.as<double>(), dt, J.seed, traj0, rngk.as<uint64_t>(), jstate.as<int>(), chosen.as<int>(),
                wbuf.as<double>(), nb, s, rho2.as<c64>(), (int64_t)L * 16, cand_LdL4.as<c64>(), cand_kind.as<int>(),
                lrnu.p ? lrnu.as<double>() : nullptr);
    if (std::getenv("YAQS_DEBUG_JUMPS")) {
        std::vector<int> hs(nb), hc(nb);
        std::vector<uint64_t> hk(nb);
        YQ_HIP(hipMemcpyAsync(hs.data(), jstate.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipMemcpyAsync(hc.data(), chosen.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipMemcpyAsync(hk.data(), rngk.p, nb * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
        std::vector<double> hw((size_t)nb * std::max(1, ncand));
        YQ_HIP(hipMemcpy(hw.data(), wbuf.p, hw.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int t = 0; t < nb; ++t)
            if (hs[t] == 1) {
                double tot = 0;
                for (int c = 0; c < ncand; ++c) tot += hw[(size_t)t * ncand + c];
                fprintf(stderr, "gpu jump traj=%llu k=%llu cand=%d probs:", (unsigned long long)(traj0 + t),
                        (unsigned long long)hk[t], hc[t]);
                for (int c = 0; c < ncand; ++c) fprintf(stderr, " %.15g", hw[(size_t)t * ncand + c] / tot);
                fprintf(stderr, "\n");
            }
    }
    for (int i = 0; i < L; ++i) {
        jump_opidx(jstate.as<int>(), chosen.as<int>(), cand_site.as<int>(), cand_op.as<int>(), cand_site2.as<int>(),
                   cand_op2.as<int>(), i, opidx.as<int>(), nb, s);
        local_op<T>(A(S, i), slotA[i], S.wd[i], S.wd[i + 1], jump_ops.as<c64>(), opidx.as<int>(), nullptr, nb, s);
    }
    // merged-pair jumps (local_2, src/StochasticProcess.jl:266-309): the
    // centre walks right by QR for the trajectories whose pair lies further
    // on; at its pair each merges it, applies the 4x4 operator and splits
    // with the relative rule T-C (U on the first site, S V† on the second)
    if (std::any_of(pair_cand_site.begin(), pair_cand_site.end(), [](char c) { return c != 0; })) {
        std::vector<int> hs(nb), hc(nb);
        YQ_HIP(hipMemcpyAsync(hs.data(), jstate.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipMemcpyAsync(hc.data(), chosen.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
        std::vector<char> at(L, 0);
        int pmax = -1;
        for (int t = 0; t < nb; ++t)
            if (hs[t] == 1 && hc[t] >= 0 && hc[t] < ncand && h_cand_pair[hc[t]] >= 0) {
                const int p = h_cand_site[hc[t]];
                at[p] = 1;
                pmax = std::max(pmax, p);
            }
        for (int p = 0; p <= pmax; ++p) {
            if (at[p]) {
                jump_pair_mask(jstate.as<int>(), chosen.as<int>(), cand_site.as<int>(), cand_pair.as<int>(), p, 0,
                               mask2.as<int>(), opidx4.as<int>(), nb, s);
                merge(S, p);
                local_op4_sel<T>(theta.as<CT>(), nmax, S.wd[p], S.wd[p + 2], pair_ops.as<c64>(), opidx4.as<int>(), nb,
                                 s);
                split(S, p, true, 1, J.thr, mask2.as<int>());
            }
            if (p < pmax) {
                jump_pair_mask(jstate.as<int>(), chosen.as<int>(), cand_site.as<int>(), cand_pair.as<int>(), p, 1,
                               mask2.as<int>(), nullptr, nb, s);
                qr_step(S, p, mask2.as<int>());
            }
        }
    }
    // long-range MPO jumps (apply_mpo_jump!), one candidate at a time
    if (!lr_cands.empty()) {
        std::vector<int> hs(nb), hc(nb);
        YQ_HIP(hipMemcpyAsync(hs.data(), jstate.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipMemcpyAsync(hc.data(), chosen.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
        for (size_t li = 0; li < lr_cands.size(); ++li) {
            bool any = false;
            for (int t = 0; t < nb && !any; ++t) any = hs[t] == 1 && hc[t] == lr_cands[li].cand;
            if (!any) continue;
            mask_chosen(jstate.as<int>(), chosen.as<int>(), lr_cands[li].cand, mask2.as<int>(), nb, s);
            mpo_jump(S, (int)li, mask2.as<int>());
        }
    }
    fill_int(count.as<int>(), 0, 1, s);
    mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
    res.jumps += sync_count(count.as<int>());
    mask_from_state(jstate.as<int>(), -1, mask.as<int>(), nullptr, nb, s);
    lq_sweep(S, L - 1, mask.as<int>());
    norm2<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], norm2v.as<double>(), mask.as<int>(), nb, s);
    jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(), fac.as<double>(), nb, s);
    scale<T>(A(S, 0), slotA[0], (int64_t)S.wd[0] * 2 * S.wd[1], fac.as<double>(), mask.as<int>(), nb, s);
}

// analog_tjm_2 / analog_tjm_1 (src/AnalogTJM.jl:104-176) for one wave.
template <typename T>
void Engine<T>::run_analog(int64_t t0, int nb_, double *d_out, JobResult &res) {
    nb = nb_;
    traj0 = J.traj_offset + (uint64_t)t0;
    fill_int(nfail.as<int>(), 0, B, s);
    nsplit = 0;
    YQ_HIP(hipMemsetAsync(rngk.p, 0, (size_t)B * sizeof(uint64_t), s));
    init_state(PHI);
    const int nt = (int)J.num_times;
    const int ncol = J.sample_timesteps ? nt : 1;
    if (J.sample_timesteps) measure(PHI, 0, ncol, d_out);
    const bool noise = J.has_noise;
    auto record_bond = [&](State &S) {
        for (int b = 0; b <= L; ++b) res.max_bond = std::max<int64_t>(res.max_bond, S.wd[b]);
    };
    if (J.order == 1) {
        for (int j = 2; j <= nt; ++j) {
            tdvp(PHI, res);
            if (noise) {
                dissipate(PHI, false);
                stochastic(PHI, J.dt, &invL_full, res);
            }
            record_bond(PHI);
            if (J.sample_timesteps) measure(PHI, j - 1, ncol, d_out);
            else if (j == nt) measure(PHI, 0, ncol, d_out);
        }
        return;
    }
    if (noise) {  // initialize (src/AnalogTJM.jl:21-33)
        // centre field after the t=0 measurement: site of the last observable
        int field = 0;
        if (J.sample_timesteps && !J.obs.empty()) field = std::min(J.obs.back().s0, J.obs.back().s1);
        dissipate(PHI, true, field);
        stochastic(PHI, J.dt, field == L - 1 ? nullptr : &invL_half, res);
    }
    auto sample = [&](int j) {  // src/AnalogTJM.jl:62-81
        copy_state(PHI, PSI);
        tdvp(PSI, res);
        if (noise) {
            dissipate(PSI, true);
            stochastic(PSI, J.dt, &invL_half, res);
        }
        record_bond(PSI);
        measure(PSI, J.sample_timesteps ? j - 1 : 0, ncol, d_out);
    };
    // measurement window: steps j > bench_warmup + 1 (one step = one time index)
    GemmProf prof;
    double t_win = 0, flop0 = 0;
    int64_t mv0 = 0;
    auto window = [&](int j) {
        if (J.bench_warmup < 0) return;
        if (j == J.bench_warmup + 1) {
            YQ_HIP(hipStreamSynchronize(s));
            t_win = J.wsync ? J.wsync->arrive() : host_now();
            flop0 = res.flop;
            mv0 = res.matvecs;
            g_prof = &prof;
            prof.start(s);
        }
        if (j == nt && j > J.bench_warmup + 1) {
            YQ_HIP(hipStreamSynchronize(s));
            res.timed_s += host_now() - t_win;
            res.timed_steps += nt - 1 - J.bench_warmup;
            prof.stop();
            g_prof = nullptr;
            res.gemm_ms += prof.ms;
            res.gemm_flop += prof.flop;
            res.gemm_launches += prof.n;
            res.win_flop += res.flop - flop0;
            res.win_matvecs += res.matvecs - mv0;
        }
    };
    static const bool progress = std::getenv("YAQS_PROGRESS") != nullptr;
    const double t_run = host_now();
    auto report = [&](int j) {
        if (!progress) return;
        YQ_HIP(hipStreamSynchronize(s));
        int mb = 0;
        for (int b = 0; b <= L; ++b) mb = std::max(mb, PHI.wd[b]);
        fprintf(stderr, "[yaqs] traj %lld+%d step %d/%d  %.2fs  max bond %d  matvecs %lld\n", (long long)t0, nb,
                j - 1, nt - 1, host_now() - t_run, mb, (long long)res.matvecs);
        fflush(stderr);
    };
    window(1);
    if (J.sample_timesteps && nt > 1) sample(2);
    window(2);
    report(2);
    for (int j = 3; j <= nt; ++j) {
        tdvp(PHI, res);  // step_through (src/AnalogTJM.jl:40-55)
        if (noise) {
            dissipate(PHI, false);
            stochastic(PHI, J.dt, &invL_full, res);
        }
        if (J.sample_timesteps || j == nt) sample(j);
        window(j);
        report(j);
    }
    g_prof = nullptr;
}

// run_digital_tjm with TEBD gates (src/DigitalTJM.jl:363-450) for one wave.
template <typename T>
void Engine<T>::run_digital(int64_t t0, int nb_, double *d_out, int64_t *h_bonds, JobResult &res) {
    nb = nb_;
    traj0 = J.traj_offset + (uint64_t)t0;
    fill_int(nfail.as<int>(), 0, B, s);
    nsplit = 0;
    YQ_HIP(hipMemsetAsync(rngk.p, 0, (size_t)B * sizeof(uint64_t), s));
    init_state(PHI);
    // greedy layering (src/DigitalTJM.jl:86-145)
    std::vector<std::vector<int>> layers;
    std::vector<int> barrier_layers;
    {
        std::vector<int> cur, busy;
        for (size_t g = 0; g < J.gates.size(); ++g) {
            const GateIn &G2 = J.gates[g];
            if (G2.kind == 3) {
                if (!cur.empty()) { layers.push_back(cur); cur.clear(); busy.clear(); }
                barrier_layers.push_back((int)layers.size());
                continue;
            }
            std::vector<int> sites = {G2.s0};
            if (G2.kind == 2) sites.push_back(G2.s1);
            bool ov = false;
            for (int q : sites) ov = ov || std::find(busy.begin(), busy.end(), q) != busy.end();
            if (ov) { layers.push_back(cur); cur.clear(); busy.clear(); }
            cur.push_back((int)g);
            for (int q : sites) busy.push_back(q);
        }
        if (!cur.empty()) layers.push_back(cur);
    }
    static const bool progress = std::getenv("YAQS_PROGRESS") != nullptr;
    const double t_run = host_now();
    bool at_start = false;
    std::vector<int> after;
    for (int b : barrier_layers) {
        if (b == 0) at_start = true;
        else if (std::find(after.begin(), after.end(), b) == after.end()) after.push_back(b);
    }
    std::sort(after.begin(), after.end());
    int nsamp = (at_start ? 1 : 0) + (int)after.size();
    if (nsamp == 0) nsamp = 1;
    int centre = 0;
    int idx = 0;
    std::vector<int> hd((size_t)nb * (L + 1));
    auto move_centre = [&](int target) {
        while (centre < target) { qr_step(PHI, centre, nullptr); ++centre; }
        if (centre > target) { lq_sweep(PHI, centre, nullptr); centre = 0; }
        while (centre < target) { qr_step(PHI, centre, nullptr); ++centre; }
    };
    auto sample = [&](int col) {
        move_centre(0);
        measure(PHI, col, nsamp, d_out);
        YQ_HIP(hipMemcpyAsync(hd.data(), PHI.dims.p, hd.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        YQ_HIP(hipStreamSynchronize(s));
        for (int t = 0; t < nb; ++t) {
            int m = 0;
            for (int b = 0; b <= L; ++b) m = std::max(m, hd[(size_t)t * (L + 1) + b]);
            h_bonds[(size_t)(t0 + t) * nsamp + col] = m;
            res.max_bond = std::max<int64_t>(res.max_bond, m);
        }
    };
    if (at_start) sample(idx++);
    const c64 *gops = gate_ops.as<c64>();
    for (size_t li = 0; li < layers.size(); ++li) {
        for (int g : layers[li])
            if (J.gates[g].kind == 1) {
                const int q = J.gates[g].s0;
                local_op<T>(A(PHI, q), slotA[q], PHI.wd[q], PHI.wd[q + 1], gops + gate_off[g], nullptr, nullptr, nb, s);
            }
        for (int g : layers[li]) {
            const GateIn &G2 = J.gates[g];
            if (G2.kind != 2) continue;
            const int a = std::min(G2.s0, G2.s1), b = std::max(G2.s0, G2.s1);
            if (b > a + 1) {
                // long-range gate, TEBD method: SWAP network (src/DigitalTJM.jl:314-330)
                // moves site b next to a, the gate acts on (a, a+1), the SWAPs unwind;
                // every step is apply_local_gate_exact! (rule T-D).  TDVP method: the
                // window sweep below.
                auto nn = [&](int k, const c64 *op) {
                    move_centre(k);
                    merge(PHI, k);
                    local_op4<T>(theta.as<CT>(), nmax, PHI.wd[k], PHI.wd[k + 2], op, nb, s);
                    split(PHI, k, true, 2, J.thr);
                    centre = k + 1;
                };
                if (G2.win_long) {
                    // long_range_method :TDVP: the circuit sweep over the window a..b
                    // (src/DigitalTJM.jl:335-358, src/Algorithms.jl:606-672): forward
                    // 2-site steps by +1 with the 1-site back-evolution by -1, then
                    // the edge step by +1 split right; identity environments at the
                    // window's outer bonds (src/Algorithms.jl:801-818)
                    std::vector<CsrDev> &wv = win_csr.at(g);
                    const int n = b - a + 1;
                    auto swap_tables = [&]() {
                        for (int q = 0; q < n; ++q) {
                            std::swap(csrL[a + q], wv[q]);
                            std::swap(csrR[a + q], wv[n + q]);
                        }
                        for (int q = 0; q + 1 < n; ++q) std::swap(csr2[a + q], wv[2 * n + q]);
                    };
                    swap_tables();
                    move_centre(a);
                    identity_env<T>(ER[b + 1].as<CT>(), slotE[b + 1], PHI.wd[b + 1], 1, 1, nb, s);
                    for (int i = b; i >= a + 2; --i) env_right(PHI, i);
                    identity_env<T>(EL[a].as<CT>(), slotE[a], PHI.wd[a], 1, 0, nb, s);
                    for (int i = a; i + 2 <= b; ++i) {
                        merge(PHI, i);
                        expm(PHI, 2, i, theta.as<CT>(), nmax, 1.0, res);
                        split(PHI, i, true, 0, J.thr);
                        env_left(PHI, i);
                        expm(PHI, 1, i + 1, A(PHI, i + 1), slotA[i + 1], -1.0, res);
                    }
                    merge(PHI, b - 1);
                    expm(PHI, 2, b - 1, theta.as<CT>(), nmax, 1.0, res);
                    split(PHI, b - 1, true, 0, J.thr);
                    centre = b;
                    swap_tables();  // ownership only: the launched kernels keep their table pointers
                } else {
                    for (int k = b - 1; k >= a + 1; --k) nn(k, gops + swap_off);
                    nn(a, gops + gate_off[g]);
                    for (int k = a + 1; k <= b - 1; ++k) nn(k, gops + swap_off);
                }
                const int64_t nc = (int64_t)PHI.wd[b] * 2 * PHI.wd[b + 1];
                auto lci = lr_cand.find(a * L + b);
                if (J.has_noise && !J.noise.empty() && lci != lr_cand.end() && lci->second.n > 0) {
                    const PairCand &lc = lci->second;
                    // dissipation (dt = 1) on b (the centre) and on a, then QR steps a..b-1
                    // bring the norm back into the centre
                    if (dsite_active[b])
                        local_op<T>(A(PHI, b), slotA[b], PHI.wd[b], PHI.wd[b + 1], dsite_ops.as<c64>() + (int64_t)b * 4,
                                    nullptr, nullptr, nb, s);
                    if (lc.has_scal)  // long-range processes on (a, b): handle_longrange_scalar! at b
                        local_op<T>(A(PHI, b), slotA[b], PHI.wd[b], PHI.wd[b + 1], lc.scal.as<c64>(), nullptr, nullptr,
                                    nb, s);
                    if (dsite_active[a]) {
                        local_op<T>(A(PHI, a), slotA[a], PHI.wd[a], PHI.wd[a + 1], dsite_ops.as<c64>() + (int64_t)a * 4,
                                    nullptr, nullptr, nb, s);
                        for (int k = a; k < b; ++k) qr_step(PHI, k, nullptr);
                    }
                    norm2<T>(A(PHI, b), slotA[b], nc, norm2v.as<double>(), nullptr, nb, s);
                    jump_decide(norm2v.as<double>(), J.seed, traj0, rngk.as<uint64_t>(), 1, jstate.as<int>(),
                                fac.as<double>(), nb, s);
                    mask_from_state(jstate.as<int>(), 0, mask.as<int>(), nullptr, nb, s);
                    scale<T>(A(PHI, b), slotA[b], nc, fac.as<double>(), mask.as<int>(), nb, s);
                    fill_int(count.as<int>(), 0, 1, s);
                    mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
                    if (sync_count(count.as<int>()) > 0) {
                        // ρ_b with the centre at b, ρ_a with the centre at a
                        auto rho_site = [&](int q) {
                            const int cl = PHI.wd[q], cr = PHI.wd[q + 1];
                            copy2d<T, double>(A(PHI, q), slotA[q], 2 * cr, Xbuf.as<c64>(), slotX, 2 * cr, cl, 2 * cr,
                                              mask.as<int>(), nb, s);
                            rho<T>(Xbuf.as<c64>(), slotX, A(PHI, q), slotA[q], cl, 2, cr, rho1.as<c64>() + (int64_t)q * 4,
                                   (int64_t)L * 4, mask.as<int>(), nb, s);
                        };
                        rho_site(b);
                        // general long-range candidates: ν from two correlators (any gauge)
                        for (auto &cl : lc.lr) {
                            const ProcIn &p = J.noise[lr_cands[cl.second].proc];
                            const c64 *ops = lr_ops.as<c64>() + (int64_t)cl.second * 16;
                            correlate(PHI, a, b, ops, ops + 4, mask.as<int>(), lr_val.as<c64>());
                            correlate(PHI, a, b, ops + 8, ops + 12, mask.as<int>(), lr_val.as<c64>() + nb);
                            lr_nu(lr_val.as<c64>(), rho1.as<c64>(), (int64_t)L * 4, b, p.alpha.re, p.alpha.im,
                                  p.beta.re, p.beta.im, cl.first, lc.n, jstate.as<int>(), lrnu.as<double>(), nb, s);
                        }
                        move_centre(a);
                        rho_site(a);
                        jump_choose(rho1.as<c64>(), (int64_t)L * 4, L, lc.n, lc.site.as<int>(), lc.LdL.as<c64>(),
                                    lc.gamma.as<double>(), 1.0, J.seed, traj0, rngk.as<uint64_t>(), jstate.as<int>(),
                                    chosen.as<int>(), wbuf.as<double>(), nb, s, nullptr, 0, nullptr,
                                    lc.kind.as<int>(), lrnu.p ? lrnu.as<double>() : nullptr);
                        for (int q : {a, b}) {
                            jump_opidx(jstate.as<int>(), chosen.as<int>(), lc.site.as<int>(), lc.op.as<int>(),
                                       lc.site2.as<int>(), lc.op2.as<int>(), q, opidx.as<int>(), nb, s);
                            local_op<T>(A(PHI, q), slotA[q], PHI.wd[q], PHI.wd[q + 1], jump_ops.as<c64>(),
                                        opidx.as<int>(), nullptr, nb, s);
                        }
                        if (!lc.lr.empty()) {
                            // MPO jumps: O, truncate!, normalize! leave the centre at 0;
                            // QR steps bring it back to a like the others'
                            std::vector<int> hs(nb), hc(nb);
                            YQ_HIP(hipMemcpyAsync(hs.data(), jstate.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
                            YQ_HIP(hipMemcpyAsync(hc.data(), chosen.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
                            YQ_HIP(hipStreamSynchronize(s));
                            for (auto &cl : lc.lr) {
                                bool any = false;
                                for (int t = 0; t < nb && !any; ++t) any = hs[t] == 1 && hc[t] == cl.first;
                                if (!any) continue;
                                mask_chosen(jstate.as<int>(), chosen.as<int>(), cl.first, mask2.as<int>(), nb, s);
                                mpo_jump(PHI, cl.second, mask2.as<int>());
                                for (int k = 0; k < a; ++k) qr_step(PHI, k, mask2.as<int>());
                            }
                        }
                        fill_int(count.as<int>(), 0, 1, s);
                        mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
                        res.jumps += sync_count(count.as<int>());
                        mask_from_state(jstate.as<int>(), -1, mask.as<int>(), nullptr, nb, s);
                        for (int k = a; k < b; ++k) qr_step(PHI, k, nullptr);  // centre back at b
                        centre = b;
                        norm2<T>(A(PHI, b), slotA[b], nc, norm2v.as<double>(), mask.as<int>(), nb, s);
                        jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(),
                                    fac.as<double>(), nb, s);
                        scale<T>(A(PHI, b), slotA[b], nc, fac.as<double>(), mask.as<int>(), nb, s);
                    }
                    continue;
                }
                // no noise process acts on a or b: normalize!
                norm2<T>(A(PHI, b), slotA[b], nc, norm2v.as<double>(), nullptr, nb, s);
                jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(), fac.as<double>(), nb,
                            s);
                scale<T>(A(PHI, b), slotA[b], nc, fac.as<double>(), nullptr, nb, s);
                continue;
            }
            move_centre(a);
            merge(PHI, a);
            local_op4<T>(theta.as<CT>(), nmax, PHI.wd[a], PHI.wd[a + 2], gops + gate_off[g], nb, s);
            split(PHI, a, true, G2.rule, J.thr);  // T-D (TEBD) or T-A (window TDVP)
            centre = a + 1;
            PairCand &pc = pair_cand[a];
            const bool noisy = J.has_noise && !J.noise.empty() && pc.n > 0;
            if (noisy) {
                // dissipation on the gate's sites (dt = 1), then restore the centre at a+1
                if (dgt_scalar[a]) {
                    // f_a I on the isometry A_a and f_{a+1} I on the centre: the QR of
                    // f_a A_a would return A_a and move f_a into the centre
                    local_op<T>(A(PHI, a + 1), slotA[a + 1], PHI.wd[a + 1], PHI.wd[a + 2],
                                dscal_ops.as<c64>() + (int64_t)a * 4, nullptr, nullptr, nb, s);
                } else {
                    for (int r = 0; r < 2; ++r) {
                        const int q = a + r;
                        if (!dgt_active[2 * a + r]) continue;
                        local_op<T>(A(PHI, q), slotA[q], PHI.wd[q], PHI.wd[q + 1],
                                    dissd_ops.as<c64>() + (2 * a + r) * 4, nullptr, nullptr, nb, s);
                    }
                    qr_step(PHI, a, nullptr);  // A_a isometric again, norm in A_{a+1}
                }
                const int64_t nc = (int64_t)PHI.wd[a + 1] * 2 * PHI.wd[a + 2];
                norm2<T>(A(PHI, a + 1), slotA[a + 1], nc, norm2v.as<double>(), nullptr, nb, s);
                jump_decide(norm2v.as<double>(), J.seed, traj0, rngk.as<uint64_t>(), 1, jstate.as<int>(),
                            fac.as<double>(), nb, s);
                mask_from_state(jstate.as<int>(), 0, mask.as<int>(), nullptr, nb, s);
                scale<T>(A(PHI, a + 1), slotA[a + 1], nc, fac.as<double>(), mask.as<int>(), nb, s);
                fill_int(count.as<int>(), 0, 1, s);
                mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
                if (sync_count(count.as<int>()) > 0) {
                    // ρ of both sites from the 2-site block (left part isometric, right part canonical)
                    merge(PHI, a);
                    const int cl = PHI.wd[a], c2 = PHI.wd[a + 2];
                    copy2d<T, double>(theta.as<CT>(), nmax, 4 * c2, Xbuf.as<c64>(), slotX, 4 * c2, cl, 4 * c2,
                                      mask.as<int>(), nb, s);
                    rho<T>(Xbuf.as<c64>(), slotX, theta.as<CT>(), nmax, cl, 4, c2, rho2.as<c64>(), (int64_t)L * 16,
                           mask.as<int>(), nb, s);
                    // 1-site matrices of sites a and a+1 (partial traces of ρ2)
                    ptrace_pair(rho2.as<c64>(), (int64_t)L * 16, a, rho1.as<c64>(), (int64_t)L * 4, L, nb, s);
                    jump_choose(rho1.as<c64>(), (int64_t)L * 4, L, pc.n, pc.site.as<int>(), pc.LdL.as<c64>(),
                                pc.gamma.as<double>(), 1.0, J.seed, traj0, rngk.as<uint64_t>(), jstate.as<int>(),
                                chosen.as<int>(), wbuf.as<double>(), nb, s);
                    for (int q : {a, a + 1}) {
                        jump_opidx(jstate.as<int>(), chosen.as<int>(), pc.site.as<int>(), pc.op.as<int>(),
                                   pc.site2.as<int>(), pc.op2.as<int>(), q, opidx.as<int>(), nb, s);
                        local_op<T>(A(PHI, q), slotA[q], PHI.wd[q], PHI.wd[q + 1], jump_ops.as<c64>(), opidx.as<int>(),
                                    nullptr, nb, s);
                    }
                    fill_int(count.as<int>(), 0, 1, s);
                    mask_from_state(jstate.as<int>(), 1, mask.as<int>(), count.as<int>(), nb, s);
                    res.jumps += sync_count(count.as<int>());
                    mask_from_state(jstate.as<int>(), -1, mask.as<int>(), nullptr, nb, s);
                    qr_step(PHI, a, mask.as<int>());
                    norm2<T>(A(PHI, a + 1), slotA[a + 1], nc, norm2v.as<double>(), mask.as<int>(), nb, s);
                    jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(), fac.as<double>(),
                                nb, s);
                    scale<T>(A(PHI, a + 1), slotA[a + 1], nc, fac.as<double>(), mask.as<int>(), nb, s);
                }
            } else {
                // normalize! after a noise-free gate (src/DigitalTJM.jl:415-418)
                const int64_t nc = (int64_t)PHI.wd[a + 1] * 2 * PHI.wd[a + 2];
                norm2<T>(A(PHI, a + 1), slotA[a + 1], nc, norm2v.as<double>(), nullptr, nb, s);
                jump_decide(norm2v.as<double>(), 0, 0, rngk.as<uint64_t>(), 0, opidx.as<int>(), fac.as<double>(), nb,
                            s);
                scale<T>(A(PHI, a + 1), slotA[a + 1], nc, fac.as<double>(), nullptr, nb, s);
            }
        }
        const int lnum = (int)li + 1;
        if (std::find(after.begin(), after.end(), lnum) != after.end()) sample(idx++);
        if (progress) {
            YQ_HIP(hipStreamSynchronize(s));
            int mb = 0;
            for (int b = 0; b <= L; ++b) mb = std::max(mb, PHI.wd[b]);
            fprintf(stderr, "[yaqs] traj %lld+%d layer %d/%d  %.2fs  max bond %d\n", (long long)t0, nb, lnum,
                    (int)layers.size(), host_now() - t_run, mb);
            fflush(stderr);
        }
    }
    if (!at_start && after.empty() && !J.obs.empty()) sample(0);
}

}  // namespace

template <typename T>
void Engine<T>::collect_failed(JobResult &res) {
    std::vector<int> h(nb);
    YQ_HIP(hipMemcpyAsync(h.data(), nfail.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    YQ_HIP(hipStreamSynchronize(s));
    for (int t = 0; t < nb; ++t)
        if (h[t]) res.failed.push_back((int64_t)(traj0 - J.traj_offset) + t);
}

// ----------------------------------------------------------------- API -----
double rng_uniform_host(uint64_t seed, uint64_t traj, uint64_t k) {
    uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), (uint32_t)traj, (uint32_t)(traj >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t o0 = (uint32_t)(p1 >> 32) ^ c[1] ^ key[0], o2 = (uint32_t)(p0 >> 32) ^ c[3] ^ key[1];
        c[0] = o0;
        c[1] = (uint32_t)p1;
        c[2] = o2;
        c[3] = (uint32_t)p0;
        key[0] += 0x9E3779B9u;
        key[1] += 0xBB67AE85u;
    }
    const uint64_t bits = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
    return (double)bits * (1.0 / 9007199254740992.0);
}

void validate_job(const JobIn &job) {
    if (job.L < 1) throw std::invalid_argument("L must be >= 1");
    if ((int64_t)job.mps.size() != job.L) throw std::invalid_argument("MPS length mismatch");
    for (int64_t i = 0; i < job.L; ++i) {
        if ((int64_t)job.mps[i].size() != (int64_t)job.mps_dl[i] * 2 * job.mps_dr[i])
            throw std::invalid_argument("MPS tensors must have physical dimension 2");
        if (i + 1 < job.L && job.mps_dr[i] != job.mps_dl[i + 1]) throw std::invalid_argument("MPS bond mismatch");
    }
    if (job.mps_dl[0] != 1 || job.mps_dr[job.L - 1] != 1) throw std::invalid_argument("MPS boundary bonds must be 1");
    if (!job.digital) {
        if ((int64_t)job.mpo.size() != job.L) throw std::invalid_argument("MPO length mismatch");
        for (int64_t i = 0; i + 1 < job.L; ++i)
            if (job.mpo_dr[i] != job.mpo_dl[i + 1]) throw std::invalid_argument("MPO bond mismatch");
        if (job.num_times < 1) throw std::invalid_argument("num_times must be >= 1");
        if (job.kdim < 1 || job.kdim > 60) throw std::invalid_argument("krylov_dim must be in [1, 60]");
        int Dmax = 1;
        for (int64_t i = 0; i < job.L; ++i) Dmax = std::max(Dmax, std::max(job.mpo_dl[i], job.mpo_dr[i]));
        if (Dmax > 16) throw std::invalid_argument("MPO bond dimension > 16 is not supported");
    }
    if (job.num_traj < 1) throw std::invalid_argument("num_traj must be >= 1");
    if (job.max_bond < 1) throw std::invalid_argument("max_bond_dim must be >= 1");
    if (job.max_bond > 512) throw std::invalid_argument("max_bond_dim > 512 is not supported by the GPU path");
    for (auto &p : job.noise) {
        if (p.s0 < 0 || p.s0 >= job.L || p.s1 < 0 || p.s1 >= job.L) throw std::invalid_argument("noise site out of range");
        if (p.lr) {
            if (job.max_bond > 256)
                throw std::invalid_argument("long-range noise processes need max_bond_dim <= 256 (the jump doubles bonds)");
            continue;
        }
        if (p.nsites == 2) {
            if (std::abs(p.s1 - p.s0) != 1) throw std::invalid_argument("long-range noise processes are not supported");
            if (!p.pauli_name && job.digital)
                throw std::invalid_argument("adjacent 2-site non-Pauli noise processes are not supported by the digital "
                                            "GPU path");
        }
    }
    for (auto &o : job.obs) {
        if (o.s0 < 0 || o.s0 >= job.L || o.s1 < 0 || o.s1 >= job.L) throw std::invalid_argument("observable site out of range");
        if (o.nsites == 2 && std::abs(o.s1 - o.s0) != 1) throw std::invalid_argument("2-site observables must be adjacent");
    }
    for (auto &g : job.gates) {
        if (g.kind == 3) continue;
        if (g.s0 < 0 || g.s0 >= job.L || g.s1 < 0 || g.s1 >= job.L) throw std::invalid_argument("gate site out of range");
        if (g.kind == 2 && g.s1 == g.s0) throw std::invalid_argument("2-qubit gate on a single site");
        if (g.kind == 2 && std::abs(g.s1 - g.s0) != 1) {
            if (g.rule != 2 && !g.win_long)
                throw std::invalid_argument("bad method for a long-range 2-qubit gate");
        }
    }
}

int64_t num_columns(const JobIn &job) {
    if (!job.digital) return job.sample_timesteps ? job.num_times : 1;
    std::vector<int> busy;
    int nlayers = 0, cur = 0;
    bool at_start = false;
    std::vector<int> after;
    for (auto &g : job.gates) {
        if (g.kind == 3) {
            if (cur) { ++nlayers; cur = 0; busy.clear(); }
            if (nlayers == 0) at_start = true;
            else if (std::find(after.begin(), after.end(), nlayers) == after.end()) after.push_back(nlayers);
            continue;
        }
        std::vector<int> sites = {g.s0};
        if (g.kind == 2) sites.push_back(g.s1);
        bool ov = false;
        for (int q : sites) ov = ov || std::find(busy.begin(), busy.end(), q) != busy.end();
        if (ov) { ++nlayers; busy.clear(); cur = 0; }
        ++cur;
        for (int q : sites) busy.push_back(q);
    }
    int64_t n = (at_start ? 1 : 0) + (int64_t)after.size();
    return n == 0 ? 1 : n;
}

int64_t footprint_bytes(const JobIn &job, int64_t batch) {
    return job.precision == 1 ? Engine<float>::footprint(job, (int)batch) : Engine<double>::footprint(job, (int)batch);
}

void StreamRes::create() {
    // coherent: the Lanczos finalize kernel stores the active count straight
    // into h_pinned[1 + (j & 1)], visible to the host's spin before the kernel ends
    YQ_HIP(hipHostMalloc((void **)&h_pinned, 16 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    YQ_HIP(hipHostGetDevicePointer((void **)&d_pinned, h_pinned, 0));
    for (auto &e : ev) YQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void StreamRes::destroy() {
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    ev[0] = ev[1] = nullptr;
    if (h_pinned) (void)hipHostFree(h_pinned);
    h_pinned = d_pinned = nullptr;
}

ArenaScope::ArenaScope(void *base, size_t bytes) {
    if (g_arena) throw std::logic_error("nested scratch arenas");
    g_arena = new Arena{static_cast<char *>(base), bytes, 0};
}

ArenaScope::~ArenaScope() {
    delete g_arena;
    g_arena = nullptr;
}

size_t ArenaScope::used() const { return g_arena ? g_arena->used : 0; }

namespace {
template <typename T>
class WaveEngineT : public WaveEngine {
  public:
    WaveEngineT(const JobIn &job, int batch, hipStream_t s, StreamRes *res)
        : eng(job, batch, s, res), digital(job.digital) {}
    void run(int64_t t0, int nb, double *d_out, int64_t *h_bonds, JobResult &res) override {
        if (digital) eng.run_digital(t0, nb, d_out, h_bonds, res);
        else eng.run_analog(t0, nb, d_out, res);
        eng.collect_failed(res);
    }

  private:
    Engine<T> eng;
    bool digital;
};
}  // namespace

double WindowSync::arrive() {
    std::unique_lock<std::mutex> lk(mu);
    if (++arrived == n) {
        t0 = host_now();
        cv.notify_all();
    } else {
        cv.wait(lk, [&] { return arrived >= n; });
    }
    return t0;
}

void WindowSync::release() {
    std::lock_guard<std::mutex> lk(mu);
    if (arrived < n) {
        arrived = n;
        t0 = host_now();
    }
    cv.notify_all();
}

std::unique_ptr<WaveEngine> make_engine(const JobIn &job, int batch, hipStream_t stream, StreamRes *res) {
    validate_job(job);
    if (job.precision == 1) return std::unique_ptr<WaveEngine>(new WaveEngineT<float>(job, batch, stream, res));
    return std::unique_ptr<WaveEngine>(new WaveEngineT<double>(job, batch, stream, res));
}

void run_job(const JobIn &job, int64_t batch, hipStream_t stream, JobResult &out, std::atomic<bool> *cancel) {
    validate_job(job);
    const int64_t ntraj = job.num_traj;
    const int64_t ncol = num_columns(job);
    const int64_t nobs = (int64_t)job.obs.size();
    out.ncol = ncol;
    out.traj.assign((size_t)std::max<int64_t>(1, nobs * ncol * ntraj), 0.0);
    out.mean.assign((size_t)std::max<int64_t>(1, nobs * ncol), 0.0);
    if (job.digital) out.bonds.assign((size_t)ntraj * ncol, 0);
    const int64_t B = std::max<int64_t>(1, std::min(batch, ntraj));
    double *d_out = nullptr;
    YQ_HIP(hipMalloc(&d_out, out.traj.size() * sizeof(double)));
    try {
        YQ_HIP(hipMemsetAsync(d_out, 0, out.traj.size() * sizeof(double), stream));
        auto go = [&](auto *tag) {
            typedef std::remove_pointer_t<decltype(tag)> TT;
            Engine<TT> eng(job, (int)B, stream);
            for (int64_t t0 = 0; t0 < ntraj; t0 += B) {
                if (cancel && cancel->load()) throw std::runtime_error("cancelled");
                const int nb = (int)std::min<int64_t>(B, ntraj - t0);
                if (job.digital) eng.run_digital(t0, nb, d_out, out.bonds.data(), out);
                else eng.run_analog(t0, nb, d_out, out);
            }
        };
        if (job.precision == 1) go((float *)nullptr);
        else go((double *)nullptr);
        YQ_HIP(hipMemcpyAsync(out.traj.data(), d_out, out.traj.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        YQ_HIP(hipStreamSynchronize(stream));
    } catch (...) {
        (void)hipFree(d_out);
        throw;
    }
    (void)hipFree(d_out);
    for (int64_t q = 0; q < nobs; ++q)
        for (int64_t c = 0; c < ncol; ++c) {
            double sum = 0;
            for (int64_t t = 0; t < ntraj; ++t) sum += out.traj[((size_t)q * ncol + c) * ntraj + t];
            out.mean[(size_t)q * ncol + c] = sum / (double)ntraj;
        }
}

}  // namespace yq

extern "C" void yaqs_debug_poison_theta(int64_t traj, int64_t split, int precision) {
    yq::g_poison_split = split;
    yq::g_poison_prec = precision;
    yq::g_poison_traj = traj;
}
