// tjm_oracle.cpp — FP64 CPU restatement of Yaqs.jl's Tensor Jump Method.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the
// CPU baseline of the MI355X engine.  It is never linked into, loaded by or
// called from the product library; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, as the checker.
//
// It follows the reference line by line, in the reference's own storage
// order (column-major tensors, so every `reshape` of the Julia code is free
// here as well):
//   AnalogTJM          src/AnalogTJM.jl:21-176
//   2-site TDVP        src/Algorithms.jl:144-226, 409-445, 580-819
//   1-site TDVP        src/Algorithms.jl:222-226, 449-452, 471-532
//   MPS gauge/measure  src/MPS.jl:182-311, 528-651
//   Dissipation        src/Dissipation.jl:14-159
//   Jumps              src/StochasticProcess.jl:35-330
//   Simulator loop     src/Simulator.jl:22-125
//   DigitalTJM         src/DigitalTJM.jl:86-305, 307-358, 363-450 (TEBD, TDVP
//                      windows, SWAP networks)
// The dense kernels the reference takes from LAPACK/KrylovKit are written
// here from their textbook definitions: one-sided (Hestenes) Jacobi SVD,
// Householder QR/LQ, Lanczos with full re-orthogonalisation, and a cyclic
// Jacobi eigensolver for the small tridiagonal.
//
// Parity status of each function is pinned by tests/test_oracle_*.py
// (golden CSVs of the reference, closed forms, numpy/scipy, exact solvers).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include <dlfcn.h>

#include "../include/yaqs_hip.h"

namespace {

typedef std::complex<double> cd;
typedef std::vector<cd> cvec;

// ---------------------------------------------------------------- RNG ------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Draw k of global
// trajectory t under `seed`: key = seed, counter = (k_lo, k_hi, t_lo, t_hi);
// the uniform double is built from the top 53 bits of the first two words.
struct Philox {
    static void round(uint32_t c[4], const uint32_t k[2]) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t o[4] = {hi1 ^ c[1] ^ k[0], lo1, hi0 ^ c[3] ^ k[1], lo0};
        c[0] = o[0]; c[1] = o[1]; c[2] = o[2]; c[3] = o[3];
    }
    static double uniform(uint64_t seed, uint64_t traj, uint64_t k) {
        uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), (uint32_t)traj, (uint32_t)(traj >> 32)};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        for (int r = 0; r < 10; ++r) {
            round(c, key);
            key[0] += 0x9E3779B9u;
            key[1] += 0xBB67AE85u;
        }
        const uint64_t bits = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
        return (double)bits * (1.0 / 9007199254740992.0);
    }
};

struct TrajRng {
    uint64_t seed, traj, k = 0;
    double next() { return Philox::uniform(seed, traj, k++); }
};

// ---------------------------------------------------------- dense utils ----
// Column-major complex matrix product C(m,n) = op(A) * op(B); ta/tb: 'N',
// 'T' or 'C' (conjugate transpose).  Inner loop runs over contiguous rows.
// Optional BLAS: the reference's tensor contractions run on OpenBLAS zgemm
// (LinearAlgebra / TensorOperations, OpenBLAS 0.3.29 in its Manifest); when
// oracle_use_blas() has bound a Fortran-ABI zgemm with 64-bit integers
// (numpy's bundled OpenBLAS 0.3.29), products above a small size go there,
// single-threaded inside each trajectory's thread as the reference runs
// them (BLAS.set_num_threads(1)).  Same mathematics, another summation
// order: results move at the 1e-15 level.
typedef void (*zgemm64_fn)(const char *, const char *, const int64_t *, const int64_t *, const int64_t *,
                           const cd *, const cd *, const int64_t *, const cd *, const int64_t *, const cd *, cd *,
                           const int64_t *, long, long);
zgemm64_fn g_zgemm = nullptr;
// Optional LAPACK zgesdd (the reference's `svd`, LinearAlgebra's default
// divide-and-conquer driver, src/Algorithms.jl:420) for the timing baseline
// only: the parity tests keep the Jacobi SVD below, whose null-space
// completion the tests pin.
typedef void (*zgesdd64_fn)(const char *, const int64_t *, const int64_t *, cd *, const int64_t *, double *, cd *,
                            const int64_t *, cd *, const int64_t *, cd *, const int64_t *, double *, int64_t *,
                            int64_t *, long);
zgesdd64_fn g_zgesdd = nullptr;
// Dead directions of QR/LQ dropped (qr_thin); off only to measure how much
// the rounding-decided completion moves results (oracle_set_drop_dead).
bool g_drop_dead = true;
thread_local bool tl_keep_all = false;  // set during measurements (see expect)
// sim_params.truncation_threshold / max_bond_dim of the running trajectory,
// read by the weights of long-range MPO jumps (src/StochasticProcess.jl:63-79)
thread_local double tl_thr = 1e-12;
thread_local int tl_max_bond = 1 << 30;

void gemm(char ta, char tb, int m, int n, int k, const cd *A, int lda, const cd *B, int ldb,
          cd *C, int ldc, bool accumulate = false) {
    if (g_zgemm && (int64_t)m * n * k >= 4096) {
        const int64_t M = m, N = n, K = k, LA = std::max(lda, 1), LB = std::max(ldb, 1), LC = std::max(ldc, 1);
        const cd one(1.0, 0.0), beta(accumulate ? 1.0 : 0.0, 0.0);
        g_zgemm(&ta, &tb, &M, &N, &K, &one, A, &LA, B, &LB, &beta, C, &LC, 1, 1);
        return;
    }
    if (!accumulate)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) C[i + (size_t)ldc * j] = 0.0;
    if (m == 0 || n == 0 || k == 0) return;
    // Materialise op(A) as an m x k column-major block when transposed.
    cvec At;
    const cd *Ap = A;
    int lda2 = lda;
    if (ta != 'N') {
        At.resize((size_t)m * k);
        for (int p = 0; p < k; ++p)
            for (int i = 0; i < m; ++i) {
                cd v = A[p + (size_t)lda * i];
                At[i + (size_t)m * p] = (ta == 'C') ? std::conj(v) : v;
            }
        Ap = At.data();
        lda2 = m;
    }
    const double *a = reinterpret_cast<const double *>(Ap);
    double *c = reinterpret_cast<double *>(C);
    for (int j = 0; j < n; ++j) {
        double *cj = c + 2 * (size_t)ldc * j;
        for (int p = 0; p < k; ++p) {
            cd bv = (tb == 'N') ? B[p + (size_t)ldb * j]
                                : (tb == 'T' ? B[j + (size_t)ldb * p] : std::conj(B[j + (size_t)ldb * p]));
            const double br = bv.real(), bi = bv.imag();
            if (br == 0.0 && bi == 0.0) continue;
            const double *ap = a + 2 * (size_t)lda2 * p;
            for (int i = 0; i < m; ++i) {
                const double ar = ap[2 * i], ai = ap[2 * i + 1];
                cj[2 * i] += ar * br - ai * bi;
                cj[2 * i + 1] += ar * bi + ai * br;
            }
        }
    }
}

// Permute a column-major 4-index array: dst dims = src dims[perm[.]].
cvec permute4(const cvec &src, const int d[4], const int perm[4]) {
    int nd[4] = {d[perm[0]], d[perm[1]], d[perm[2]], d[perm[3]]};
    size_t sstride[4] = {1, (size_t)d[0], (size_t)d[0] * d[1], (size_t)d[0] * d[1] * d[2]};
    cvec dst((size_t)nd[0] * nd[1] * nd[2] * nd[3]);
    size_t idx = 0;
    for (int i3 = 0; i3 < nd[3]; ++i3)
        for (int i2 = 0; i2 < nd[2]; ++i2)
            for (int i1 = 0; i1 < nd[1]; ++i1)
                for (int i0 = 0; i0 < nd[0]; ++i0) {
                    size_t s = i0 * sstride[perm[0]] + i1 * sstride[perm[1]] + i2 * sstride[perm[2]] +
                               i3 * sstride[perm[3]];
                    dst[idx++] = src[s];
                }
    return dst;
}

double nrm2(const cvec &v) {
    double s = 0;
    for (const cd &x : v) s += std::norm(x);
    return std::sqrt(s);
}
cd dotc(const cd *a, const cd *b, size_t n) {  // Σ conj(a)·b (Julia `dot`)
    double re = 0, im = 0;
    for (size_t i = 0; i < n; ++i) {
        re += a[i].real() * b[i].real() + a[i].imag() * b[i].imag();
        im += a[i].real() * b[i].imag() - a[i].imag() * b[i].real();
    }
    return cd(re, im);
}

// Thin QR of a column-major m x n matrix: A = Q R, Q (m x r), R (r x n),
// r = min(m, n), as Julia's `Matrix(qr(A).Q)`, `qr(A).R` (src/MPS.jl:
// 200-205), up to the phase of each column of Q (the freedom of a QR; the
// state Q R is the same).  Columns are orthonormalised in order by
// classical Gram-Schmidt applied twice.  Dead directions: a column whose
// residual is at most 1e-12 ||A||_F depends on the earlier ones; LAPACK
// builds a reflector from its rounding noise there, so the column of Q it
// returns is decided by that noise (and so is the 2-site TDVP tangent space
// that follows).  This restatement drops such a direction instead: a zero
// column of Q and a zero row of R (later columns are orthogonalised against
// the live ones only; the next live column takes its slot), the same state
// with the bond at its numerical rank -- the rule the GPU engine's gauge
// steps follow.  With g_drop_dead off it
// runs Householder reflectors as LAPACK does (used to measure how far the
// noise-decided completion moves results).
void qr_householder(const cvec &A, int m, int n, cvec &Q, cvec &R, int &r) {
    r = std::min(m, n);
    cvec W = A;
    std::vector<cvec> vs(r);
    std::vector<double> taus(r, 0.0);
    for (int k = 0; k < r; ++k) {
        int len = m - k;
        cvec v(len);
        for (int i = 0; i < len; ++i) v[i] = W[(k + i) + (size_t)m * k];
        double xn = 0;
        for (auto &x : v) xn += std::norm(x);
        xn = std::sqrt(xn);
        if (xn == 0.0) { vs[k] = cvec(len, 0.0); taus[k] = 0.0; continue; }
        cd x0 = v[0];
        cd phase = (std::abs(x0) > 0) ? x0 / std::abs(x0) : cd(1.0, 0.0);
        cd alpha = -phase * xn;
        v[0] -= alpha;
        double vn2 = 0;
        for (auto &x : v) vn2 += std::norm(x);
        if (vn2 == 0.0) { vs[k] = cvec(len, 0.0); taus[k] = 0.0; continue; }
        // apply H = I - 2 v v^H / (v^H v) to W[k:, k:]
        for (int j = k; j < n; ++j) {
            cd s = 0;
            for (int i = 0; i < len; ++i) s += std::conj(v[i]) * W[(k + i) + (size_t)m * j];
            s *= 2.0 / vn2;
            for (int i = 0; i < len; ++i) W[(k + i) + (size_t)m * j] -= v[i] * s;
        }
        vs[k] = v;
        taus[k] = 2.0 / vn2;
    }
    R.assign((size_t)r * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= std::min(j, r - 1); ++i) R[i + (size_t)r * j] = W[i + (size_t)m * j];
    // Q = H_0 ... H_{r-1} [I_r; 0]
    Q.assign((size_t)m * r, 0.0);
    for (int j = 0; j < r; ++j) Q[j + (size_t)m * j] = 1.0;
    for (int k = r - 1; k >= 0; --k) {
        if (taus[k] == 0.0) continue;
        const cvec &v = vs[k];
        int len = m - k;
        for (int j = 0; j < r; ++j) {
            cd s = 0;
            for (int i = 0; i < len; ++i) s += std::conj(v[i]) * Q[(k + i) + (size_t)m * j];
            s *= taus[k];
            for (int i = 0; i < len; ++i) Q[(k + i) + (size_t)m * j] -= v[i] * s;
        }
    }
}

void qr_thin(const cvec &A, int m, int n, cvec &Q, cvec &R, int &r) {
    if (!g_drop_dead || tl_keep_all) {
        qr_householder(A, m, n, Q, R, r);
        return;
    }
    r = std::min(m, n);
    double fro = 0;
    for (auto &x : A) fro += std::norm(x);
    fro = std::sqrt(fro);
    const double dead = 1e-12 * fro;
    Q.assign((size_t)m * r, 0.0);
    R.assign((size_t)r * n, 0.0);
    // a dead column takes no column of Q: the next live one does (with r
    // slots for r = min(m, n) columns, a live direction after a dead column
    // would otherwise find no slot); unused slots stay zero
    int slot = 0;
    for (int j = 0; j < n; ++j) {
        cvec v(A.begin() + (size_t)m * j, A.begin() + (size_t)m * (j + 1));
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < slot; ++k) {
                const cd *q = &Q[(size_t)m * k];
                const cd h = dotc(q, v.data(), m);
                R[k + (size_t)r * j] += h;
                for (int i = 0; i < m; ++i) v[i] -= h * q[i];
            }
        if (slot < r) {
            const double nr = nrm2(v);
            if (nr > dead && nr > 0.0) {
                R[slot + (size_t)r * j] = nr;
                for (int i = 0; i < m; ++i) Q[i + (size_t)m * slot] = v[i] / nr;
                ++slot;
            }
        }
    }
}

// LQ of a column-major m x n matrix: A = L Q with Q (r x n) orthonormal rows,
// computed as the QR of A^H (Julia's `lq`, src/MPS.jl:236-244).
void lq_thin(const cvec &A, int m, int n, cvec &L, cvec &Q, int &r) {
    cvec AH((size_t)n * m);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) AH[j + (size_t)n * i] = std::conj(A[i + (size_t)m * j]);
    cvec Qh, Rh;
    qr_thin(AH, n, m, Qh, Rh, r);  // AH = Qh (n x r) Rh (r x m)
    L.assign((size_t)m * r, 0.0);   // L = Rh^H (m x r)
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < m; ++j) L[j + (size_t)m * i] = std::conj(Rh[i + (size_t)r * j]);
    Q.assign((size_t)r * n, 0.0);   // Q = Qh^H (r x n)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < r; ++j) Q[j + (size_t)r * i] = std::conj(Qh[i + (size_t)n * j]);
}

// One-sided (Hestenes) Jacobi SVD of a column-major m x n matrix with m >= n:
// rotate column pairs until mutually orthogonal.  Returns U (m x n), S (n,
// descending), V (n x n) with A = U diag(S) V^H.
void svd_tall(const cvec &A, int m, int n, cvec &U, std::vector<double> &S, cvec &V) {
    cvec X = A;
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
    const double eps = 2.220446049250313e-16;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                cd *cp = &X[(size_t)m * p], *cq = &X[(size_t)m * q];
                double a = 0, b = 0;
                for (int i = 0; i < m; ++i) { a += std::norm(cp[i]); b += std::norm(cq[i]); }
                cd g = dotc(cp, cq, m);
                double ag = std::abs(g);
                if (ag == 0.0 || ag <= eps * std::sqrt(a * b)) continue;
                off = std::max(off, ag / std::sqrt(a * b));
                cd ph = g / ag;
                double zeta = (b - a) / (2.0 * ag);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                // [cp', cq'] = [cp, cq] * [[c, s e^{iφ}], [-s e^{-iφ}, c]]
                cd m01 = s * ph, m10 = -s * std::conj(ph);
                for (int i = 0; i < m; ++i) {
                    cd xp = cp[i], xq = cq[i];
                    cp[i] = c * xp + m10 * xq;
                    cq[i] = m01 * xp + c * xq;
                }
                cd *vp = &V[(size_t)n * p], *vq = &V[(size_t)n * q];
                for (int i = 0; i < n; ++i) {
                    cd xp = vp[i], xq = vq[i];
                    vp[i] = c * xp + m10 * xq;
                    vq[i] = m01 * xp + c * xq;
                }
            }
        if (off <= eps * 4) break;
    }
    std::vector<double> sv(n);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s += std::norm(X[i + (size_t)m * j]);
        sv[j] = std::sqrt(s);
    }
    std::vector<int> order(n);
    for (int j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sv[x] > sv[y]; });
    U.assign((size_t)m * n, 0.0);
    cvec V2((size_t)n * n);
    S.assign(n, 0.0);
    const double smax = n ? sv[order[0]] : 0.0;
    for (int jj = 0; jj < n; ++jj) {
        int j = order[jj];
        S[jj] = sv[j];
        for (int i = 0; i < n; ++i) V2[i + (size_t)n * jj] = V[i + (size_t)n * j];
        if (sv[j] > smax * 1e-300 && sv[j] > 0)
            for (int i = 0; i < m; ++i) U[i + (size_t)m * jj] = X[i + (size_t)m * j] / sv[j];
    }
    V = V2;
    // Orthonormal completion of U for zero singular values (LAPACK returns
    // some orthonormal basis there; any choice reproduces the same state).
    for (int jj = 0; jj < n; ++jj) {
        if (S[jj] > 0) continue;
        for (int e = 0; e < m; ++e) {
            cvec u(m, 0.0);
            u[e] = 1.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < n; ++k) {
                    if (k == jj) continue;
                    cd *uk = &U[(size_t)m * k];
                    cd h = dotc(uk, u.data(), m);
                    for (int i = 0; i < m; ++i) u[i] -= h * uk[i];
                }
            double un = nrm2(u);
            if (un > 1e-8) {
                for (int i = 0; i < m; ++i) U[i + (size_t)m * jj] = u[i] / un;
                break;
            }
        }
    }
}

// Thin SVD A = U S Vt of a column-major m x n matrix (Julia `svd`,
// src/Algorithms.jl:420).  U (m x r), S (r), Vt (r x n), r = min(m, n).
void svd_thin(const cvec &A, int m, int n, cvec &U, std::vector<double> &S, cvec &Vt) {
    if (g_zgesdd && m > 0 && n > 0) {
        const int64_t M = m, N = n, r = std::min(m, n), mx = std::max(m, n);
        cvec X = A, work(1);
        U.assign((size_t)m * r, 0.0);
        Vt.assign((size_t)r * n, 0.0);
        S.assign(r, 0.0);
        std::vector<double> rwork((size_t)(2 * mx * r + 5 * r * r + 7 * r + 1));
        std::vector<int64_t> iwork((size_t)8 * r);
        int64_t lwork = -1, info = 0, ldu = m, ldvt = r;
        g_zgesdd("S", &M, &N, X.data(), &M, S.data(), U.data(), &ldu, Vt.data(), &ldvt, work.data(), &lwork,
                 rwork.data(), iwork.data(), &info, 1);
        lwork = std::max<int64_t>(1, (int64_t)work[0].real());
        work.resize((size_t)lwork);
        g_zgesdd("S", &M, &N, X.data(), &M, S.data(), U.data(), &ldu, Vt.data(), &ldvt, work.data(), &lwork,
                 rwork.data(), iwork.data(), &info, 1);
        if (info == 0) return;
    }
    if (m >= n) {
        cvec V;
        svd_tall(A, m, n, U, S, V);
        Vt.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Vt[i + (size_t)n * j] = std::conj(V[j + (size_t)n * i]);
    } else {
        cvec AH((size_t)n * m);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) AH[j + (size_t)n * i] = std::conj(A[i + (size_t)m * j]);
        cvec U2, V2;
        svd_tall(AH, n, m, U2, S, V2);  // A^H = U2 S V2^H  ->  A = V2 S U2^H
        U = V2;                          // m x m
        Vt.assign((size_t)m * n, 0.0);   // U2^H (m x n)
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) Vt[i + (size_t)m * j] = std::conj(U2[j + (size_t)n * i]);
    }
}

// Cyclic Jacobi eigensolver of a real symmetric k x k matrix (column-major).
void sym_eig(std::vector<double> A, int k, std::vector<double> &w, std::vector<double> &Q) {
    Q.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) Q[i + (size_t)k * i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) {
                double v = A[i + (size_t)k * j] * A[i + (size_t)k * j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                double apq = A[p + (size_t)k * q];
                if (apq == 0.0) continue;
                double app = A[p + (size_t)k * p], aqq = A[q + (size_t)k * q];
                double zeta = (aqq - app) / (2.0 * apq);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < k; ++r) {  // columns p, q
                    double x = A[r + (size_t)k * p], y = A[r + (size_t)k * q];
                    A[r + (size_t)k * p] = c * x - s * y;
                    A[r + (size_t)k * q] = s * x + c * y;
                }
                for (int r = 0; r < k; ++r) {  // rows p, q
                    double x = A[p + (size_t)k * r], y = A[q + (size_t)k * r];
                    A[p + (size_t)k * r] = c * x - s * y;
                    A[q + (size_t)k * r] = s * x + c * y;
                }
                for (int r = 0; r < k; ++r) {
                    double x = Q[r + (size_t)k * p], y = Q[r + (size_t)k * q];
                    Q[r + (size_t)k * p] = c * x - s * y;
                    Q[r + (size_t)k * q] = s * x + c * y;
                }
            }
    }
    w.resize(k);
    for (int i = 0; i < k; ++i) w[i] = A[i + (size_t)k * i];
}

// c = exp(-i * tau * T) e_0 for the symmetric tridiagonal T(alpha, beta).
void tridiag_exp_e0(const std::vector<double> &alpha, const std::vector<double> &beta, int k, double tau,
                    cvec &c) {
    std::vector<double> T((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) {
        T[i + (size_t)k * i] = alpha[i];
        if (i + 1 < k) {
            T[i + (size_t)k * (i + 1)] = beta[i + 1];
            T[(i + 1) + (size_t)k * i] = beta[i + 1];
        }
    }
    std::vector<double> w, Q;
    sym_eig(T, k, w, Q);
    c.assign(k, 0.0);
    for (int m = 0; m < k; ++m) {
        cd f = std::exp(cd(0.0, -tau * w[m])) * Q[0 + (size_t)k * m];
        for (int i = 0; i < k; ++i) c[i] += Q[i + (size_t)k * m] * f;
    }
}

// exp(-i dt A) v by Lanczos with full re-orthogonalisation — the role of
// KrylovKit.exponentiate(A, -i*dt, v; tol, krylovdim, maxiter=1,
// ishermitian=true) at src/Algorithms.jl:144-187.  Stopping rule (shared with
// the GPU engine): after step j, with c = exp(-i dt T_j) e_0,
//   ‖v‖ * β_{j+1} * |c_j| <= tol   or   j+1 == krylovdim.
// ‖v‖ == 0 returns v unchanged (src/Algorithms.jl:145-148).
template <class F>
cvec expm_krylov(F &&apply, const cvec &v, double dt, int kmax, double tol, int *iters = nullptr) {
    const size_t n = v.size();
    double b0 = nrm2(v);
    if (b0 == 0.0) { if (iters) *iters = 0; return v; }
    std::vector<cvec> V;
    V.reserve(kmax + 1);
    cvec v0(n);
    for (size_t i = 0; i < n; ++i) v0[i] = v[i] / b0;
    V.push_back(v0);
    std::vector<double> alpha, beta(1, 0.0);
    cvec c;
    for (int j = 0; j < kmax; ++j) {
        cvec w = apply(V[j]);
        double a = dotc(V[j].data(), w.data(), n).real();
        for (size_t i = 0; i < n; ++i) w[i] -= a * V[j][i];
        if (j > 0)
            for (size_t i = 0; i < n; ++i) w[i] -= beta[j] * V[j - 1][i];
        // one classical Gram-Schmidt pass against the whole basis
        std::vector<cd> h(j + 1);
        for (int q = 0; q <= j; ++q) h[q] = dotc(V[q].data(), w.data(), n);
        for (int q = 0; q <= j; ++q)
            for (size_t i = 0; i < n; ++i) w[i] -= h[q] * V[q][i];
        alpha.push_back(a);
        double b = nrm2(w);
        tridiag_exp_e0(alpha, beta, j + 1, dt, c);
        double err = b0 * b * std::abs(c[j]);
        if (err <= tol || b == 0.0 || j + 1 == kmax) {
            if (iters) *iters = j + 1;
            cvec out(n, 0.0);
            for (int q = 0; q <= j; ++q) {
                cd f = b0 * c[q];
                for (size_t i = 0; i < n; ++i) out[i] += f * V[q][i];
            }
            return out;
        }
        beta.push_back(b);
        for (size_t i = 0; i < n; ++i) w[i] /= b;
        V.push_back(std::move(w));
    }
    return v;  // unreachable
}

// ----------------------------------------------------------- tensors -------
struct T3 {  // column-major (dl, d, dr), src/MPS.jl:27-44
    int dl = 1, d = 2, dr = 1;
    cvec v;
    T3() {}
    T3(int a, int b, int c) : dl(a), d(b), dr(c), v((size_t)a * b * c, 0.0) {}
    cd &at(int a, int s, int b) { return v[a + (size_t)dl * (s + (size_t)d * b)]; }
};
struct T4 {  // column-major (dl, dout, din, dr), src/MPO.jl:14-25
    int dl = 1, po = 2, pi = 2, dr = 1;
    cvec v;
};

struct MPSState {  // src/MPS.jl:33-44
    std::vector<T3> t;
    int L = 0;
    int oc = 1;  // 1-based orth_center field
};

// shift_orthogonality_center! (src/MPS.jl:182-266)
void shift_oc(MPSState &s, int nc) {
    int oc = s.oc;
    if (oc == nc) return;
    if (nc > oc) {
        for (int i = oc; i <= nc - 1; ++i) {
            T3 &A = s.t[i - 1];
            int L = A.dl, d = A.d, R = A.dr, r;
            cvec Q, Rm;
            qr_thin(A.v, L * d, R, Q, Rm, r);
            T3 nA(L, d, r);
            nA.v = Q;
            T3 &N = s.t[i];
            T3 nN(r, N.d, N.dr);
            gemm('N', 'N', r, N.d * N.dr, R, Rm.data(), r, N.v.data(), R, nN.v.data(), r);
            s.t[i - 1] = std::move(nA);
            s.t[i] = std::move(nN);
        }
    } else {
        for (int i = oc; i >= nc + 1; --i) {
            T3 &A = s.t[i - 1];
            int L = A.dl, d = A.d, R = A.dr, r;
            cvec Lm, Q;
            lq_thin(A.v, L, d * R, Lm, Q, r);
            T3 nA(r, d, R);
            nA.v = Q;
            T3 &P = s.t[i - 2];
            T3 nP(P.dl, P.d, r);
            gemm('N', 'N', P.dl * P.d, r, L, P.v.data(), P.dl * P.d, Lm.data(), L, nP.v.data(), P.dl * P.d);
            s.t[i - 1] = std::move(nA);
            s.t[i - 2] = std::move(nP);
        }
    }
    s.oc = nc;
}

// normalize! (src/MPS.jl:274-311): LQ sweep L->2, scale site 1 if ‖·‖>1e-13.
void normalize_mps(MPSState &s) {
    for (int i = s.L; i >= 2; --i) {
        T3 &A = s.t[i - 1];
        int L = A.dl, d = A.d, R = A.dr, r;
        cvec Lm, Q;
        lq_thin(A.v, L, d * R, Lm, Q, r);
        T3 nA(r, d, R);
        nA.v = Q;
        T3 &P = s.t[i - 2];
        T3 nP(P.dl, P.d, r);
        gemm('N', 'N', P.dl * P.d, r, L, P.v.data(), P.dl * P.d, Lm.data(), L, nP.v.data(), P.dl * P.d);
        s.t[i - 1] = std::move(nA);
        s.t[i - 2] = std::move(nP);
    }
    double nrm = nrm2(s.t[0].v);
    if (nrm > 1e-13)
        for (auto &x : s.t[0].v) x *= 1.0 / nrm;
    s.oc = 1;
}

// scalar_product(ψ, ψ) (src/MPS.jl:528-585).
cd scalar_product(const MPSState &a, const MPSState &b) {
    cvec E(1, 1.0);
    int ea = 1, eb = 1;
    for (int i = 0; i < a.L; ++i) {
        const T3 &A = a.t[i], &B = b.t[i];
        // T1 = E^T * A_mat  (Lb, d*Ra)
        cvec T1((size_t)eb * A.d * A.dr);
        gemm('T', 'N', eb, A.d * A.dr, ea, E.data(), ea, A.v.data(), A.dl, T1.data(), eb);
        // E_next = T1r^T * conj(B_mat), T1r (Lb*d, Ra), B_mat (Lb*d, Rb)
        cvec Bc(B.v.size());
        for (size_t q = 0; q < Bc.size(); ++q) Bc[q] = std::conj(B.v[q]);
        cvec En((size_t)A.dr * B.dr);
        gemm('T', 'N', A.dr, B.dr, eb * A.d, T1.data(), eb * A.d, Bc.data(), B.dl * B.d, En.data(), A.dr);
        E = En;
        ea = A.dr;
        eb = B.dr;
    }
    return E[0];
}

// local_expect (src/MPS.jl:608-624): shift centre, real(dot(A_perm, op*A_perm)).
cd local_expect(MPSState &s, const cd op[4], int site) {
    shift_oc(s, site);
    T3 &A = s.t[site - 1];
    cd acc = 0;
    for (int b = 0; b < A.dr; ++b)
        for (int a = 0; a < A.dl; ++a)
            for (int p = 0; p < A.d; ++p) {
                cd o = 0;
                for (int k = 0; k < A.d; ++k) o += op[p + 2 * k] * A.at(a, k, b);
                acc += std::conj(A.at(a, p, b)) * o;
            }
    return acc;
}

// Θ = A1 · A2 merged over the shared bond: (l, p1, p2, r) column-major.
cvec merge2(const T3 &A, const T3 &B) {
    cvec th((size_t)A.dl * A.d * B.d * B.dr);
    gemm('N', 'N', A.dl * A.d, B.d * B.dr, A.dr, A.v.data(), A.dl * A.d, B.v.data(), B.dl, th.data(),
         A.dl * A.d);
    return th;
}

// Apply a 4x4 (column-major, combined index p1 + 2 p2) to Θ (l, p1, p2, r).
cvec apply_op2(const cd *op, const cvec &th, int l, int r) {
    cvec out(th.size(), 0.0);
    for (int b = 0; b < r; ++b)
        for (int a = 0; a < l; ++a)
            for (int po = 0; po < 4; ++po) {
                cd acc = 0;
                for (int pi = 0; pi < 4; ++pi)
                    acc += op[po + 4 * pi] * th[a + (size_t)l * ((pi & 1) + 2 * ((pi >> 1) + 2 * (size_t)b))];
                out[a + (size_t)l * ((po & 1) + 2 * ((po >> 1) + 2 * (size_t)b))] = acc;
            }
    return out;
}

// local_expect_two_site (src/MPS.jl:626-651).
cd local_expect2(MPSState &s, const cd op[16], int s1) {
    shift_oc(s, s1);
    const T3 &A = s.t[s1 - 1], &B = s.t[s1];
    cvec th = merge2(A, B);
    cvec o = apply_op2(op, th, A.dl, B.dr);
    return dotc(th.data(), o.data(), th.size());
}

// ------------------------------------------------------------ TDVP ---------
// make_identity_env (src/Algorithms.jl:191-197): E[i, 1, i] = 1.
cvec identity_env(int dim, int mpo) {
    cvec E((size_t)dim * mpo * dim, 0.0);
    for (int i = 0; i < dim; ++i) E[i + (size_t)dim * (0 + (size_t)mpo * i)] = 1.0;
    return E;
}

struct Env { int b, m, k; cvec v; };  // E[bra, mpo, ket] column-major

// T2[bl, kr, pout, mr] = Σ L[bl,ml,k] A[k,pin,kr] W[ml,pout,pin,mr] (shared by
// project_site and update_left_environment, src/Algorithms.jl:199-218,354-357).
cvec left_T2(const Env &E, const cvec &A, int Akl, int pin, int Akr, const T4 &W) {
    int bl = E.b, ml = E.m;
    cvec T1((size_t)bl * ml * pin * Akr);
    gemm('N', 'N', bl * ml, pin * Akr, Akl, E.v.data(), bl * ml, A.data(), Akl, T1.data(), bl * ml);
    int d1[4] = {bl, ml, pin, Akr};
    int p1[4] = {0, 3, 1, 2};  // -> [bl, kr, ml, pin]
    cvec T1p = permute4(T1, d1, p1);
    int dw[4] = {W.dl, W.po, W.pi, W.dr};
    int pw[4] = {0, 2, 1, 3};  // -> [ml, pin, pout, mr]
    cvec Wp = permute4(W.v, dw, pw);
    cvec T2((size_t)bl * Akr * W.po * W.dr);
    gemm('N', 'N', bl * Akr, W.po * W.dr, ml * pin, T1p.data(), bl * Akr, Wp.data(), ml * pin, T2.data(),
         bl * Akr);
    return T2;
}

// project_site (src/Algorithms.jl:215-220, 330-361): H_eff · A.
cvec project_site(const cvec &A, int Akl, int pin, int Akr, const Env &Lenv, const Env &Renv, const T4 &W) {
    cvec T2 = left_T2(Lenv, A, Akl, pin, Akr, W);
    int bl = Lenv.b, po = W.po, mr = W.dr, br = Renv.b;
    int d2[4] = {bl, Akr, po, mr};
    int p2[4] = {0, 2, 1, 3};  // -> [bl, pout, kr, mr]
    cvec T2p = permute4(T2, d2, p2);
    int dr_[4] = {Renv.b, Renv.m, Renv.k, 1};
    int pr[4] = {2, 1, 0, 3};  // R[br, mr, kr] -> [kr, mr, br]
    cvec Rp = permute4(Renv.v, dr_, pr);
    cvec out((size_t)bl * po * br);
    gemm('N', 'N', bl * po, br, Akr * mr, T2p.data(), bl * po, Rp.data(), Akr * mr, out.data(), bl * po);
    return out;
}

// update_left_environment (src/Algorithms.jl:199-204).
Env update_left_env(const T3 &A, const T4 &W, const Env &E) {
    cvec T2 = left_T2(E, A.v, A.dl, A.d, A.dr, W);
    int bl = E.b, kr = A.dr, po = W.po, mr = W.dr;
    int d2[4] = {bl, kr, po, mr};
    int p2[4] = {0, 2, 1, 3};  // -> [bl, pout, kr, mr]
    cvec T2p = permute4(T2, d2, p2);
    cvec X((size_t)A.dr * kr * mr);  // [br, kr, mr] = conj(A)^T · T2p
    gemm('C', 'N', A.dr, kr * mr, bl * po, A.v.data(), A.dl * A.d, T2p.data(), bl * po, X.data(), A.dr);
    int dx[4] = {A.dr, kr, mr, 1};
    int px[4] = {0, 2, 1, 3};  // -> [br, mr, kr]
    Env En{A.dr, mr, kr, permute4(X, dx, px)};
    return En;
}

// update_right_environment (src/Algorithms.jl:206-211).
Env update_right_env(const T3 &A, const T4 &W, const Env &E) {
    int kl = A.dl, pin = A.d, br = E.b, mr = E.m;
    cvec T1((size_t)kl * pin * br * mr);  // [kl, pin, br, mr] = A_mat · E_mat^T
    gemm('N', 'T', kl * pin, br * mr, A.dr, A.v.data(), kl * pin, E.v.data(), br * mr, T1.data(), kl * pin);
    int d1[4] = {kl, pin, br, mr};
    int p1[4] = {1, 3, 0, 2};  // -> [pin, mr, kl, br]
    cvec T1p = permute4(T1, d1, p1);
    int dw[4] = {W.dl, W.po, W.pi, W.dr};
    int pw[4] = {0, 1, 2, 3};
    cvec Wm = permute4(W.v, dw, pw);  // (ml*pout) x (pin*mr)
    cvec T2((size_t)W.dl * W.po * kl * br);
    gemm('N', 'N', W.dl * W.po, kl * br, pin * mr, Wm.data(), W.dl * W.po, T1p.data(), pin * mr, T2.data(),
         W.dl * W.po);
    int d2[4] = {W.dl, W.po, kl, br};
    int p2[4] = {1, 3, 0, 2};  // -> [pout, br, ml, kl]
    cvec T2p = permute4(T2, d2, p2);
    // A permuted to [bl, pout, br] is A itself; conj(A) as bl x (pout*br)
    cvec Ac(A.v.size());
    for (size_t q = 0; q < Ac.size(); ++q) Ac[q] = std::conj(A.v[q]);
    Env En{A.dl, W.dl, kl, cvec((size_t)A.dl * W.dl * kl)};
    gemm('N', 'N', A.dl, W.dl * kl, W.po * br, Ac.data(), A.dl, T2p.data(), W.po * br, En.v.data(), A.dl);
    return En;
}

// W_group of two MPO sites (src/Algorithms.jl:639-643): (l, p1o*p2o, p1i*p2i, r)
// with combined indices p1 + 2 p2.
T4 merge_mpo(const T4 &W1, const T4 &W2) {
    T4 G;
    G.dl = W1.dl;
    G.po = W1.po * W2.po;
    G.pi = W1.pi * W2.pi;
    G.dr = W2.dr;
    G.v.assign((size_t)G.dl * G.po * G.pi * G.dr, 0.0);
    for (int l = 0; l < W1.dl; ++l)
        for (int k = 0; k < W1.dr; ++k)
            for (int r = 0; r < W2.dr; ++r)
                for (int a = 0; a < W1.po; ++a)
                    for (int b = 0; b < W1.pi; ++b) {
                        cd w1 = W1.v[l + (size_t)W1.dl * (a + (size_t)W1.po * (b + (size_t)W1.pi * k))];
                        if (w1 == 0.0) continue;
                        for (int c = 0; c < W2.po; ++c)
                            for (int e = 0; e < W2.pi; ++e) {
                                cd w2 = W2.v[k + (size_t)W2.dl * (c + (size_t)W2.po * (e + (size_t)W2.pi * r))];
                                int po = a + W1.po * c, pi = b + W1.pi * e;
                                G.v[l + (size_t)G.dl * (po + (size_t)G.po * (pi + (size_t)G.pi * r))] += w1 * w2;
                            }
                    }
    return G;
}

struct Cfg {
    double dt;
    int max_bond;
    double thr;
    int kdim;
    double ktol;
};

struct Counters {
    int64_t matvec2 = 0, matvec1 = 0, svd = 0, env = 0;
    double flop = 0;
};

// split_mps_tensor_svd with rule T-A (src/Algorithms.jl:409-445).
void split_svd(cvec &th, int l, int p1, int p2, int r, const Cfg &cfg, cvec &U, std::vector<double> &S,
               cvec &Vt, int &keep) {
    for (auto &x : th)
        if (!std::isfinite(x.real()) || !std::isfinite(x.imag())) x = 0.0;  // NaN guard :414-417
    int m = l * p1, n = p2 * r;
    cvec Uf, Vf;
    std::vector<double> Sf;
    svd_thin(th, m, n, Uf, Sf, Vf);
    int ns = (int)Sf.size();
    double disc = 0.0;
    int keep_rank = ns;
    for (int k = ns; k >= 1; --k) {
        disc += Sf[k - 1] * Sf[k - 1];
        if (disc >= cfg.thr) { keep_rank = std::max(k, 2); break; }
    }
    keep = std::max(1, std::min(keep_rank, cfg.max_bond));
    keep = std::min(keep, ns);
    U.assign(Uf.begin(), Uf.begin() + (size_t)m * keep);
    S.assign(Sf.begin(), Sf.begin() + keep);
    Vt.assign((size_t)keep * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < keep; ++i) Vt[i + (size_t)keep * j] = Vf[i + (size_t)ns * j];
}

double flop_2site(int cl, int cr, int D) {
    return 8.0 * (4.0 * D * cl * cr * (double)(cl + cr) + 16.0 * D * D * (double)cl * cr);
}
double flop_1site(int cl, int cr, int D) {
    return 8.0 * (2.0 * D * cl * cr * (double)(cl + cr) + 4.0 * D * D * (double)cl * cr);
}

struct TdvpCtx {
    const std::vector<T4> *H;
    Cfg cfg;
    Counters *cnt;
};

cvec expm_site(const TdvpCtx &C, const cvec &A, int kl, int p, int kr, const Env &Le, const Env &Re,
               const T4 &W, double dt, bool two) {
    int iters = 0;
    auto f = [&](const cvec &x) { return project_site(x, kl, p, kr, Le, Re, W); };
    cvec out = expm_krylov(f, A, dt, C.cfg.kdim, C.cfg.ktol, &iters);
    if (C.cnt) {
        double per = two ? flop_2site(kl, kr, W.dl) : flop_1site(kl, kr, W.dl);
        C.cnt->flop += per * iters;
        (two ? C.cnt->matvec2 : C.cnt->matvec1) += iters;
    }
    return out;
}

// _two_site_update_forward! (src/Algorithms.jl:631-672)
void two_site_forward(MPSState &s, const TdvpCtx &C, std::vector<Env> &EL, std::vector<Env> &ER, int i,
                      double dts, bool back) {
    const std::vector<T4> &H = *C.H;
    T3 &A1 = s.t[i - 1], &A2 = s.t[i];
    int l = A1.dl, p1 = A1.d, p2 = A2.d, r = A2.dr;
    cvec th = merge2(A1, A2);
    T4 Wg = merge_mpo(H[i - 1], H[i]);
    cvec thn = expm_site(C, th, l, p1 * p2, r, EL[i - 1], ER[i + 1], Wg, dts, true);
    cvec U, Vt;
    std::vector<double> S;
    int keep;
    split_svd(thn, l, p1, p2, r, C.cfg, U, S, Vt, keep);
    if (C.cnt) C.cnt->svd++;
    T3 nA1(l, p1, keep);
    nA1.v = U;
    s.t[i - 1] = nA1;
    EL[i] = update_left_env(s.t[i - 1], H[i - 1], EL[i - 1]);
    if (C.cnt) C.cnt->env++;
    T3 A2t(keep, p2, r);
    for (int j = 0; j < p2 * r; ++j)
        for (int q = 0; q < keep; ++q) A2t.v[q + (size_t)keep * j] = S[q] * Vt[q + (size_t)keep * j];
    if (back) {
        T3 A2n(keep, p2, r);
        A2n.v = expm_site(C, A2t.v, keep, p2, r, EL[i], ER[i + 1], H[i], -dts, false);
        s.t[i] = A2n;
    } else {
        s.t[i] = A2t;
    }
}

// Shared tail of the edge and backward updates: evolve Θ, split left
// (src/Algorithms.jl:674-711, 713-760).
void two_site_split_left(MPSState &s, const TdvpCtx &C, std::vector<Env> &EL, std::vector<Env> &ER, int i,
                         double dts) {
    const std::vector<T4> &H = *C.H;
    T3 &A1 = s.t[i - 1], &A2 = s.t[i];
    int l = A1.dl, p1 = A1.d, p2 = A2.d, r = A2.dr;
    cvec th = merge2(A1, A2);
    T4 Wg = merge_mpo(H[i - 1], H[i]);
    cvec thn = expm_site(C, th, l, p1 * p2, r, EL[i - 1], ER[i + 1], Wg, dts, true);
    cvec U, Vt;
    std::vector<double> S;
    int keep;
    split_svd(thn, l, p1, p2, r, C.cfg, U, S, Vt, keep);
    if (C.cnt) C.cnt->svd++;
    T3 nA2(keep, p2, r);
    nA2.v = Vt;
    s.t[i] = nA2;
    ER[i] = update_right_env(s.t[i], H[i], ER[i + 1]);
    if (C.cnt) C.cnt->env++;
    T3 nA1(l, p1, keep);
    for (int q = 0; q < keep; ++q)
        for (int a = 0; a < l * p1; ++a) nA1.v[a + (size_t)l * p1 * q] = U[a + (size_t)l * p1 * q] * S[q];
    s.t[i - 1] = nA1;
}

// _tdvp_sweep_hamiltonian_2site! (src/Algorithms.jl:580-603) with _init_envs
// (src/Algorithms.jl:802-819).
void two_site_tdvp(MPSState &s, const TdvpCtx &C) {
    const std::vector<T4> &H = *C.H;
    shift_oc(s, 1);
    const int L = s.L;
    const double dt = C.cfg.dt;
    std::vector<Env> EL(L + 1), ER(L + 1);
    {
        int rb = s.t[L - 1].dr, rm = H[L - 1].dr;
        ER[L] = Env{rb, rm, rb, identity_env(rb, rm)};
        for (int i = L; i >= 2; --i) {
            ER[i - 1] = update_right_env(s.t[i - 1], H[i - 1], ER[i]);
            if (C.cnt) C.cnt->env++;
        }
        int lb = s.t[0].dl, lm = H[0].dl;
        EL[0] = Env{lb, lm, lb, identity_env(lb, lm)};
    }
    for (int i = 1; i <= L - 2; ++i) two_site_forward(s, C, EL, ER, i, dt / 2, true);
    if (L >= 2) two_site_split_left(s, C, EL, ER, L - 1, dt);
    for (int i = L - 2; i >= 1; --i) {
        // 1. pre-correct site i+1 by -dt/2 with E_left[i+1], E_right[i+2]
        T3 &A2 = s.t[i];
        T3 nA2(A2.dl, A2.d, A2.dr);
        nA2.v = expm_site(C, A2.v, A2.dl, A2.d, A2.dr, EL[i], ER[i + 1], H[i], -dt / 2, false);
        s.t[i] = nA2;
        two_site_split_left(s, C, EL, ER, i, dt / 2);
    }
}

// project_bond (src/Algorithms.jl:222-226): C (k x kr, column-major) ->
// C_new[bra_l, bra_r] = sum L[bra_l, mpo, k] C[k, ket_r] R[bra_r, mpo, ket_r].
cvec project_bond(const cvec &Cm, int k, int kr, const Env &Le, const Env &Re) {
    cvec T1((size_t)Le.b * Le.m * kr);  // [bra_l, mpo, ket_r]
    gemm('N', 'N', Le.b * Le.m, kr, k, Le.v.data(), Le.b * Le.m, Cm.data(), k, T1.data(), Le.b * Le.m);
    cvec out((size_t)Le.b * Re.b);
    gemm('N', 'T', Le.b, Re.b, Le.m * kr, T1.data(), Le.b, Re.v.data(), Re.b, out.data(), Le.b);
    return out;
}

cvec expm_bond(const TdvpCtx &C, const cvec &Cm, int k, int kr, const Env &Le, const Env &Re, double dt) {
    int iters = 0;
    auto f = [&](const cvec &x) { return project_bond(x, k, kr, Le, Re); };
    cvec out = expm_krylov(f, Cm, dt, C.cfg.kdim, C.cfg.ktol, &iters);
    if (C.cnt) C.cnt->flop += 8.0 * 2.0 * Le.m * (double)k * k * kr * iters;
    return out;
}

// _tdvp_sweep_hamiltonian_1site! (src/Algorithms.jl:471-532): forward, each
// site evolves by dt/2, then (but the last) its QR moves the centre right and
// the R factor evolves by -dt/2 under the bond Hamiltonian before it is
// absorbed into the next site; backward the same with LQ from site L.
void single_site_tdvp(MPSState &s, const TdvpCtx &C) {
    const std::vector<T4> &H = *C.H;
    shift_oc(s, 1);
    const int L = s.L;
    const double dt = C.cfg.dt;
    std::vector<Env> EL(L + 1), ER(L + 1);
    {
        int rb = s.t[L - 1].dr, rm = H[L - 1].dr;
        ER[L] = Env{rb, rm, rb, identity_env(rb, rm)};
        for (int i = L; i >= 2; --i) {
            ER[i - 1] = update_right_env(s.t[i - 1], H[i - 1], ER[i]);
            if (C.cnt) C.cnt->env++;
        }
        int lb = s.t[0].dl, lm = H[0].dl;
        EL[0] = Env{lb, lm, lb, identity_env(lb, lm)};
    }
    for (int i = 1; i <= L; ++i) {  // site i (1-based): EL[i-1], ER[i]
        {
            T3 &A = s.t[i - 1];
            A.v = expm_site(C, A.v, A.dl, A.d, A.dr, EL[i - 1], ER[i], H[i - 1], dt / 2, false);
        }
        if (i == L) break;
        const T3 &A = s.t[i - 1];
        const int l = A.dl, p = A.d, r = A.dr;
        int rr;
        cvec Q, R;
        qr_thin(A.v, l * p, r, Q, R, rr);
        T3 nA(l, p, rr);
        nA.v = Q;
        s.t[i - 1] = std::move(nA);
        EL[i] = update_left_env(s.t[i - 1], H[i - 1], EL[i - 1]);
        if (C.cnt) C.cnt->env++;
        const cvec Cn = expm_bond(C, R, rr, r, EL[i], ER[i], -dt / 2);
        const T3 &N = s.t[i];
        T3 nN(rr, N.d, N.dr);
        gemm('N', 'N', rr, N.d * N.dr, r, Cn.data(), rr, N.v.data(), r, nN.v.data(), rr);
        s.t[i] = std::move(nN);
    }
    for (int i = L; i >= 1; --i) {
        {
            T3 &A = s.t[i - 1];
            A.v = expm_site(C, A.v, A.dl, A.d, A.dr, EL[i - 1], ER[i], H[i - 1], dt / 2, false);
        }
        if (i == 1) break;
        const T3 &A = s.t[i - 1];
        const int l = A.dl, p = A.d, r = A.dr;
        int rr;
        cvec Lm, Q;
        lq_thin(A.v, l, p * r, Lm, Q, rr);
        T3 nA(rr, p, r);
        nA.v = Q;
        s.t[i - 1] = std::move(nA);
        ER[i - 1] = update_right_env(s.t[i - 1], H[i - 1], ER[i]);
        if (C.cnt) C.cnt->env++;
        const cvec Cn = expm_bond(C, Lm, l, rr, EL[i - 1], ER[i - 1], -dt / 2);
        const T3 &P = s.t[i - 2];
        T3 nP(P.dl, P.d, rr);
        gemm('N', 'N', P.dl * P.d, rr, l, P.v.data(), P.dl * P.d, Cn.data(), l, nP.v.data(), P.dl * P.d);
        s.t[i - 2] = std::move(nP);
    }
}

// The sweep the analog path runs: two_site_tdvp! in the reference
// (src/AnalogTJM.jl:44,68,161); single_site_tdvp! when the job asks for it.
void tdvp_step(MPSState &s, const TdvpCtx &C, int sites) {
    if (sites == 1) single_site_tdvp(s, C);
    else two_site_tdvp(s, C);
}

// ------------------------------------------------------- noise model ------
struct Proc {
    int nsites;
    int s0, s1;
    double gamma;
    bool pauli_name;  // name-based test, src/Dissipation.jl:14-21
    cd m[16];         // 2x2 (first 4) or 4x4, column-major
    // long-range process (MPONoiseProcess, src/Noise.jl:22-33) whose MPO is
    // build_mpo_phys's O = alpha I + beta sig_{s0} tau_{s1} (src/Noise.jl:
    // 77-128); pauli_name then marks the "pauli_long_range" candidate
    bool lr = false;
    cd sig[4], tau[4], alpha = 0.0, beta = 0.0;
};

bool is_pauli_matrix(const Proc &p) {  // src/StochasticProcess.jl:337-342
    int d = p.nsites == 1 ? 2 : 4;
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            cd acc = 0;
            for (int k = 0; k < d; ++k) acc += std::conj(p.m[k + d * i]) * p.m[k + d * j];
            if (std::abs(acc - (i == j ? 1.0 : 0.0)) > 1e-10) return false;
        }
    return true;
}
bool is_pauli_proc(const Proc &p) { return p.pauli_name || is_pauli_matrix(p); }  // :344-353

// exp(-0.5 dt γ L^†L) for a d x d (d = 2, 4) operator via eigen-decomposition
// of the Hermitian L^†L (the role of `exp(::Matrix)`, src/Dissipation.jl:63-64,95).
void herm_expm_damp(const cd *Lop, int d, double scale, cd *out) {
    // M = L^†L (Hermitian); embed as 2d x 2d real symmetric [[Re, -Im],[Im, Re]].
    cd M[16];
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            cd acc = 0;
            for (int k = 0; k < d; ++k) acc += std::conj(Lop[k + d * i]) * Lop[k + d * j];
            M[i + d * j] = acc;
        }
    int n2 = 2 * d;
    std::vector<double> R((size_t)n2 * n2);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            R[i + n2 * j] = M[i + d * j].real();
            R[(i + d) + n2 * (j + d)] = M[i + d * j].real();
            R[(i + d) + n2 * j] = M[i + d * j].imag();
            R[i + n2 * (j + d)] = -M[i + d * j].imag();
        }
    std::vector<double> w, Q;
    sym_eig(R, n2, w, Q);
    // exp(-scale M) in the embedding, then read back the complex block.
    std::vector<double> E((size_t)n2 * n2, 0.0);
    for (int m = 0; m < n2; ++m) {
        double f = std::exp(-scale * w[m]);
        for (int i = 0; i < n2; ++i)
            for (int j = 0; j < n2; ++j) E[i + n2 * j] += Q[i + n2 * m] * f * Q[j + n2 * m];
    }
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) out[i + d * j] = cd(E[i + n2 * j], E[(i + d) + n2 * j]);
}

// Relative-threshold rank, rule T-C (src/Dissipation.jl:123-136,
// src/StochasticProcess.jl:288-300).
int rank_relative(const std::vector<double> &S, double thr, int max_bond) {
    int rank = (int)S.size();
    double ns = 0;
    for (double x : S) ns += x * x;
    double cur = 0;
    for (int k = (int)S.size(); k >= 1; --k) {
        cur += S[k - 1] * S[k - 1];
        if (cur > thr * ns) { rank = k; break; }
    }
    return std::max(1, std::min(rank, max_bond));
}

void split_relative(MPSState &s, int i1, const cvec &thn, int l, int r, double thr, int max_bond) {
    cvec U, Vt;
    std::vector<double> S;
    svd_thin(thn, l * 2, 2 * r, U, S, Vt);
    int ns = (int)S.size();
    int rank = std::min(rank_relative(S, thr, max_bond), ns);
    T3 A(l, 2, rank), B(rank, 2, r);
    for (size_t q = 0; q < (size_t)l * 2 * rank; ++q) A.v[q] = U[q];
    for (int j = 0; j < 2 * r; ++j)
        for (int q = 0; q < rank; ++q) B.v[q + (size_t)rank * j] = S[q] * Vt[q + (size_t)ns * j];
    s.t[i1 - 1] = A;
    s.t[i1] = B;
}

void apply_1site_op(T3 &A, const cd *op) {  // T_new[l,p,r] = Σ_k op[p,k] T[l,k,r]
    T3 n(A.dl, A.d, A.dr);
    for (int b = 0; b < A.dr; ++b)
        for (int p = 0; p < A.d; ++p)
            for (int k = 0; k < A.d; ++k) {
                cd o = op[p + 2 * k];
                if (o == 0.0) continue;
                for (int a = 0; a < A.dl; ++a) n.at(a, p, b) += o * A.at(a, k, b);
            }
    A = n;
}

// apply_dissipation (src/Dissipation.jl:37-159).
void apply_dissipation(MPSState &s, const std::vector<Proc> *noise, double dt, double thr, int max_bond) {
    bool none = (noise == nullptr);
    if (!none) {
        bool allz = true;
        for (auto &p : *noise) allz = allz && (p.gamma == 0.0);
        none = allz;  // `all(p.strength == 0)` is true for an empty list too
    }
    if (none) {
        for (int i = s.L; i >= 2; --i) shift_oc(s, i - 1);
        return;
    }
    for (int i = s.L; i >= 1; --i) {
        for (const Proc &p : *noise) {
            if (p.nsites == 1 && p.s0 == i) {
                if (p.pauli_name) {
                    double f = std::exp(-0.5 * dt * p.gamma);
                    for (auto &x : s.t[i - 1].v) x *= f;
                } else {
                    cd op[4];
                    herm_expm_damp(p.m, 2, 0.5 * dt * p.gamma, op);
                    apply_1site_op(s.t[i - 1], op);
                }
            }
        }
        if (i > 1) {
            for (const Proc &p : *noise) {
                if (p.nsites != 2 || std::max(p.s0, p.s1) != i) continue;
                if (p.lr) {
                    // handle_longrange_scalar! (src/Dissipation.jl:161-204):
                    // exp(-dt/2 γ) per process (a projector pair's one factor
                    // exp(-dt/2 (γ+ + γ-)) is the product of its two)
                    double f = std::exp(-0.5 * dt * p.gamma);
                    for (auto &x : s.t[i - 1].v) x *= f;
                } else if (p.pauli_name) {
                    double f = std::exp(-0.5 * dt * p.gamma);
                    for (auto &x : s.t[i - 1].v) x *= f;
                } else {
                    cd op[16];
                    herm_expm_damp(p.m, 4, 0.5 * dt * p.gamma, op);
                    const T3 &A = s.t[i - 2], &B = s.t[i - 1];
                    int l = A.dl, r = B.dr;
                    cvec th = merge2(A, B);
                    cvec thn = apply_op2(op, th, l, r);
                    split_relative(s, i - 1, thn, l, r, thr, max_bond);
                }
            }
        }
        if (i > 1) shift_oc(s, i - 1);
    }
}

// MPS.truncate! (src/MPS.jl:320-441): the centre to site 1, then one SVD
// sweep 1 -> L-1: Θ = A_i A_{i+1} scaled by 1/max|Θ| for the SVD (singular
// values scaled back), kept rank by the absolute rule (discarded weight
// >= threshold, at least 2) capped at max_bond, A_i = U, A_{i+1} = S V^†;
// then normalize!.  Returns the discarded weight.
double truncate_mps(MPSState &s, double thr, int max_bond) {
    shift_oc(s, 1);
    double err = 0;
    for (int i = 1; i <= s.L - 1; ++i) {
        const T3 &A = s.t[i - 1], &B = s.t[i];
        const int l = A.dl, d1 = A.d, d2 = B.d, r = B.dr;
        cvec th = merge2(A, B);  // (l d1) x (d2 r), column-major
        double scale = 0;
        for (auto &x : th) scale = std::max(scale, std::abs(x));
        if (!std::isfinite(scale)) throw std::runtime_error("truncate!: non-finite entries encountered in Theta");
        if (scale != 0)
            for (auto &x : th) x *= 1.0 / scale;
        cvec U, Vt;
        std::vector<double> S;
        svd_thin(th, l * d1, d2 * r, U, S, Vt);
        const int ns = (int)S.size();
        if (scale != 0)
            for (auto &x : S) x *= scale;
        double disc = 0;
        int keep = ns;
        for (int k = ns; k >= 1; --k) {
            disc += S[k - 1] * S[k - 1];
            if (disc >= thr) { keep = std::max(k, 2); break; }
        }
        keep = std::min(keep, max_bond);
        for (int k = keep; k < ns; ++k) err += S[k] * S[k];
        T3 nA(l, d1, keep), nB(keep, d2, r);
        for (size_t q = 0; q < (size_t)l * d1 * keep; ++q) nA.v[q] = U[q];
        for (int j = 0; j < d2 * r; ++j)
            for (int q = 0; q < keep; ++q) nB.v[q + (size_t)keep * j] = S[q] * Vt[q + (size_t)ns * j];
        s.t[i - 1] = std::move(nA);
        s.t[i] = std::move(nB);
    }
    s.oc = s.L;
    normalize_mps(s);
    return err;
}

// build_mpo_phys's tensor at site k (1-based) of O = alpha I + beta sig_a
// tau_b (src/Noise.jl:77-128), (Left, Out, In, Right) column-major.
T4 lr_mpo_site(const Proc &p, int k) {
    const int a = std::min(p.s0, p.s1), b = std::max(p.s0, p.s1);
    const cd I2[4] = {1.0, 0.0, 0.0, 1.0};
    auto put = [](T4 &W, int l, int r, const cd *op, cd f) {
        for (int o = 0; o < 2; ++o)
            for (int in = 0; in < 2; ++in) W.v[l + W.dl * (o + 2 * (in + 2 * r))] = f * op[o + 2 * in];
    };
    T4 W;
    if (k < a || k > b) {
        W = T4{1, 2, 2, 1, cvec(4)};
        put(W, 0, 0, I2, 1.0);
    } else if (k == a) {
        W = T4{1, 2, 2, 2, cvec(8)};
        put(W, 0, 0, I2, 1.0);
        put(W, 0, 1, p.sig, 1.0);
    } else if (k < b) {
        W = T4{2, 2, 2, 2, cvec(16)};
        put(W, 0, 0, I2, 1.0);
        put(W, 1, 1, I2, 1.0);
    } else {
        W = T4{2, 2, 2, 1, cvec(8)};
        put(W, 0, 0, I2, p.alpha);
        put(W, 1, 0, p.tau, p.beta);
    }
    return W;
}

// apply_mpo_jump! (src/StochasticProcess.jl:46-82): every site contracted
// with its MPO tensor over the physical index, C[la, lw, po, ra, rw] =
// Σ_pi A[la, pi, ra] W[lw, po, pi, rw] reshaped (la·lw, po, ra·rw) (la and
// ra fastest), then truncate! with the job's threshold and bond cap.
void apply_mpo_jump(MPSState &s, const Proc &p, double thr, int max_bond) {
    for (int k = 1; k <= s.L; ++k) {
        T3 &A = s.t[k - 1];
        const T4 W = lr_mpo_site(p, k);
        T3 C(A.dl * W.dl, W.po, A.dr * W.dr);
        for (int rw = 0; rw < W.dr; ++rw)
            for (int ra = 0; ra < A.dr; ++ra)
                for (int po = 0; po < W.po; ++po)
                    for (int lw = 0; lw < W.dl; ++lw)
                        for (int pi = 0; pi < W.pi; ++pi) {
                            const cd w = W.v[lw + W.dl * (po + W.po * (pi + W.pi * rw))];
                            if (w == 0.0) continue;
                            for (int la = 0; la < A.dl; ++la)
                                C.at(la + A.dl * lw, po, ra + A.dr * rw) += A.at(la, pi, ra) * w;
                        }
        s.t[k - 1] = std::move(C);
    }
    truncate_mps(s, thr, max_bond);
}

struct Candidate {
    int proc;
    int type;  // 1 = local_1, 2 = local_2, 3 = pauli_long_range, 4 = mpo_general
    int site;
};

// create_probability_distribution (src/StochasticProcess.jl:89-212).
void prob_distribution(MPSState &s, const std::vector<Proc> &noise, double dt, std::vector<double> &probs,
                       std::vector<Candidate> &cands) {
    // its centre walk is a gauge move the GPU engine does not make (it reads
    // the weights from transfer matrices): it keeps every direction
    struct KeepAll {
        KeepAll() { tl_keep_all = true; }
        ~KeepAll() { tl_keep_all = false; }
    } keep;
    probs.clear();
    cands.clear();
    if (noise.empty()) return;
    std::vector<double> dps;
    for (int site = 1; site <= s.L; ++site) {
        shift_oc(s, site);
        for (size_t q = 0; q < noise.size(); ++q) {
            const Proc &p = noise[q];
            if (p.nsites == 1 && p.s0 == site) {
                double nrm;
                if (is_pauli_matrix(p)) {
                    nrm = 0;
                    for (auto &x : s.t[site - 1].v) nrm += std::norm(x);
                } else {
                    T3 tn = s.t[site - 1];
                    apply_1site_op(tn, p.m);
                    nrm = 0;
                    for (auto &x : tn.v) nrm += std::norm(x);
                }
                dps.push_back(dt * p.gamma * nrm);
                cands.push_back({(int)q, 1, site});
            }
        }
        if (site < s.L) {
            for (size_t q = 0; q < noise.size(); ++q) {
                const Proc &p = noise[q];
                if (p.nsites != 2) continue;
                int a = std::min(p.s0, p.s1), b = std::max(p.s0, p.s1);
                if (!(a == site && b == site + 1) || p.s0 != site) continue;
                double nrm = 0;
                if (is_pauli_proc(p)) {
                    for (auto &x : s.t[site - 1].v) nrm += std::norm(x);
                } else {
                    const T3 &A = s.t[site - 1], &B = s.t[site];
                    cvec th = merge2(A, B);
                    cvec thn = apply_op2(p.m, th, A.dl, B.dr);
                    for (auto &x : thn) nrm += std::norm(x);
                }
                dps.push_back(dt * p.gamma * nrm);
                cands.push_back({(int)q, 2, site});
            }
        }
    }
    // long-range candidates in process order (src/StochasticProcess.jl:
    // 168-198), the centre at L: a "pauli_long_range" one weighs the centre
    // norm; a general one the norm of a copy after apply_mpo_jump!
    for (size_t q = 0; q < noise.size(); ++q) {
        const Proc &p = noise[q];
        if (!p.lr) continue;
        double nrm = 0;
        if (p.pauli_name) {
            for (auto &x : s.t[s.oc - 1].v) nrm += std::norm(x);
            dps.push_back(dt * p.gamma * nrm);
            cands.push_back({(int)q, 3, 0});
        } else {
            MPSState tmp = s;
            apply_mpo_jump(tmp, p, tl_thr, tl_max_bond);
            for (auto &x : tmp.t[tmp.oc - 1].v) nrm += std::norm(x);
            dps.push_back(dt * p.gamma * nrm);
            cands.push_back({(int)q, 4, 0});
        }
    }
    double tot = 0;
    for (double x : dps) tot += x;
    if (tot == 0) { cands.clear(); return; }
    for (double x : dps) probs.push_back(x / tot);
}

// stochastic_process! (src/StochasticProcess.jl:219-330).  Returns 1 if a
// jump was applied.
int stochastic_process(MPSState &s, const std::vector<Proc> &noise, double dt, double thr, int max_bond,
                       TrajRng &rng) {
    tl_thr = thr;
    tl_max_bond = max_bond;
    double dp = 1.0 - scalar_product(s, s).real();
    if (noise.empty() || rng.next() >= dp) {
        shift_oc(s, 1);
        normalize_mps(s);
        return 0;
    }
    std::vector<double> probs;
    std::vector<Candidate> cands;
    prob_distribution(s, noise, dt, probs, cands);
    if (probs.empty()) {
        shift_oc(s, 1);
        normalize_mps(s);
        return 0;
    }
    double r = rng.next();
    double acc = 0;
    size_t chosen = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
        acc += probs[i];
        if (r <= acc) { chosen = i; break; }
    }
    const Candidate &c = cands[chosen];
    const Proc &p = noise[c.proc];
    if (std::getenv("ORACLE_DEBUG_JUMPS")) {
        fprintf(stderr, "oracle jump traj=%llu k=%llu cand=%zu site=%d dp=%.17g r2=%.17g probs:",
                (unsigned long long)rng.traj, (unsigned long long)rng.k, chosen, c.site, dp, r);
        for (double x : probs) fprintf(stderr, " %.15g", x);
        fprintf(stderr, "\n");
    }
    if (c.type == 3) {
        // "pauli_long_range": no branch of the reference applies it
        // (src/StochasticProcess.jl:255-327); only normalize! follows
    } else if (c.type == 4) {
        apply_mpo_jump(s, p, thr, max_bond);
    } else if (c.type == 1) {
        shift_oc(s, c.site);
        apply_1site_op(s.t[c.site - 1], p.m);
    } else {
        int s1 = c.site;
        shift_oc(s, s1);
        const T3 &A = s.t[s1 - 1], &B = s.t[s1];
        int l = A.dl, r2 = B.dr;
        cvec th = merge2(A, B);
        cvec thn = apply_op2(p.m, th, l, r2);
        split_relative(s, s1, thn, l, r2, thr, max_bond);
    }
    normalize_mps(s);
    return 1;
}

struct Obs {
    int nsites;
    int s0, s1;
    cd op[16];
};

// The centre shifts of a measurement are gauge moves on the state itself
// (src/MPS.jl:608-651; analog_tjm_1 measures the evolving state): they keep
// every direction (Householder, as LAPACK), where the sweeps of the
// algorithm drop dead ones; the GPU engine measures through transfer
// matrices and leaves the state's gauge alone.
double expect(MPSState &s, const Obs &o) {  // src/SimulationConfigs.jl:42-56
    struct KeepAll {
        KeepAll() { tl_keep_all = true; }
        ~KeepAll() { tl_keep_all = false; }
    } keep;
    if (o.nsites == 1) return local_expect(s, o.op, o.s0).real();
    int a = std::min(o.s0, o.s1);
    return local_expect2(s, o.op, a).real();
}

// ---------------------------------------------------------- job input ------
void load_c(const double *src, cd *dst, size_t n) {
    for (size_t i = 0; i < n; ++i) dst[i] = cd(src[2 * i], src[2 * i + 1]);
}

MPSState load_mps(int64_t L, const yaqs_mps_tensor *t) {
    MPSState s;
    s.L = (int)L;
    s.oc = 1;
    s.t.resize(L);
    for (int64_t i = 0; i < L; ++i) {
        T3 A((int)t[i].dl, (int)t[i].d, (int)t[i].dr);
        load_c(t[i].data, A.v.data(), A.v.size());
        s.t[i] = A;
    }
    return s;
}

std::vector<Proc> load_noise(int64_t n, const yaqs_noise_process *p) {
    std::vector<Proc> out;
    for (int64_t i = 0; i < n; ++i) {
        Proc q;
        q.nsites = p[i].nsites;
        q.s0 = (int)p[i].site0;
        q.s1 = (int)(p[i].nsites == 2 ? p[i].site1 : p[i].site0);
        q.gamma = p[i].strength;
        q.pauli_name = p[i].pauli_name != 0;
        if (p[i].kind == YAQS_NOISE_LONG_RANGE) {
            q.lr = true;
            load_c(p[i].matrix, q.sig, 4);
            load_c(p[i].matrix + 8, q.tau, 4);
            q.alpha = cd(p[i].matrix[16], p[i].matrix[17]);
            q.beta = cd(p[i].matrix[18], p[i].matrix[19]);
            for (auto &z : q.m) z = 0.0;
            out.push_back(q);
            continue;
        }
        int d = q.nsites == 1 ? 2 : 4;
        load_c(p[i].matrix, q.m, (size_t)d * d);
        out.push_back(q);
    }
    return out;
}

std::vector<Obs> load_obs(int64_t n, const yaqs_observable *o) {
    std::vector<Obs> out;
    for (int64_t i = 0; i < n; ++i) {
        Obs q;
        q.nsites = o[i].nsites;
        q.s0 = (int)o[i].site0;
        q.s1 = (int)(o[i].nsites == 2 ? o[i].site1 : o[i].site0);
        load_c(o[i].op, q.op, (size_t)(q.nsites == 1 ? 4 : 16));
        out.push_back(q);
    }
    return out;
}

// ------------------------------------------------------------ drivers ------
struct TrajOut {
    std::vector<double> res;  // [obs][time]
    int64_t jumps = 0;
    Counters cnt;
    double w0 = 0.0, w1 = 0.0;  // wall clock at the start of the timed window and at the end
};

// Timed window for the CPU baseline: steps after the first g_window_warmup
// time steps of each trajectory (negative: whole trajectory).
int g_window_warmup = -1;
double wall_now() {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

// analog_tjm_2 / analog_tjm_1 (src/AnalogTJM.jl:104-176) for one trajectory.
void run_trajectory(const yaqs_analog_job *job, const MPSState &psi0, const std::vector<T4> &H,
                    const std::vector<Proc> *noise, const std::vector<Obs> &obs, int64_t t, TrajOut &out) {
    Cfg cfg{job->dt, (int)job->max_bond_dim, job->truncation_threshold, job->krylov_dim, job->krylov_tol};
    TdvpCtx C{&H, cfg, &out.cnt};
    TrajRng rng{job->seed, job->traj_offset + (uint64_t)t, 0};
    const int nobs = (int)obs.size();
    const int nt = (int)job->num_times;
    const int ncol = job->sample_timesteps ? nt : 1;
    out.res.assign((size_t)nobs * ncol, 0.0);
    auto measure = [&](MPSState &st, int col) {
        for (int q = 0; q < nobs; ++q) out.res[(size_t)q * ncol + col] = expect(st, obs[q]);
    };
    const double dt = job->dt;
    const int mb = (int)job->max_bond_dim;
    const double thr = job->truncation_threshold;
    MPSState state = psi0;
    out.w0 = wall_now();
    if (job->sample_timesteps) measure(state, 0);
    auto stoch = [&](MPSState &st, double dts) {
        if (noise) out.jumps += stochastic_process(st, *noise, dts, thr, mb, rng);
    };
    if (job->order == 1) {
        for (int j = 2; j <= nt; ++j) {
            if (j == g_window_warmup + 2) out.w0 = wall_now();
            tdvp_step(state, C, job->tdvp_sites);
            if (noise) {
                apply_dissipation(state, noise, dt, thr, mb);
                stoch(state, dt);
            }
            if (job->sample_timesteps) measure(state, j - 1);
            else if (j == nt) measure(state, 0);
        }
        out.w1 = wall_now();
        return;
    }
    // initialize (src/AnalogTJM.jl:21-33)
    apply_dissipation(state, noise, dt / 2, thr, mb);
    stoch(state, dt);
    auto sample = [&](int j) {  // src/AnalogTJM.jl:62-81
        MPSState psi = state;
        tdvp_step(psi, C, job->tdvp_sites);
        apply_dissipation(psi, noise, dt / 2, thr, mb);
        stoch(psi, dt);
        measure(psi, job->sample_timesteps ? j - 1 : 0);
    };
    if (job->sample_timesteps && nt > 1) sample(2);
    for (int j = 3; j <= nt; ++j) {
        if (j == g_window_warmup + 2) out.w0 = wall_now();
        tdvp_step(state, C, job->tdvp_sites);  // step_through (src/AnalogTJM.jl:40-55)
        apply_dissipation(state, noise, dt, thr, mb);
        stoch(state, dt);
        if (job->sample_timesteps || j == nt) sample(j);
    }
    out.w1 = wall_now();
}

// ----------------------------------------------------------- digital ------
struct Gate {
    int kind;
    int method = 0;  // yaqs_gate_method: the matrix is the gate (0) or the window Hamiltonian (1)
    int s0, s1;
    cd m[16];
};

// process_circuit greedy layering (src/DigitalTJM.jl:86-145).
void layer_circuit(const std::vector<Gate> &gates, std::vector<std::vector<int>> &layers,
                   std::vector<int> &sample_after, bool &sample_at_start) {
    layers.clear();
    std::vector<int> cur;
    std::vector<int> busy;
    std::vector<int> barrier_layers;  // layer index (count of finished layers) per sample barrier
    for (size_t g = 0; g < gates.size(); ++g) {
        const Gate &G = gates[g];
        if (G.kind == YAQS_GATE_SAMPLE) {
            if (!cur.empty()) { layers.push_back(cur); cur.clear(); busy.clear(); }
            barrier_layers.push_back((int)layers.size());
            continue;
        }
        std::vector<int> sites = {G.s0};
        if (G.kind == YAQS_GATE_2Q) sites.push_back(G.s1);
        bool overlap = false;
        for (int q : sites)
            if (std::find(busy.begin(), busy.end(), q) != busy.end()) overlap = true;
        if (overlap) { layers.push_back(cur); cur.clear(); busy.clear(); }
        cur.push_back((int)g);
        for (int q : sites) busy.push_back(q);
    }
    if (!cur.empty()) layers.push_back(cur);
    sample_at_start = false;
    sample_after.clear();
    for (int b : barrier_layers) {
        if (b == 0) sample_at_start = true;
        else if (std::find(sample_after.begin(), sample_after.end(), b) == sample_after.end())
            sample_after.push_back(b);
    }
    std::sort(sample_after.begin(), sample_after.end());
}

int max_bond_dim(const MPSState &s) {  // write_max_bond_dim (src/MPS.jl:163-172)
    int m = 0;
    for (auto &A : s.t) m = std::max(m, std::max(A.dl, A.dr));
    return m;
}

// apply_local_gate_exact! with rule T-D (src/DigitalTJM.jl:251-305).
void tebd_gate(MPSState &s, const cd *op, int s1, int s2, double thr, int max_bond) {
    shift_oc(s, s1);
    const T3 &A = s.t[s1 - 1], &B = s.t[s2 - 1];
    int l = A.dl, r = B.dr;
    cvec th = merge2(A, B);
    cvec thn = apply_op2(op, th, l, r);
    cvec U, Vt;
    std::vector<double> S;
    svd_thin(thn, l * 2, 2 * r, U, S, Vt);
    int ns = (int)S.size();
    double cur = 0;
    int keep = ns;
    for (int k = ns; k >= 1; --k) {
        if (cur + S[k - 1] * S[k - 1] > thr) break;
        cur += S[k - 1] * S[k - 1];
        keep -= 1;
    }
    keep = std::max(1, std::min(keep, max_bond));
    T3 nA(l, 2, keep), nB(keep, 2, r);
    for (size_t q = 0; q < (size_t)l * 2 * keep; ++q) nA.v[q] = U[q];
    for (int j = 0; j < 2 * r; ++j)
        for (int q = 0; q < keep; ++q) nB.v[q + (size_t)keep * j] = S[q] * Vt[q + (size_t)ns * j];
    s.t[s1 - 1] = nA;
    s.t[s2 - 1] = nB;
    s.oc = s2;
}

// apply_window! with local_method :TDVP on an adjacent pair (a, a+1)
// (src/DigitalTJM.jl:335-358).  With padding 0 the window is the pair: the
// centre moves to a, the window MPO is coeff g_a (x) g_b with 1x1 bonds
// (construct_window_mpo, :192-236; `h` here is that 4x4 operator in the
// gate-matrix index order) and the window state keeps the real outer bonds,
// whose environments are identities (_init_envs, src/Algorithms.jl:801-818).
// _tdvp_sweep_circuit_2site! on two sites is its edge step alone
// (src/Algorithms.jl:606-627, 762-800): Krylov exp(-i * 1.0 * H_eff) Θ with
// krylovdim 25, then split_mps_tensor_svd; A_a = U, A_b = S V†, centre b.
void window_tdvp_gate(MPSState &s, const cd *h, int a, int b, double thr, int max_bond) {
    shift_oc(s, a);
    const T3 &A = s.t[a - 1], &B = s.t[b - 1];
    const int l = A.dl, r = B.dr;
    cvec th = merge2(A, B);
    auto heff = [&](const cvec &x) { return apply_op2(h, x, l, r); };
    cvec thn = expm_krylov(heff, th, 1.0, 25, 1e-10);
    Cfg cfg;
    cfg.max_bond = max_bond;
    cfg.thr = thr;
    cvec U, Vt;
    std::vector<double> S;
    int keep = 0;
    split_svd(thn, l, 2, 2, r, cfg, U, S, Vt, keep);
    T3 nA(l, 2, keep), nB(keep, 2, r);
    for (size_t q = 0; q < (size_t)l * 2 * keep; ++q) nA.v[q] = U[q];
    for (int j = 0; j < 2 * r; ++j)
        for (int q = 0; q < keep; ++q) nB.v[q + (size_t)keep * j] = S[q] * Vt[q + (size_t)keep * j];
    s.t[a - 1] = nA;
    s.t[b - 1] = nB;
    s.oc = b;
}

// apply_window! :TDVP on a long-range pair a < b - 1 (src/DigitalTJM.jl:335-358):
// the window a..b with the window MPO (src/DigitalTJM.jl:192-236: coeff g_a on
// a, g_b on b, identities between, bonds of dimension 1) and the circuit
// sweep (src/Algorithms.jl:606-627): forward 2-site updates with +1 and the
// 1-site back-evolution by -1, then the edge step (+1, split right).  `h` is
// coeff g_a (x) g_b in the gate-matrix index order; its Kronecker factors
// (unique up to a scalar that cancels in every effective operator) give
// the two MPO tensors.
void window_tdvp_long(MPSState &s, const cd *h, int a, int b, double thr, int max_bond) {
    int im = 0;
    for (int q = 1; q < 16; ++q)
        if (std::abs(h[q]) > std::abs(h[im])) im = q;
    const int r0 = im % 4, c0 = im / 4;  // row p1 + 2 p2, column p1' + 2 p2'
    const int p1s = r0 % 2, p2s = r0 / 2, p1cs = c0 % 2, p2cs = c0 / 2;
    T4 Wa, Wb, Wi;
    for (T4 *W : {&Wa, &Wb, &Wi}) {
        W->dl = W->dr = 1;
        W->po = W->pi = 2;
        W->v.assign(4, 0.0);
    }
    Wi.v[0] = Wi.v[3] = 1.0;
    const cd piv = h[im];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
            Wa.v[x + 2 * y] = h[(x + 2 * p2s) + 4 * (y + 2 * p2cs)];          // A[x, y] B[p2s, p2cs]
            Wb.v[x + 2 * y] = h[(p1s + 2 * x) + 4 * (p1cs + 2 * y)] / piv;    // A[p1s, p1cs] B[x, y] / (A B)[piv]
        }
    shift_oc(s, a);
    const int Lw = b - a + 1;
    MPSState w;
    w.L = Lw;
    w.oc = 1;
    for (int q = a; q <= b; ++q) w.t.push_back(s.t[q - 1]);
    std::vector<T4> H(Lw, Wi);
    H[0] = Wa;
    H[Lw - 1] = Wb;
    TdvpCtx C;
    C.H = &H;
    C.cfg.dt = 1.0;
    C.cfg.max_bond = max_bond;
    C.cfg.thr = thr;
    C.cfg.kdim = 25;
    C.cfg.ktol = 1e-10;
    C.cnt = nullptr;
    std::vector<Env> EL(Lw + 1), ER(Lw + 1);
    const int rb = w.t[Lw - 1].dr, lb = w.t[0].dl;
    ER[Lw] = Env{rb, 1, rb, identity_env(rb, 1)};
    for (int i = Lw; i >= 2; --i) ER[i - 1] = update_right_env(w.t[i - 1], H[i - 1], ER[i]);
    EL[0] = Env{lb, 1, lb, identity_env(lb, 1)};
    for (int i = 1; i <= Lw - 2; ++i) two_site_forward(w, C, EL, ER, i, 1.0, true);
    two_site_forward(w, C, EL, ER, Lw - 1, 1.0, false);  // _two_site_update_edge_circuit!
    for (int q = 0; q < Lw; ++q) s.t[a - 1 + q] = w.t[q];
    s.oc = b;
}

void run_digital_trajectory(const yaqs_digital_job *job, const MPSState &psi0, const std::vector<Gate> &gates,
                            const std::vector<std::vector<int>> &layers, const std::vector<int> &sample_after,
                            bool sample_at_start, int nsamp, const std::vector<Proc> *noise,
                            const std::vector<Obs> &obs, int64_t t, std::vector<double> &res,
                            std::vector<int64_t> &bonds, int64_t &jumps) {
    TrajRng rng{job->seed, job->traj_offset + (uint64_t)t, 0};
    const int nobs = (int)obs.size();
    const double thr = job->truncation_threshold;
    const int mb = (int)job->max_bond_dim;
    MPSState st = psi0;
    res.assign((size_t)nobs * nsamp, 0.0);
    bonds.assign(nsamp, 0);
    int idx = 0;
    auto measure = [&](int k) {
        for (int q = 0; q < nobs; ++q) res[(size_t)q * nsamp + k] = expect(st, obs[q]);
        bonds[k] = max_bond_dim(st);
    };
    if (sample_at_start) measure(idx++);
    for (size_t li = 0; li < layers.size(); ++li) {
        for (int g : layers[li])
            if (gates[g].kind == YAQS_GATE_1Q) apply_1site_op(st.t[gates[g].s0 - 1], gates[g].m);
        for (int g : layers[li]) {
            const Gate &G = gates[g];
            if (G.kind != YAQS_GATE_2Q) continue;
            // apply_window! sorts the sites and applies the matrix unchanged
            // on (min, max) (src/DigitalTJM.jl:307-334)
            int a = std::min(G.s0, G.s1), b = std::max(G.s0, G.s1);
            if (b > a + 1 && G.method == YAQS_GATE_WINDOW_TDVP) {
                window_tdvp_long(st, G.m, a, b, thr, mb);
            } else if (b > a + 1) {
                // long_range_method :TEBD: SWAP network (src/DigitalTJM.jl:314-330)
                static const cd SW[16] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
                for (int k = b - 1; k >= a + 1; --k) tebd_gate(st, SW, k, k + 1, thr, mb);
                tebd_gate(st, G.m, a, a + 1, thr, mb);
                for (int k = a + 1; k <= b - 1; ++k) tebd_gate(st, SW, k, k + 1, thr, mb);
            } else if (G.method == YAQS_GATE_WINDOW_TDVP) {
                window_tdvp_gate(st, G.m, a, b, thr, mb);
            } else {
                tebd_gate(st, G.m, a, b, thr, mb);
            }
            if (noise && !noise->empty()) {
                std::vector<Proc> local;  // create_local_noise_model (:172-188)
                for (const Proc &p : *noise) {
                    if (p.nsites == 1 && (p.s0 == a || p.s0 == b)) local.push_back(p);
                    else if (p.nsites == 2 && ((p.s0 == a && p.s1 == b) || (p.s0 == b && p.s1 == a)))
                        local.push_back(p);
                }
                if (!local.empty()) {
                    apply_dissipation(st, &local, 1.0, thr, mb);
                    jumps += stochastic_process(st, local, 1.0, thr, mb, rng);
                } else {
                    normalize_mps(st);
                }
            } else {
                normalize_mps(st);
            }
        }
        int lnum = (int)li + 1;
        if (std::find(sample_after.begin(), sample_after.end(), lnum) != sample_after.end()) measure(idx++);
    }
    if (!sample_at_start && sample_after.empty() && nobs > 0) measure(0);
}

thread_local std::string g_err;
double g_last_window = 0.0;

}  // namespace

// ================================================================ C API ====
extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

double oracle_rng_uniform(uint64_t seed, uint64_t traj, uint64_t k) { return Philox::uniform(seed, traj, k); }

// Raw Philox4x32-10 block (for known-answer tests of the generator).
void oracle_philox_raw(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        Philox::round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

int64_t oracle_analog_num_samples(const yaqs_analog_job *job) {
    return job->sample_timesteps ? job->num_times : 1;
}

// Runs every trajectory of an analog job.  traj_out [obs][col][traj]
// (column-major (num_traj, num_cols) per observable), mean_out [obs][col].
// stats_out (may be NULL): [0] FLOP model, [1] 2-site matvecs, [2] 1-site
// matvecs, [3] SVDs, [4] env updates, [5] jumps.
int oracle_run_analog(const yaqs_analog_job *job, double *traj_out, double *mean_out, int nthreads,
                      double *stats_out) {
    if (!job || job->L < 1 || !job->mps || !job->mpo || job->num_traj < 1 || job->num_times < 1) {
        g_err = "invalid analog job";
        return YAQS_EINVAL;
    }
    MPSState psi0 = load_mps(job->L, job->mps);
    std::vector<T4> H(job->L);
    for (int64_t i = 0; i < job->L; ++i) {
        const yaqs_mpo_tensor &w = job->mpo[i];
        H[i].dl = (int)w.dl; H[i].po = (int)w.dout; H[i].pi = (int)w.din; H[i].dr = (int)w.dr;
        H[i].v.resize((size_t)w.dl * w.dout * w.din * w.dr);
        load_c(w.data, H[i].v.data(), H[i].v.size());
    }
    std::vector<Proc> noise_v = load_noise(job->has_noise_model ? job->n_noise : 0, job->noise);
    const std::vector<Proc> *noise = job->has_noise_model ? &noise_v : nullptr;
    std::vector<Obs> obs = load_obs(job->n_obs, job->obs);
    const int64_t ntraj = job->num_traj;
    const int64_t ncol = oracle_analog_num_samples(job);
    const int64_t nobs = job->n_obs;
    std::vector<TrajOut> outs(ntraj);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t t = 0; t < ntraj; ++t) run_trajectory(job, psi0, H, noise, obs, t, outs[t]);
    double st[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t q = 0; q < nobs; ++q)
        for (int64_t c = 0; c < ncol; ++c) {
            double s = 0;
            for (int64_t t = 0; t < ntraj; ++t) {
                double v = outs[t].res[(size_t)q * ncol + c];
                if (traj_out) traj_out[((size_t)q * ncol + c) * ntraj + t] = v;
                s += v;
            }
            if (mean_out) mean_out[(size_t)q * ncol + c] = s / (double)ntraj;
        }
    for (auto &o : outs) {
        st[0] += o.cnt.flop; st[1] += o.cnt.matvec2; st[2] += o.cnt.matvec1;
        st[3] += o.cnt.svd; st[4] += o.cnt.env; st[5] += o.jumps;
    }
    if (stats_out) std::copy(st, st + 6, stats_out);
    g_last_window = 0.0;
    for (auto &o : outs) g_last_window = std::max(g_last_window, o.w1 - o.w0);
    return YAQS_OK;
}

// CPU-baseline timing: time steps warmup+1 .. end of every trajectory of the
// next oracle_run_analog call (warmup < 0: whole trajectories); the window
// reported is the longest over trajectories (they run concurrently, one per
// thread).
void oracle_set_window(int warmup_steps) { g_window_warmup = warmup_steps; }

// Binds zgemm (symbol `sym`, Fortran ABI with 64-bit integers) from the
// shared library at `path`, and sets its thread count to one through
// `threads_sym` when given.  path == NULL unbinds.  Returns 0 on success.
int oracle_use_blas(const char *path, const char *sym, const char *threads_sym) {
    if (!path) {
        g_zgemm = nullptr;
        g_zgesdd = nullptr;
        return 0;
    }
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    void *f = dlsym(h, sym);
    if (!f) return 2;
    if (threads_sym) {
        typedef void (*setn_fn)(int64_t);
        if (void *t = dlsym(h, threads_sym)) reinterpret_cast<setn_fn>(t)(1);
    }
    g_zgemm = reinterpret_cast<zgemm64_fn>(f);
    return 0;
}

// Binds zgesdd (symbol `sym`, Fortran ABI with 64-bit integers) from the
// library at `path` for the SVDs; NULL unbinds.  Returns 0 on success.
int oracle_use_lapack_svd(const char *path, const char *sym) {
    if (!path) {
        g_zgesdd = nullptr;
        return 0;
    }
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    void *f = dlsym(h, sym);
    if (!f) return 2;
    g_zgesdd = reinterpret_cast<zgesdd64_fn>(f);
    return 0;
}

int oracle_blas_bound(void) { return g_zgemm ? 1 : 0; }

void oracle_set_drop_dead(int on) { g_drop_dead = on != 0; }
double oracle_last_window_seconds(void) { return g_last_window; }

int64_t oracle_digital_num_samples(const yaqs_digital_job *job) {
    std::vector<Gate> gates(job->n_gates);
    for (int64_t g = 0; g < job->n_gates; ++g) {
        gates[g].kind = job->gates[g].kind;
        gates[g].s0 = (int)job->gates[g].site0;
        gates[g].s1 = (int)job->gates[g].site1;
    }
    std::vector<std::vector<int>> layers;
    std::vector<int> sa;
    bool s0;
    layer_circuit(gates, layers, sa, s0);
    int64_t n = (s0 ? 1 : 0) + (int64_t)sa.size();
    return n == 0 ? 1 : n;
}

int oracle_run_digital(const yaqs_digital_job *job, double *traj_out, double *mean_out, int64_t *bond_out,
                       int nthreads, double *stats_out) {
    if (!job || job->L < 1 || !job->mps || job->num_traj < 1) {
        g_err = "invalid digital job";
        return YAQS_EINVAL;
    }
    MPSState psi0 = load_mps(job->L, job->mps);
    std::vector<Gate> gates(job->n_gates);
    for (int64_t g = 0; g < job->n_gates; ++g) {
        const yaqs_gate &G = job->gates[g];
        gates[g].kind = G.kind;
        gates[g].method = G.method;
        if (G.method != YAQS_GATE_EXACT && (G.method != YAQS_GATE_WINDOW_TDVP || G.kind != YAQS_GATE_2Q)) {
            g_err = "bad gate method";
            return YAQS_EINVAL;
        }
        gates[g].s0 = (int)G.site0;
        gates[g].s1 = (int)G.site1;
        if (G.kind == YAQS_GATE_1Q) load_c(G.matrix, gates[g].m, 4);
        if (G.kind == YAQS_GATE_2Q) {
            load_c(G.matrix, gates[g].m, 16);
            if (G.site1 == G.site0) {
                g_err = "2-qubit gate on a single site";
                return YAQS_EINVAL;
            }
        }
    }
    std::vector<std::vector<int>> layers;
    std::vector<int> sample_after;
    bool sample_at_start;
    layer_circuit(gates, layers, sample_after, sample_at_start);
    int nsamp = (sample_at_start ? 1 : 0) + (int)sample_after.size();
    if (nsamp == 0) nsamp = 1;
    std::vector<Proc> noise_v = load_noise(job->has_noise_model ? job->n_noise : 0, job->noise);
    const std::vector<Proc> *noise = job->has_noise_model ? &noise_v : nullptr;
    std::vector<Obs> obs = load_obs(job->n_obs, job->obs);
    const int64_t ntraj = job->num_traj, nobs = job->n_obs;
    std::vector<std::vector<double>> res(ntraj);
    std::vector<std::vector<int64_t>> bonds(ntraj);
    std::vector<int64_t> jumps(ntraj, 0);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t t = 0; t < ntraj; ++t)
        run_digital_trajectory(job, psi0, gates, layers, sample_after, sample_at_start, nsamp, noise, obs, t,
                               res[t], bonds[t], jumps[t]);
    for (int64_t q = 0; q < nobs; ++q)
        for (int c = 0; c < nsamp; ++c) {
            double s = 0;
            for (int64_t t = 0; t < ntraj; ++t) {
                double v = res[t][(size_t)q * nsamp + c];
                if (traj_out) traj_out[((size_t)q * nsamp + c) * ntraj + t] = v;
                s += v;
            }
            if (mean_out) mean_out[(size_t)q * nsamp + c] = s / (double)ntraj;
        }
    if (bond_out)
        for (int64_t t = 0; t < ntraj; ++t)
            for (int c = 0; c < nsamp; ++c) bond_out[(size_t)t * nsamp + c] = bonds[t][c];
    if (stats_out) {
        double j = 0;
        for (auto x : jumps) j += x;
        stats_out[0] = j;
    }
    return YAQS_OK;
}

// ---- unit entry points for the oracle's own pinning tests (column-major,
// interleaved complex) -------------------------------------------------------
void oracle_svd(int m, int n, const double *A, double *U, double *S, double *Vt) {
    cvec a((size_t)m * n);
    load_c(A, a.data(), a.size());
    cvec u, vt;
    std::vector<double> s;
    svd_thin(a, m, n, u, s, vt);
    std::memcpy(U, u.data(), u.size() * sizeof(cd));
    std::memcpy(S, s.data(), s.size() * sizeof(double));
    std::memcpy(Vt, vt.data(), vt.size() * sizeof(cd));
}

int oracle_qr(int m, int n, const double *A, double *Q, double *R) {
    cvec a((size_t)m * n), q, r;
    load_c(A, a.data(), a.size());
    int rr;
    qr_thin(a, m, n, q, r, rr);
    std::memcpy(Q, q.data(), q.size() * sizeof(cd));
    std::memcpy(R, r.data(), r.size() * sizeof(cd));
    return rr;
}

int oracle_lq(int m, int n, const double *A, double *L, double *Q) {
    cvec a((size_t)m * n), l, q;
    load_c(A, a.data(), a.size());
    int rr;
    lq_thin(a, m, n, l, q, rr);
    std::memcpy(L, l.data(), l.size() * sizeof(cd));
    std::memcpy(Q, q.data(), q.size() * sizeof(cd));
    return rr;
}

// exp(-i dt H) v for a dense column-major Hermitian H (n x n) by the
// oracle's Lanczos; returns the number of Krylov steps.
int oracle_expm_krylov_dense(int n, const double *H, const double *v, double dt, int kmax, double tol,
                             double *out) {
    cvec h((size_t)n * n), vv(n);
    load_c(H, h.data(), h.size());
    load_c(v, vv.data(), n);
    auto f = [&](const cvec &x) {
        cvec y(n);
        gemm('N', 'N', n, 1, n, h.data(), n, x.data(), n, y.data(), n);
        return y;
    };
    int iters = 0;
    cvec o = expm_krylov(f, vv, dt, kmax, tol, &iters);
    std::memcpy(out, o.data(), (size_t)n * sizeof(cd));
    return iters;
}

void oracle_tridiag_exp(int k, const double *alpha, const double *beta, double tau, double *out) {
    std::vector<double> a(alpha, alpha + k), b(k, 0.0);
    for (int i = 1; i < k; ++i) b[i] = beta[i];
    cvec c;
    tridiag_exp_e0(a, b, k, tau, c);
    std::memcpy(out, c.data(), (size_t)k * sizeof(cd));
}

// H_eff · x for one site/pair (project_site); all tensors column-major in
// the reference's index order: x (kl, p, kr), L (bl, ml, kl), W (ml, po, pi,
// mr), R (br, mr, kr); out (bl, po, br).
void oracle_project_site(int kl, int p, int kr, int ml, int mr, const double *x, const double *Lv,
                         const double *Wv, const double *Rv, double *out) {
    Env Le{kl, ml, kl, cvec((size_t)kl * ml * kl)}, Re{kr, mr, kr, cvec((size_t)kr * mr * kr)};
    load_c(Lv, Le.v.data(), Le.v.size());
    load_c(Rv, Re.v.data(), Re.v.size());
    T4 W;
    W.dl = ml; W.po = p; W.pi = p; W.dr = mr;
    W.v.resize((size_t)ml * p * p * mr);
    load_c(Wv, W.v.data(), W.v.size());
    cvec xv((size_t)kl * p * kr);
    load_c(x, xv.data(), xv.size());
    cvec o = project_site(xv, kl, p, kr, Le, Re, W);
    std::memcpy(out, o.data(), o.size() * sizeof(cd));
}

void oracle_env_left(int al, int p, int ar, int ml, int mr, const double *A, const double *Wv, const double *Ev,
                     double *out) {
    T3 a(al, p, ar);
    load_c(A, a.v.data(), a.v.size());
    T4 W;
    W.dl = ml; W.po = p; W.pi = p; W.dr = mr;
    W.v.resize((size_t)ml * p * p * mr);
    load_c(Wv, W.v.data(), W.v.size());
    Env E{al, ml, al, cvec((size_t)al * ml * al)};
    load_c(Ev, E.v.data(), E.v.size());
    Env En = update_left_env(a, W, E);
    std::memcpy(out, En.v.data(), En.v.size() * sizeof(cd));
}

void oracle_env_right(int al, int p, int ar, int ml, int mr, const double *A, const double *Wv, const double *Ev,
                      double *out) {
    T3 a(al, p, ar);
    load_c(A, a.v.data(), a.v.size());
    T4 W;
    W.dl = ml; W.po = p; W.pi = p; W.dr = mr;
    W.v.resize((size_t)ml * p * p * mr);
    load_c(Wv, W.v.data(), W.v.size());
    Env E{ar, mr, ar, cvec((size_t)ar * mr * ar)};
    load_c(Ev, E.v.data(), E.v.size());
    Env En = update_right_env(a, W, E);
    std::memcpy(out, En.v.data(), En.v.size() * sizeof(cd));
}

// dp = 1 - <ψ|ψ> (calculate_stochastic_factor, src/StochasticProcess.jl:35-38)
double oracle_stochastic_factor(int64_t L, const yaqs_mps_tensor *t) {
    MPSState s = load_mps(L, t);
    return 1.0 - scalar_product(s, s).real();
}

// create_probability_distribution on a given state; returns the number of
// candidates, writes probs and (type, site, proc) triples.
int oracle_prob_distribution(int64_t L, const yaqs_mps_tensor *t, int64_t n_noise,
                             const yaqs_noise_process *noise, double dt, double *probs, int *cand) {
    MPSState s = load_mps(L, t);
    std::vector<Proc> nv = load_noise(n_noise, noise);
    std::vector<double> p;
    std::vector<Candidate> c;
    prob_distribution(s, nv, dt, p, c);
    for (size_t i = 0; i < p.size(); ++i) {
        probs[i] = p[i];
        cand[3 * i] = c[i].type;
        cand[3 * i + 1] = c[i].site;
        cand[3 * i + 2] = c[i].proc;
    }
    return (int)p.size();
}

// apply_mpo_jump! of one long-range process (mode 0) or MPS.truncate! alone
// (mode 1, `proc` unused) on a copy of the MPS; writes (dl, d, dr) per site
// to dims and the tensors (interleaved, column-major, one after the other)
// to data.  Returns the doubles written, or -1 when `cap` is too small.
int64_t oracle_mpo_jump(int64_t L, const yaqs_mps_tensor *mps, const yaqs_noise_process *proc, double thr,
                        int max_bond, int mode, int64_t *dims, double *data, int64_t cap) {
    MPSState s = load_mps(L, mps);
    s.oc = (int)L;  // the centre field create_probability_distribution leaves
    if (mode == 0) {
        std::vector<Proc> p = load_noise(1, proc);
        apply_mpo_jump(s, p[0], thr, max_bond);
    } else {
        truncate_mps(s, thr, max_bond);
    }
    int64_t n = 0;
    for (int i = 0; i < s.L; ++i) {
        dims[3 * i] = s.t[i].dl;
        dims[3 * i + 1] = s.t[i].d;
        dims[3 * i + 2] = s.t[i].dr;
        if (n + 2 * (int64_t)s.t[i].v.size() > cap) return -1;
        for (const cd &z : s.t[i].v) {
            data[n++] = z.real();
            data[n++] = z.imag();
        }
    }
    return n;
}

// measure_shots (src/MPS.jl:742-749): `shots` calls of single_shot_measure
// (:683-738) on copies of ψ (taken as given, centre at site 1, as the
// reference's shift_orthogonality_center!(psi, 1) leaves a state whose
// orth_center field is 1): per site the slice norms of the current tensor
// are the outcome weights, the draw is the Philox uniform (seed, shot, site),
// the outcome the first with r <= cumulative weight (else the last), the
// projected slice is divided by sqrt(weight) when that exceeds 1e-12 and
// contracted into the next site.  bits: site i (1-based) in bit i-1.
int oracle_measure_shots(int64_t L, const yaqs_mps_tensor *mps, int64_t shots, uint64_t seed, int64_t *bits) {
    const MPSState s0 = load_mps(L, mps);
    for (int64_t t = 0; t < shots; ++t) {
        cvec v(1, cd(1.0, 0.0));  // the projected row carried into the next site (L = 1 after site 1)
        int64_t b = 0;
        for (int i = 0; i < s0.L; ++i) {
            const T3 &A = s0.t[i];
            const int R = A.dr, d = A.d, Lb = A.dl;
            // current tensor row: C[p][r] = sum_a v[a] A[a, p, r]
            cvec C((size_t)d * R, 0.0);
            for (int p = 0; p < d; ++p)
                for (int r = 0; r < R; ++r) {
                    cd acc = 0;
                    for (int a = 0; a < Lb; ++a) acc += v[a] * A.v[a + (size_t)Lb * (p + (size_t)d * r)];
                    C[p + (size_t)d * r] = acc;
                }
            std::vector<double> probs(d, 0.0);
            for (int p = 0; p < d; ++p)
                for (int r = 0; r < R; ++r) probs[p] += std::norm(C[p + (size_t)d * r]);
            const double rr = Philox::uniform(seed, (uint64_t)t, (uint64_t)i);
            double accum = 0;
            int chosen = d - 1;
            for (int k = 0; k < d; ++k) {
                accum += probs[k];
                if (rr <= accum) { chosen = k; break; }
            }
            b += (int64_t)chosen << i;
            const double nrm = std::sqrt(probs[chosen]);
            v.assign(R, 0.0);
            for (int r = 0; r < R; ++r) v[r] = C[chosen + (size_t)d * r] * (nrm > 1e-12 ? 1.0 / nrm : 1.0);
        }
        bits[t] = b;
    }
    return 0;
}

// Noise-free TDVP: `steps` sweeps of two_site_tdvp! (sites = 2) or
// single_site_tdvp! (sites = 1) from the given state; returns
// scalar_product(ψ(t), ψ0) = <ψ0|ψ(t)> (src/MPS.jl:528-585) in out[0..1]
// (test/test_Algorithms.jl:37-93).
int oracle_tdvp_overlap(int64_t L, const yaqs_mps_tensor *mps, const yaqs_mpo_tensor *mpo, double dt, int steps,
                        int sites, int max_bond, double thr, double *out) {
    MPSState s = load_mps(L, mps);
    const MPSState s0 = s;
    std::vector<T4> H(L);
    for (int64_t i = 0; i < L; ++i) {
        const yaqs_mpo_tensor &w = mpo[i];
        H[i].dl = (int)w.dl; H[i].po = (int)w.dout; H[i].pi = (int)w.din; H[i].dr = (int)w.dr;
        H[i].v.resize((size_t)w.dl * w.dout * w.din * w.dr);
        load_c(w.data, H[i].v.data(), H[i].v.size());
    }
    Cfg cfg{dt, max_bond, thr, 25, 1e-10};
    TdvpCtx C{&H, cfg, nullptr};
    for (int k = 0; k < steps; ++k) tdvp_step(s, C, sites);
    const cd ov = scalar_product(s, s0);
    out[0] = ov.real();
    out[1] = ov.imag();
    return 0;
}

// Noise-free 2-site TDVP: `steps` sweeps of two_site_tdvp! from the given
// state, measuring 1-site observables after each (test/benchmark_tdvp.jl:196-206).
int oracle_tdvp_evolve(int64_t L, const yaqs_mps_tensor *mps, const yaqs_mpo_tensor *mpo, double dt, int steps,
                       int max_bond, double thr, int n_obs, const yaqs_observable *obs, double *out) {
    MPSState s = load_mps(L, mps);
    std::vector<T4> H(L);
    for (int64_t i = 0; i < L; ++i) {
        const yaqs_mpo_tensor &w = mpo[i];
        H[i].dl = (int)w.dl; H[i].po = (int)w.dout; H[i].pi = (int)w.din; H[i].dr = (int)w.dr;
        H[i].v.resize((size_t)w.dl * w.dout * w.din * w.dr);
        load_c(w.data, H[i].v.data(), H[i].v.size());
    }
    std::vector<Obs> ob = load_obs(n_obs, obs);
    Cfg cfg{dt, max_bond, thr, 25, 1e-10};
    TdvpCtx C{&H, cfg, nullptr};
    for (int q = 0; q < n_obs; ++q) out[(size_t)q * (steps + 1)] = expect(s, ob[q]);
    for (int k = 1; k <= steps; ++k) {
        two_site_tdvp(s, C);
        for (int q = 0; q < n_obs; ++q) out[(size_t)q * (steps + 1) + k] = expect(s, ob[q]);
        if (std::getenv("ORACLE_DEBUG_BONDS")) {
            for (int i = 0; i < s.L; ++i) fprintf(stderr, "%d ", s.t[i].dr);
            fprintf(stderr, "\n");
        }
    }
    return 0;
}

}  // extern "C"
