// Device kernels of the batched TJM engine (gfx950).
//
// Every kernel works on a batch of trajectories laid out slot by slot:
// trajectory t of a buffer starts at base + t * stride.  Storage precision T
// is float (mixed mode) or double; norms, Lanczos scalars, Gram matrices,
// eigen-decompositions, transfer matrices and jump probabilities are always
// computed in double.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "reduce.h"
#include "rng.h"

#include <hip/hip_runtime.h>

namespace yq {

// ---------------------------------------------------------------- helpers --
template <typename T>
__device__ __forceinline__ double2 ld2(const cplx<T> &z) {
    return make_double2((double)z.re, (double)z.im);
}
template <typename T>
__device__ __forceinline__ void st2(cplx<T> &z, double2 v) {
    z.re = (T)v.x;
    z.im = (T)v.y;
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }


// Block-wide sum of a double (blockDim.x multiple of 64, <= 1024).
__device__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < nw; ++i) r += sh[i];
    __syncthreads();
    return r;
}

__device__ double2 block_sum2(double2 v, double *sh) {
    double a = block_sum(v.x, sh);
    double b = block_sum(v.y, sh);
    return make_double2(a, b);
}

// ------------------------------------------------------------------ GEMM ---
// Generic batched complex GEMM, 64x64 tile per 256-thread workgroup, each
// thread accumulating a 4x4 block in Tacc.  Used for every contraction whose
// shape or precision mix the MFMA kernels do not cover.
template <typename TA, typename TB, typename TC, typename Tacc>
__global__ __launch_bounds__(256) void gemm_generic_kernel(GemmDesc d, const cplx<TA> *__restrict__ A,
                                                           const cplx<TB> *__restrict__ B, cplx<TC> *__restrict__ C) {
    constexpr int BM = 64, BN = 64, BK = 16;
    const int t = blockIdx.z, jb = blockIdx.y;
    if (d.active && !d.active[t]) return;
    const int tilesN = (d.N + BN - 1) / BN;
    const int m0 = (blockIdx.x / tilesN) * BM, n0 = (blockIdx.x % tilesN) * BN;
    if (d.lower && n0 >= m0 + BM) return;  // the tile lies above the diagonal
    const cplx<TA> *Ab = A + t * d.sA_outer + jb * d.sA_inner;
    const cplx<TB> *Bb = B + t * d.sB_outer + jb * d.sB_inner;
    cplx<TC> *Cb = C + t * d.sC_outer + jb * d.sC_inner;
    __shared__ Tacc As_r[BK][BM + 1], As_i[BK][BM + 1], Bs_r[BK][BN + 1], Bs_i[BK][BN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    Tacc acc_r[4][4], acc_i[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc_r[i][j] = acc_i[i][j] = (Tacc)0;
    for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + 256 * r;
            // A tile: element (row, kk)
            int row, kk;
            if (d.opA == OP_N) { row = e / BK; kk = e % BK; }
            else { kk = e / BM; row = e % BM; }
            const int gi = m0 + row, gk = k0 + kk;
            Tacc vr = 0, vi = 0;
            if (gi < d.M && gk < d.K) {
                cplx<TA> z = (d.opA == OP_N) ? Ab[(int64_t)gi * d.lda + gk] : Ab[(int64_t)gk * d.lda + gi];
                vr = (Tacc)z.re;
                vi = (d.opA == OP_C) ? -(Tacc)z.im : (Tacc)z.im;
            }
            As_r[kk][row] = vr;
            As_i[kk][row] = vi;
            // B tile: element (kk, col)
            int col;
            if (d.opB == OP_N) { kk = e / BN; col = e % BN; }
            else { col = e / BK; kk = e % BK; }
            const int gj = n0 + col, gk2 = k0 + kk;
            Tacc wr = 0, wi = 0;
            if (gj < d.N && gk2 < d.K) {
                cplx<TB> z = (d.opB == OP_N) ? Bb[(int64_t)gk2 * d.ldb + gj] : Bb[(int64_t)gj * d.ldb + gk2];
                wr = (Tacc)z.re;
                wi = (d.opB == OP_C) ? -(Tacc)z.im : (Tacc)z.im;
            }
            Bs_r[kk][col] = wr;
            Bs_i[kk][col] = wi;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            Tacc ar[4], ai[4], br[4], bi[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[i] = As_r[kk][ty * 4 + i];
                ai[i] = As_i[kk][ty * 4 + i];
                br[i] = Bs_r[kk][tx * 4 + i];
                bi[i] = Bs_i[kk][tx * 4 + i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc_r[i][j] += ar[i] * br[j] - ai[i] * bi[j];
                    acc_i[i][j] += ar[i] * bi[j] + ai[i] * br[j];
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = m0 + ty * 4 + i;
        if (gi >= d.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = n0 + tx * 4 + j;
            if (gj >= d.N) continue;
            cplx<TC> &c = Cb[(int64_t)gi * d.ldc + gj];
            if (d.accumulate) {
                c.re = (TC)((Tacc)c.re + acc_r[i][j]);
                c.im = (TC)((Tacc)c.im + acc_i[i][j]);
            } else {
                c.re = (TC)acc_r[i][j];
                c.im = (TC)acc_i[i][j];
            }
        }
    }
}

template <typename TA, typename TB, typename TC, typename Tacc>
void gemm(const GemmDesc &d, const cplx<TA> *A, const cplx<TB> *B, cplx<TC> *C, hipStream_t s) {
    if (d.M <= 0 || d.N <= 0 || d.outer <= 0 || d.inner <= 0) return;
    if (d.K <= 0) throw HipError("gemm with K = 0");
    const int tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64);
    dim3 grid(tiles, d.inner, d.outer);
    hipLaunchKernelGGL((gemm_generic_kernel<TA, TB, TC, Tacc>), grid, dim3(256), 0, s, d, A, B, C);
    YQ_HIP(hipGetLastError());
}

template void gemm<float, float, float, float>(const GemmDesc &, const c32 *, const c32 *, c32 *, hipStream_t);
template void gemm<double, double, double, double>(const GemmDesc &, const c64 *, const c64 *, c64 *, hipStream_t);
template void gemm<float, float, double, double>(const GemmDesc &, const c32 *, const c32 *, c64 *, hipStream_t);
template void gemm<double, float, double, double>(const GemmDesc &, const c64 *, const c32 *, c64 *, hipStream_t);
template void gemm<float, double, double, double>(const GemmDesc &, const c32 *, const c64 *, c64 *, hipStream_t);
template void gemm<double, double, float, double>(const GemmDesc &, const c64 *, const c64 *, c32 *, hipStream_t);

// ------------------------------------------------------------- W step ------
// Y[t][p][j][q] = Σ_i W[j][i] X[t][p][i][q] with W in CSR (rows j).  The MPO
// contraction of the effective Hamiltonian; memory-bound, coalesced along q.
// Accumulated in the contraction precision T (W rounded to T), the order the
// fused GEMM epilogue repeats.
template <typename T>
__global__ __launch_bounds__(256) void wstep_kernel(const cplx<T> *__restrict__ X, cplx<T> *__restrict__ Y, int P, int I,
                                                    int J, int Q, int64_t sX, int64_t sY, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const c64 *__restrict__ val,
                                                    const int *active) {
    const int t = blockIdx.y;
    if (active && !active[t]) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)P * Q) return;
    const int p = (int)(idx / Q), q = (int)(idx % Q);
    const cplx<T> *x = X + t * sX + (int64_t)p * I * Q + q;
    cplx<T> *y = Y + t * sY + (int64_t)p * J * Q + q;
    for (int j = 0; j < J; ++j) {
        T ax = (T)0, ay = (T)0;
        for (int z = rowptr[j]; z < rowptr[j + 1]; ++z) {
            const c64 w = val[z];
            const T wr = (T)w.re, wi = (T)w.im;
            const cplx<T> xv = x[(int64_t)col[z] * Q];
            // explicit FMAs: the fused GEMM epilogue repeats this order exactly
            ax = fma_t(wr, xv.re, ax);
            ax = fma_t(-wi, xv.im, ax);
            ay = fma_t(wr, xv.im, ay);
            ay = fma_t(wi, xv.re, ay);
        }
        y[(int64_t)j * Q] = cmake<T>(ax, ay);
    }
}

template <typename T>
void wstep(const cplx<T> *X, cplx<T> *Y, int P, int I, int J, int Q, int64_t sX, int64_t sY, const int *rowptr,
           const int *col, const c64 *val, const int *active, int B, hipStream_t s) {
    const int64_t n = (int64_t)P * Q;
    dim3 grid((unsigned)((n + 255) / 256), B);
    hipLaunchKernelGGL(wstep_kernel<T>, grid, dim3(256), 0, s, X, Y, P, I, J, Q, sX, sY, rowptr, col, val, active);
    YQ_HIP(hipGetLastError());
}
template void wstep<float>(const c32 *, c32 *, int, int, int, int, int64_t, int64_t, const int *, const int *,
                           const c64 *, const int *, int, hipStream_t);
template void wstep<double>(const c64 *, c64 *, int, int, int, int, int64_t, int64_t, const int *, const int *,
                            const c64 *, const int *, int, hipStream_t);

// ------------------------------------------------------------- Lanczos -----
// lz_init: β0 = ‖x‖, V0 = x/β0, active = β0 > 0; block 0 clears the
// per-iteration active counters nactive[0 .. nslots).
template <typename T>
__global__ __launch_bounds__(256) void lz_init_kernel(const cplx<T> *x, int64_t sx, cplx<T> *V, int64_t sV, int64_t n,
                                                      double *beta0, int *active, int *nit, int *nactive, int nslots) {
    __shared__ double sh[16];
    const int t = blockIdx.x;
    if (t == 0)
        for (int i = threadIdx.x; i < nslots; i += blockDim.x) nactive[i] = 0;
    const cplx<T> *xt = x + t * sx;
    cplx<T> *v = V + t * sV;
    double s = 0;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        double2 z = ld2(xt[e]);
        s += z.x * z.x + z.y * z.y;
    }
    s = block_sum(s, sh);
    const double b0 = sqrt(s);
    if (threadIdx.x == 0) {
        beta0[t] = b0;
        active[t] = b0 > 0.0 ? 1 : 0;
        nit[t] = 0;
    }
    if (b0 > 0.0) {
        const double inv = 1.0 / b0;
        for (int64_t e = threadIdx.x; e < n; e += blockDim.x) st2(v[e], cscale(ld2(xt[e]), inv));
    }
}

// lz_dots: part[t][q][c] = Σ_{e in chunk c} conj(V_q[e]) w[e] for q <= j,
// w = V_{j+1} (the raw matvec output).
constexpr int LZ_CHUNK = 2048;  // elements per block (256 threads x 4 pairs)

// Krylov vectors are read and written as pairs of consecutive elements
// (one 16-byte access per lane for complex-FP32): every vector length n is
// even (2 or 4 physical indices) and every vector starts at an even
// element (even nmax and slot strides).
template <typename T>
struct LzPair;
template <>
struct LzPair<float> {
    static __device__ __forceinline__ void ld(const c32 *p, int64_t e2, double2 &a, double2 &b) {
        const float4 v = reinterpret_cast<const float4 *>(p)[e2];
        a = make_double2(v.x, v.y);
        b = make_double2(v.z, v.w);
    }
    static __device__ __forceinline__ void st(c32 *p, int64_t e2, double2 a, double2 b) {
        reinterpret_cast<float4 *>(p)[e2] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
};
template <>
struct LzPair<double> {
    static __device__ __forceinline__ void ld(const c64 *p, int64_t e2, double2 &a, double2 &b) {
        const double2 *q = reinterpret_cast<const double2 *>(p) + 2 * e2;
        a = q[0];
        b = q[1];
    }
    static __device__ __forceinline__ void st(c64 *p, int64_t e2, double2 a, double2 b) {
        double2 *q = reinterpret_cast<double2 *>(p) + 2 * e2;
        q[0] = a;
        q[1] = b;
    }
};

// Deferred normalisation: the basis is stored unnormalised, V_q = v_q / s_q
// with s_0 = 1 (lz_init stores v_0) and s_q = 1 / β_q (β_q = beta[q], the
// norm of the q-th orthogonalised matvec output, stored as it is).
__device__ __forceinline__ double lz_scale_of(const double *be, int q) { return q == 0 ? 1.0 : 1.0 / be[q]; }

// Each wavefront reduces its partial of every q with DPP and parks it in
// LDS; one barrier, then lanes q <= j add the four wavefront partials.
template <typename T>
__global__ __launch_bounds__(256) void lz_dots_kernel(const cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j,
                                                      c64 *part, int nchunk, int kslots, const int *active) {
    __shared__ double2 sh[4][64];
    const int t = blockIdx.y, c = blockIdx.x;
    if (!active[t]) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const cplx<T> *Vt = V + t * sV;
    const cplx<T> *w = Vt + (int64_t)(j + 1) * nmax;
    const int64_t n2 = n >> 1, base = (int64_t)c * (LZ_CHUNK / 2);
    double2 wa[4], wb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t e2 = base + threadIdx.x + 256 * r;
        wa[r] = wb[r] = make_double2(0, 0);
        if (e2 < n2) LzPair<T>::ld(w, e2, wa[r], wb[r]);
    }
    for (int q = 0; q <= j; ++q) {
        const cplx<T> *vq = Vt + (int64_t)q * nmax;
        double2 acc = make_double2(0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t e2 = base + threadIdx.x + 256 * r;
            if (e2 < n2) {
                double2 va, vb;
                LzPair<T>::ld(vq, e2, va, vb);
                acc = cadd(acc, cmulc(va, wa[r]));
                acc = cadd(acc, cmulc(vb, wb[r]));
            }
        }
        acc.x = wave_sum(acc.x);
        acc.y = wave_sum(acc.y);
        if (lane == 0) sh[wid][q] = acc;
    }
    __syncthreads();
    if (threadIdx.x <= j) {
        const int q = threadIdx.x;
        c64 o;
        o.re = sh[0][q].x + sh[1][q].x + sh[2][q].x + sh[3][q].x;
        o.im = sh[0][q].y + sh[1][q].y + sh[2][q].y + sh[3][q].y;
        part[((int64_t)t * kslots + q) * nchunk + c] = o;
    }
}

// lz_update: g_q = Σ_c part; w -= Σ_q g_q V_q; partn[t][c] = ‖w_chunk‖².
// q runs outside the pairs, so each basis vector is read by four
// independent loads per lane (the subtraction order per element is q
// ascending either way).
template <typename T>
__global__ __launch_bounds__(256) void lz_update_kernel(cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j,
                                                        const c64 *part, int nchunk, int kslots, double *partn,
                                                        double *galpha, const int *active, const double *beta,
                                                        int kmax) {
    __shared__ double sh[16];
    __shared__ double2 g[64];
    const int t = blockIdx.y, c = blockIdx.x;
    if (!active[t]) return;
    // stored V_q = v_q / s_q, w = H V_j: <v_q, H v_j> = s_q s_j <V_q, w>, and
    // H v_j - Σ <v_q, H v_j> v_q = s_j w - Σ h_q V_q with h_q = s_q <v_q, H v_j>
    const double *be = beta + (int64_t)t * (kmax + 1);
    const double sj = lz_scale_of(be, j);
    if (threadIdx.x <= j) {
        const int q = threadIdx.x;
        double2 s = make_double2(0, 0);
        const c64 *p = part + ((int64_t)t * kslots + q) * nchunk;
        for (int i = 0; i < nchunk; ++i) s = cadd(s, make_double2(p[i].re, p[i].im));
        const double sq = lz_scale_of(be, q);
        s = cscale(s, sq * sj);
        if (q == j && c == 0) galpha[t] = s.x;
        g[q] = cscale(s, sq);
    }
    __syncthreads();
    cplx<T> *Vt = V + t * sV;
    cplx<T> *w = Vt + (int64_t)(j + 1) * nmax;
    const int64_t n2 = n >> 1, base = (int64_t)c * (LZ_CHUNK / 2);
    double2 xa[4], xb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t e2 = base + threadIdx.x + 256 * r;
        xa[r] = xb[r] = make_double2(0, 0);
        if (e2 < n2) {
            LzPair<T>::ld(w, e2, xa[r], xb[r]);
            xa[r] = cscale(xa[r], sj);
            xb[r] = cscale(xb[r], sj);
        }
    }
    for (int q = 0; q <= j; ++q) {
        const double2 gq = g[q];
        const cplx<T> *vq = Vt + (int64_t)q * nmax;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t e2 = base + threadIdx.x + 256 * r;
            if (e2 < n2) {
                double2 va, vb;
                LzPair<T>::ld(vq, e2, va, vb);
                xa[r] = csub(xa[r], cmul(gq, va));
                xb[r] = csub(xb[r], cmul(gq, vb));
            }
        }
    }
    double nrm = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t e2 = base + threadIdx.x + 256 * r;
        if (e2 < n2) {
            LzPair<T>::st(w, e2, xa[r], xb[r]);
            // the norm of the stored (storage-precision) values
            double2 ba, bb;
            LzPair<T>::ld(w, e2, ba, bb);
            nrm += ba.x * ba.x + ba.y * ba.y + bb.x * bb.x + bb.y * bb.y;
        }
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) partn[(int64_t)t * nchunk + c] = nrm;
}

// exp(-i tau T) e_0 for the symmetric tridiagonal T (k <= 64), one lane per
// component (diagonal a, couplings bl to lane - 1 and br to lane + 1, zero
// on lanes >= k).  Neighbours are read by DPP wavefront shifts (lane 0 /
// lane 63 read 0).
//
// Chebyshev expansion (Jacobi-Anger): with the Gershgorin interval
// [c - r, c + r] of T, T' = (T - c) / r and x = tau r,
//   exp(-i tau T) e0 = e^{-i tau c} sum_k eps_k (-i)^k J_k(x) T_k(T') e0,
// eps_0 = 1, eps_k = 2 (k >= 1).  T_k(T') e0 comes from the three-term
// recurrence on real vectors (|T_k(T')| <= 1 on the spectrum), the Bessel
// values J_k(x) from Miller's backward recurrence normalised by
// J_0 + 2 sum J_2k = 1, started where the bound |J_k(x)| <= (|x|/2)^k / k!
// is below 1e-30; the series stops where the bound is below 1e-17.  When
// that start would pass lane 60 (|x| above about 25) the Taylor sub-steps
// below run instead.
__device__ double2 tridiag_exp_taylor(double a, double bl, double br, int k, double tau, int lane) {
    double rown = fabs(a) + fabs(bl) + fabs(br);
    rown = wave_max(rown);
    // sub-steps with |h| rown <= 1/2; NaN gives one, and the count is capped
    // so that a non-finite T (a trajectory the split guard flags) cannot
    // keep the wavefront looping
    const double sd = ceil(fabs(tau) * rown / 0.5);
    const int s = sd >= 1.0 ? (sd < 65536.0 ? (int)sd : 65536) : 1;
    const double h = tau / s;
    double2 v = make_double2(lane == 0 ? 1.0 : 0.0, 0.0);
    for (int st = 0; st < s; ++st) {
        double2 term = v, acc = v;
        for (int m = 1; m <= 40; ++m) {
            // term <- (-i h T) term / m
            const double tlx = dpp_move<0x138>(term.x), tly = dpp_move<0x138>(term.y);   // lane - 1
            const double trx = dpp_move<0x130>(term.x), try_ = dpp_move<0x130>(term.y);  // lane + 1
            double ux = a * term.x + bl * tlx + br * trx;
            double uy = a * term.y + bl * tly + br * try_;
            if (lane >= k) { ux = 0; uy = 0; }
            // (-i h / m) * (ux + i uy) = (h/m) (uy - i ux)
            const double f = h / m;
            term = make_double2(f * uy, -f * ux);
            acc = cadd(acc, term);
            // stop once no lane's term reaches 1e-20 (one compare and a ballot)
            if (__ballot(fabs(term.x) + fabs(term.y) >= 1e-20) == 0) break;
        }
        v = acc;
    }
    return lane < k ? v : make_double2(0, 0);
}

__device__ double2 tridiag_exp_lane(const double *alpha, const double *beta, int k, double tau, int lane,
                                    double a_last) {
    const bool in = lane < k;
    const double a = in ? (lane == k - 1 ? a_last : alpha[lane]) : 0.0;
    const double bl = (lane >= 1 && in) ? beta[lane] : 0.0;  // coupling to lane-1
    const double br = (lane + 1 < k) ? beta[lane + 1] : 0.0;  // coupling to lane+1
    const double rad = fabs(bl) + fabs(br);
    const double lo = wave_min(in ? a - rad : 1e300), hi = wave_max(in ? a + rad : -1e300);
    const double c = 0.5 * (lo + hi), r = 0.5 * (hi - lo);
    const double x = tau * r, hx = 0.5 * fabs(x);
    // number of terms (bound below 1e-17) and Miller start (bound below 1e-30)
    int K = 0, N = 0;
    if (x != 0.0) {
        double bnd = 1.0;
        for (N = 1; N <= 61; ++N) {
            bnd *= hx / N;
            if (K == 0 && bnd < 1e-17) K = N;
            if (bnd < 1e-30) break;
        }
        if (N > 60) return tridiag_exp_taylor(a, bl, br, k, tau, lane);
    }
    // lane q keeps J_q(x)
    double jq = lane == 0 ? 1.0 : 0.0;
    if (K > 0) {
        double jp = 0.0, jc = 1e-250, nsum = (N & 1) ? 0.0 : 2.0 * jc;
        jq = lane == N ? jc : 0.0;
        const double tx = 2.0 / x;
        for (int q = N; q >= 1; --q) {
            const double jm = (tx * q) * jc - jp;  // J_{q-1}
            jp = jc;
            jc = jm;
            if (lane == q - 1) jq = jm;
            if (((q - 1) & 1) == 0) nsum += q == 1 ? jm : 2.0 * jm;
        }
        jq /= nsum;
    }
    // T_q(T') e0 on real vectors, accumulated with the coefficients
    const double ap = in && K > 0 ? (a - c) / r : 0.0;
    const double blp = K > 0 ? bl / r : 0.0, brp = K > 0 ? br / r : 0.0;
    double w0 = lane == 0 ? 1.0 : 0.0;
    double2 acc = make_double2(lane_read(jq, 0) * w0, 0.0);
    if (K > 0) {
        double w1 = ap * w0 + blp * dpp_move<0x138>(w0) + brp * dpp_move<0x130>(w0);
        acc.y -= 2.0 * lane_read(jq, 1) * w1;
        for (int q = 2; q <= K; ++q) {
            const double tw = ap * w1 + blp * dpp_move<0x138>(w1) + brp * dpp_move<0x130>(w1);
            const double w2 = 2.0 * tw - w0;
            w0 = w1;
            w1 = w2;
            const double cq = 2.0 * lane_read(jq, q);
            switch (q & 3) {
                case 0: acc.x += cq * w2; break;
                case 1: acc.y -= cq * w2; break;
                case 2: acc.x -= cq * w2; break;
                default: acc.y += cq * w2; break;
            }
        }
    }
    double sn, cs;
    sincos(tau * c, &sn, &cs);
    const double2 v = make_double2(cs * acc.x + sn * acc.y, cs * acc.y - sn * acc.x);
    return in ? v : make_double2(0, 0);
}

// lz_finalize: one wavefront per trajectory.  β = ‖w‖ from the chunk
// partials (w, stored unnormalised, becomes V_{j+1} with s_{j+1} = 1/β),
// c = exp(-i dt T) e0 and the stopping rule β0·β·|c_j| <= tol (shared with
// the oracle).  A trajectory that stops keeps nit = j + 1 and leaves the
// active set; trajectories that go on are counted in nactive[j], first per
// workgroup in LDS, then with one global atomic per workgroup (a thousand
// same-address atomics from single waves cost about 10 us).  With a host
// word, each workgroup also adds (1 << 32) | its count to the 64-bit
// `fin_done`: one atomic on one word carries the ticket (high half) and the
// running total (low half), so the workgroup that draws the last ticket has
// every count in its return value, with no fence; it stores the total into
// the host's pinned word (the host needs the count alone, so a relaxed store
// suffices: device-scope and system-scope release fences write back the L2
// of the XCD, which cost most of this kernel's time).
#ifndef LZ_FIN_THREADS
#define LZ_FIN_THREADS 256
#endif
__global__ __launch_bounds__(LZ_FIN_THREADS) void lz_finalize_kernel(int j, int kmax, double dt, double tol,
                                                                     const double *partn, int nchunk,
                                                                     const double *galpha, double *alpha,
                                                                     double *beta, const double *beta0, c64 *coef,
                                                                     int *nit, int *active, int *nactive, int B,
                                                                     unsigned long long *fin_done, int *host_count) {
    __shared__ int going;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (LZ_FIN_THREADS / 64) + (threadIdx.x >> 6);
    if (threadIdx.x == 0) going = 0;
    __syncthreads();
    if (t < B && active[t]) {
        double s = 0;
        for (int i = lane; i < nchunk; i += 64) s += partn[(int64_t)t * nchunk + i];
        s = wave_sum(s);
        const double b = sqrt(s);
        double *al = alpha + (int64_t)t * kmax;
        double *be = beta + (int64_t)t * (kmax + 1);
        const double aj = galpha[t];
        if (lane == 0) al[j] = aj;
        const double2 cv = tridiag_exp_lane(al, be, j + 1, dt, lane, aj);
        const double cjx = lane_read(cv.x, j), cjy = lane_read(cv.y, j);
        const double err = beta0[t] * b * sqrt(cjx * cjx + cjy * cjy);
        const bool done = (err <= tol) || (b == 0.0) || (j + 1 == kmax);
        if (done) {
            if (lane <= j) {
                c64 o;
                o.re = cv.x;
                o.im = cv.y;
                coef[(int64_t)t * kmax + lane] = o;
            }
            if (lane == 0) {
                nit[t] = j + 1;
                active[t] = 0;
            }
        } else if (lane == 0) {
            be[j + 1] = b;
            atomicAdd(&going, 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (going) atomicAdd(nactive + j, going);
        if (host_count) {
            const unsigned long long old = atomicAdd(fin_done, (1ull << 32) | (unsigned long long)going);
            if ((unsigned)(old >> 32) == gridDim.x - 1) {
                const int total = (int)(old & 0xffffffffull) + going;
                atomicExch(fin_done, 0ull);
                __hip_atomic_store(host_count, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// x = β0 Σ_{q<nit} c_q v_q = β0 Σ c_q s_q V_q  (trajectories with β0 > 0); the basis is read
// in element pairs with q outside the pairs (per element the sum runs over
// q ascending); x is written in pairs when its slots allow (xpair).
template <typename T>
__global__ __launch_bounds__(256) void lz_combine_kernel(cplx<T> *x, int64_t sx, const cplx<T> *V, int64_t sV,
                                                         int64_t nmax, int64_t n, const double *beta0, const c64 *coef,
                                                         const int *nit, int kmax, const int *mask, int xpair,
                                                         const double *beta) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const double b0 = beta0[t];
    if (!(b0 > 0.0)) return;
    const int k = nit[t];
    const cplx<T> *Vt = V + t * sV;
    const c64 *ct = coef + (int64_t)t * kmax;
    const int64_t n2 = n >> 1, base = (int64_t)blockIdx.x * (LZ_CHUNK / 2);
    double2 xa[4], xb[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) xa[r] = xb[r] = make_double2(0, 0);
    const double *be = beta + (int64_t)t * (kmax + 1);
    for (int q = 0; q < k; ++q) {
        const double2 c = cscale(make_double2(ct[q].re, ct[q].im), lz_scale_of(be, q));
        const cplx<T> *vq = Vt + (int64_t)q * nmax;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t e2 = base + threadIdx.x + 256 * r;
            if (e2 < n2) {
                double2 va, vb;
                LzPair<T>::ld(vq, e2, va, vb);
                xa[r] = cadd(xa[r], cmul(c, va));
                xb[r] = cadd(xb[r], cmul(c, vb));
            }
        }
    }
    cplx<T> *xt = x + t * sx;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t e2 = base + threadIdx.x + 256 * r;
        if (e2 >= n2) continue;
        if (xpair) {
            LzPair<T>::st(xt, e2, cscale(xa[r], b0), cscale(xb[r], b0));
        } else {
            st2(xt[2 * e2], cscale(xa[r], b0));
            st2(xt[2 * e2 + 1], cscale(xb[r], b0));
        }
    }
}

template <typename T>
void lz_init(const cplx<T> *x, int64_t sx, cplx<T> *V, int64_t sV, int64_t n, double *beta0, int *active, int *nit,
             int *nactive, int nslots, int B, hipStream_t s) {
    hipLaunchKernelGGL(lz_init_kernel<T>, dim3(B), dim3(256), 0, s, x, sx, V, sV, n, beta0, active, nit, nactive,
                       nslots);
    YQ_HIP(hipGetLastError());
}
template <typename T>
void lz_dots(const cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j, c64 *part, int kslots, const int *active,
             int B, hipStream_t s) {
    if ((n | nmax | sV) & 1) throw std::invalid_argument("Lanczos vectors must have even length and offsets");
    const int nchunk = (int)((n + LZ_CHUNK - 1) / LZ_CHUNK);
    hipLaunchKernelGGL(lz_dots_kernel<T>, dim3(nchunk, B), dim3(256), 0, s, V, sV, nmax, n, j, part, nchunk, kslots,
                       active);
    YQ_HIP(hipGetLastError());
}
template <typename T>
void lz_update(cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j, const c64 *part, int kslots, double *partn,
               double *galpha, const int *active, const double *beta, int kmax, int B, hipStream_t s) {
    if ((n | nmax | sV) & 1) throw std::invalid_argument("Lanczos vectors must have even length and offsets");
    const int nchunk = (int)((n + LZ_CHUNK - 1) / LZ_CHUNK);
    hipLaunchKernelGGL(lz_update_kernel<T>, dim3(nchunk, B), dim3(256), 0, s, V, sV, nmax, n, j, part, nchunk, kslots,
                       partn, galpha, active, beta, kmax);
    YQ_HIP(hipGetLastError());
}
void lz_finalize(int j, int kmax, double dt, double tol, const double *partn, int64_t n, const double *galpha,
                 double *alpha, double *beta, const double *beta0, c64 *coef, int *nit, int *active, int *nactive, int B,
                 unsigned long long *fin_done, int *host_count, hipStream_t s) {
    if (host_count && !fin_done) throw std::invalid_argument("lz_finalize: host_count needs fin_done");
    const int nchunk = (int)((n + LZ_CHUNK - 1) / LZ_CHUNK);
    constexpr int per = LZ_FIN_THREADS / 64;
    hipLaunchKernelGGL(lz_finalize_kernel, dim3((B + per - 1) / per), dim3(LZ_FIN_THREADS), 0, s, j, kmax, dt, tol,
                       partn, nchunk, galpha, alpha, beta, beta0, coef, nit, active, nactive, B, fin_done,
                       host_count);
    YQ_HIP(hipGetLastError());
}
template <typename T>
void lz_combine(cplx<T> *x, int64_t sx, const cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, const double *beta0,
                const c64 *coef, const int *nit, const double *beta, int kmax, const int *mask, int B, hipStream_t s) {
    if ((n | nmax | sV) & 1) throw std::invalid_argument("Lanczos vectors must have even length and offsets");
    const int xpair = ((sx & 1) == 0 && ((uintptr_t)x & (2 * sizeof(cplx<T>) - 1)) == 0) ? 1 : 0;
    const int nchunk = (int)((n + LZ_CHUNK - 1) / LZ_CHUNK);
    hipLaunchKernelGGL(lz_combine_kernel<T>, dim3(nchunk, B), dim3(256), 0, s, x, sx, V, sV, nmax, n, beta0, coef, nit,
                       kmax, mask, xpair, beta);
    YQ_HIP(hipGetLastError());
}

#define YQ_LZ_INST(T)                                                                                              \
    template void lz_init<T>(const cplx<T> *, int64_t, cplx<T> *, int64_t, int64_t, double *, int *, int *, int *,  \
                             int, int, hipStream_t);                                                               \
    template void lz_dots<T>(const cplx<T> *, int64_t, int64_t, int64_t, int, c64 *, int, const int *, int,         \
                             hipStream_t);                                                                         \
    template void lz_update<T>(cplx<T> *, int64_t, int64_t, int64_t, int, const c64 *, int, double *, double *,     \
                               const int *, const double *, int, int, hipStream_t);                                \
    template void lz_combine<T>(cplx<T> *, int64_t, const cplx<T> *, int64_t, int64_t, int64_t, const double *,     \
                                const c64 *, const int *, const double *, int, const int *, int, hipStream_t);
YQ_LZ_INST(float)
YQ_LZ_INST(double)

// Write the kept eigenvectors as an isometry in storage precision.
//   left  (mode 0): Out[r][k] = U[r][perm[k]]        (r < n, k < kw), 0 for k >= keep
//   right (mode 1): Out[k][r] = conj(U[r][perm[k]])  (k < kw, r < n), 0 for k >= keep
template <typename T>
__global__ void write_iso_kernel(const c64 *U, int64_t sU, const int *perm, int64_t sE, const int *keep, int n, int kw,
                                 int mode, cplx<T> *Out, int64_t sO, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * kw) return;
    int r, k;
    if (mode == 0) { r = (int)(e / kw); k = (int)(e % kw); }
    else { k = (int)(e / n); r = (int)(e % n); }
    double2 v = make_double2(0, 0);
    if (k < keep[t]) {
        const c64 u = U[t * sU + (int64_t)r * n + perm[t * sE + k]];
        v = make_double2(u.re, mode == 0 ? u.im : -u.im);
    }
    st2(Out[t * sO + e], v);
}

template <typename T>
void write_iso(const c64 *U, int64_t sU, const int *perm, int64_t sE, const int *keep, int n, int kw, int mode,
               cplx<T> *Out, int64_t sO, const int *mask, int B, hipStream_t s) {
    const int64_t tot = (int64_t)n * kw;
    hipLaunchKernelGGL(write_iso_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, U, sU, perm, sE,
                       keep, n, kw, mode, Out, sO, mask);
    YQ_HIP(hipGetLastError());
}
template void write_iso<float>(const c64 *, int64_t, const int *, int64_t, const int *, int, int, int, c32 *, int64_t,
                               const int *, int, hipStream_t);
template void write_iso<double>(const c64 *, int64_t, const int *, int64_t, const int *, int, int, int, c64 *, int64_t,
                                const int *, int, hipStream_t);

// ------------------------------------------------- CGS2 orthonormalisation --
// Orthonormalises nvec vectors v_k[e] = A[k*vs + e*es] (e < len) in order by
// classical Gram-Schmidt applied twice, writing the coefficients
// Cf[k][j] = <q_j, a_k> (j <= k, row-major nvec x nvec, zero above the
// diagonal).  A vector whose remaining norm is <= 1e-12 ‖A‖_F becomes
// exactly zero (a dead direction, e.g. a zero-padded bond index).  Rows give
// LQ (vs = row stride, es = 1), columns QR (vs = 1, es = row stride).
//
// Blocked (panels of pw vectors): each panel is first projected twice
// against all earlier vectors with two batched matrix-core GEMMs per pass
// (H = Q^H P, then P -= Q H), then orthonormalised within itself by
// cgs2_panel_kernel with the panel held in LDS in FP64.

// ‖A‖_F per trajectory (FP64).
template <typename T>
__global__ __launch_bounds__(256) void cgs_fro_kernel(const cplx<T> *A, int64_t sA, int nvec, int len, int64_t vs,
                                                      int64_t es, double *fro, const int *mask) {
    __shared__ double sh[16];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const cplx<T> *M = A + t * sA;
    double f = 0;
    for (int64_t e = threadIdx.x; e < (int64_t)nvec * len; e += blockDim.x) {
        const int k = (int)(e / len), i = (int)(e % len);
        const double2 z = ld2(M[k * vs + i * es]);
        f += z.x * z.x + z.y * z.y;
    }
    f = block_sum(f, sh);
    if (threadIdx.x == 0) fro[t] = sqrt(f);
}

// Coefficients of one projection pass of panel [p0, p0+pw): H[j][q] (j < p0)
// from the GEMM scratch S (QR: S = H, p0 x pw; LQ: S = H^T, pw x p0), added
// to Cf[p0 + q][j] (stored on pass 0), and S replaced by the operand of the
// update GEMM: -H (QR) or -H^T (LQ).
template <typename T>
__global__ __launch_bounds__(256) void cgs_coef_kernel(cplx<T> *S, int64_t sS, cplx<T> *Cf, int64_t sC, int nvec,
                                                       int p0, int pw, int lq, int pass, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p0 * pw) return;
    cplx<T> *St = S + t * sS;
    cplx<T> *C = Cf + t * sC;
    int j, q;
    double2 h;
    if (lq) {
        q = e / p0;
        j = e % p0;
        h = ld2(St[e]);
    } else {
        j = e / pw;
        q = e % pw;
        h = ld2(St[e]);
    }
    st2(St[e], make_double2(-h.x, -h.y));
    cplx<T> &c = C[(int64_t)(p0 + q) * nvec + j];
    st2(c, pass ? cadd(ld2(c), h) : h);
}

// CGS2 of the vectors [p0, p0+pw) among themselves (they are already
// orthogonal to the earlier ones), panel in LDS as FP64 [pw][ld] with
// ld = len + 1 (consecutive vectors one 16-byte slot apart, so the four
// lanes of an element quad, which read four consecutive vectors, hit
// different banks); writes Cf[k][j] for p0 <= j <= k and zeroes Cf[k][j]
// for j > k.  Normalisation is deferred: vector j stays unscaled in LDS,
// q_j = s_j u_j with s_j its inverse norm (0 for a dead direction), and the
// panel is scaled as it is written back.  Per vector q and pass: wavefront
// w forms <u_j, v> for j = w (mod waves), giving the coefficient
// h_j = s_j <u_j, v> against q_j; barrier; element quad i (lane & 3 = part)
// sums s_j h_j u_j[i] over j = part (mod 4), adds the four sums by DPP and
// subtracts them from v[i]; barrier.  The second pass also forms the
// norm's per-wave partials, so a vector takes four barriers.  Blocks of up
// to 1024 threads (four per element, in multiples of 64).
template <typename T>
__global__ __launch_bounds__(1024) void cgs2_panel_kernel(cplx<T> *A, int64_t sA, int nvec, int len, int64_t vs,
                                                          int64_t es, int p0, int pw, cplx<T> *Cf, int64_t sC,
                                                          const double *fro, const int *mask) {
    extern __shared__ double2 pan[];  // [pw][ld]
    __shared__ double2 h[32], hs[32];  // h_j and s_j h_j of the current pass
    __shared__ double sc[32];          // s_j of the finished vectors
    __shared__ double npart[16];       // per-wave partials of the norm
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    cplx<T> *M = A + t * sA;
    cplx<T> *C = Cf + t * sC;
    const double dead = 1e-12 * fro[t];
    const int ld = len + 1;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int part = threadIdx.x & 3, nquad = blockDim.x >> 2;
    // load, coalesced along whichever index is contiguous
    if (vs == 1) {
        for (int e = threadIdx.x; e < pw * len; e += blockDim.x) {
            const int i = e / pw, q = e % pw;
            pan[q * ld + i] = ld2(M[(int64_t)(p0 + q) + (int64_t)i * es]);
        }
    } else {
        for (int e = threadIdx.x; e < pw * len; e += blockDim.x) {
            const int q = e / len, i = e % len;
            pan[q * ld + i] = ld2(M[(int64_t)(p0 + q) * vs + (int64_t)i * es]);
        }
    }
    for (int e = threadIdx.x; e < pw * nvec; e += blockDim.x) {
        const int q = e / nvec, j = e % nvec;
        if (j >= p0 + q) C[(int64_t)(p0 + q) * nvec + j] = cmake<T>((T)0, (T)0);  // diagonal set below
    }
    __syncthreads();
    double s_prev = 0.0;  // s_{q-1}, held by every thread (sc[q-1] is not yet visible)
    for (int q = 0; q < pw; ++q) {
        double2 *v = pan + q * ld;
        double nrm = 0.0;
        for (int pass = (q == 0 ? 1 : 0); pass < 2; ++pass) {
            for (int j = wid; j < q; j += nw) {
                const double2 *u = pan + j * ld;
                double2 acc = make_double2(0, 0);
                for (int i = lane; i < len; i += 64) acc = cadd(acc, cmulc(u[i], v[i]));
                acc.x = wave_sum(acc.x);
                acc.y = wave_sum(acc.y);
                const double sj = j == q - 1 ? s_prev : sc[j];
                const double2 hj = cscale(acc, sj);
                if (lane == 0) {
                    h[j] = hj;
                    hs[j] = cscale(hj, sj);
                }
            }
            __syncthreads();
            for (int j = threadIdx.x; j < q; j += blockDim.x) {
                cplx<T> &c = C[(int64_t)(p0 + q) * nvec + p0 + j];
                st2(c, pass ? cadd(ld2(c), h[j]) : h[j]);
            }
            for (int i0 = 0; i0 < len; i0 += nquad) {
                const int i = i0 + (threadIdx.x >> 2);
                double2 acc = make_double2(0, 0);
                if (i < len)
                    for (int j = part; j < q; j += 4) acc = cadd(acc, cmul(hs[j], pan[j * ld + i]));
                acc.x += dpp_move<0xb1>(acc.x);  // lanes xor 1
                acc.y += dpp_move<0xb1>(acc.y);
                acc.x += dpp_move<0x4e>(acc.x);  // lanes xor 2
                acc.y += dpp_move<0x4e>(acc.y);
                if (part == 0 && i < len) {
                    const double2 x = csub(v[i], acc);
                    v[i] = x;
                    if (pass == 1) nrm += x.x * x.x + x.y * x.y;
                }
            }
            if (pass == 1) {
                nrm = wave_sum(nrm);
                if (lane == 0) npart[wid] = nrm;
            }
            __syncthreads();
        }
        double nr = 0.0;
        for (int w = 0; w < nw; ++w) nr += npart[w];
        nr = sqrt(nr);
        const bool live = nr > dead && nr > 0.0;
        s_prev = live ? 1.0 / nr : 0.0;
        if (threadIdx.x == 0) {
            sc[q] = s_prev;
            C[(int64_t)(p0 + q) * nvec + p0 + q] = cmake<T>((T)(live ? nr : 0.0), (T)0);
        }
    }
    __syncthreads();
    if (vs == 1) {
        for (int e = threadIdx.x; e < pw * len; e += blockDim.x) {
            const int i = e / pw, q = e % pw;
            st2(M[(int64_t)(p0 + q) + (int64_t)i * es], cscale(pan[q * ld + i], sc[q]));
        }
    } else {
        for (int e = threadIdx.x; e < pw * len; e += blockDim.x) {
            const int q = e / len, i = e % len;
            st2(M[(int64_t)(p0 + q) * vs + (int64_t)i * es], cscale(pan[q * ld + i], sc[q]));
        }
    }
}

// Panel width: the panel (pw x len FP64 complex) fits in 128 KB of LDS (its
// padded rows add pw slots of 16 bytes).
inline int cgs_panel_width(int len) {
    int pw = 32;
    while (pw > 1 && (int64_t)pw * len * 16 > 128 * 1024) pw /= 2;
    return pw;
}

int64_t cgs2_scratch_elems(int nvec, int len) { return (int64_t)nvec * cgs_panel_width(len); }

template <typename T>
void cgs2(cplx<T> *A, int64_t sA, int nvec, int len, int64_t vs, int64_t es, cplx<T> *Cf, int64_t sC, cplx<T> *S,
          int64_t sS, double *fro, const int *mask, int B, hipStream_t s) {
    if (len > 8192) throw HipError("cgs2: vectors longer than 8192");
    if (vs != 1 && es != 1) throw HipError("cgs2: one of the strides must be 1");
    const bool lq = vs != 1;
    const int PB = cgs_panel_width(len);
    const int panel_threads = std::min(1024, std::max(64, (4 * len + 63) / 64 * 64));
    hipLaunchKernelGGL(cgs_fro_kernel<T>, dim3(B), dim3(256), 0, s, A, sA, nvec, len, vs, es, fro, mask);
    YQ_HIP(hipGetLastError());
    for (int p0 = 0; p0 < nvec; p0 += PB) {
        const int pw = std::min(PB, nvec - p0);
        if (p0 > 0) {
            // H = Q^H P over the earlier vectors Q = [0, p0), then P -= Q H, twice
            GemmDesc g1{}, g2{};
            g1.inner = g2.inner = 1;
            g1.outer = g2.outer = B;
            g1.active = g2.active = mask;
            g1.sA_outer = g1.sB_outer = sA;
            g1.sC_outer = sS;
            g2.sA_outer = lq ? sS : sA;
            g2.sB_outer = lq ? sA : sS;
            g2.sC_outer = sA;
            g2.accumulate = 1;
            const cplx<T> *g1a, *g1b, *g2a, *g2b;
            cplx<T> *P = lq ? A + (int64_t)p0 * vs : A + p0;
            if (!lq) {
                // A is len x nvec row-major (ld es): H (p0 x pw) = Q^H P
                g1.M = p0; g1.N = pw; g1.K = len; g1.opA = OP_C; g1.opB = OP_N;
                g1.lda = es; g1.ldb = es; g1.ldc = pw;
                g1a = A; g1b = P;
                // P (len x pw) += Q (len x p0) (-H) (p0 x pw)
                g2.M = len; g2.N = pw; g2.K = p0; g2.opA = OP_N; g2.opB = OP_N;
                g2.lda = es; g2.ldb = pw; g2.ldc = es;
                g2a = A; g2b = S;
            } else {
                // A is nvec x len row-major (ld vs): S (pw x p0) = P Q^H = H^T
                g1.M = pw; g1.N = p0; g1.K = len; g1.opA = OP_N; g1.opB = OP_C;
                g1.lda = vs; g1.ldb = vs; g1.ldc = p0;
                g1a = P; g1b = A;
                // P (pw x len) += (-H^T) (pw x p0) Q (p0 x len)
                g2.M = pw; g2.N = len; g2.K = p0; g2.opA = OP_N; g2.opB = OP_N;
                g2.lda = p0; g2.ldb = vs; g2.ldc = vs;
                g2a = S; g2b = A;
            }
            // C tiles matched to the panel width (32 columns or rows: the
            // 64 x 32 / 32 x 64 tiles, bit-identical to the 64 x 64 one;
            // they exist for op N and op C right operands)
            auto tile_for = [](const GemmDesc &g) {
                if (g.opB == OP_T) return 0;
                if (g.M <= 32 && g.N <= 32) return 3;
                if (g.N <= 32) return 2;
                if (g.M <= 32) return 1;
                return 0;
            };
            const int t1 = tile_for(g1), t2 = tile_for(g2);
            for (int pass = 0; pass < 2; ++pass) {
                if (!gemm_mfma<T>(g1, g1a, g1b, S, s, t1)) gemm<T, T, T, T>(g1, g1a, g1b, S, s);
                hipLaunchKernelGGL(cgs_coef_kernel<T>, dim3((p0 * pw + 255) / 256, B), dim3(256), 0, s, S, sS, Cf, sC,
                                   nvec, p0, pw, lq ? 1 : 0, pass, mask);
                YQ_HIP(hipGetLastError());
                if (!gemm_mfma<T>(g2, g2a, g2b, P, s, t2)) gemm<T, T, T, T>(g2, g2a, g2b, P, s);
            }
        }
        hipLaunchKernelGGL(cgs2_panel_kernel<T>, dim3(B), dim3(panel_threads), (size_t)pw * (len + 1) * sizeof(double2),
                           s, A, sA, nvec, len, vs, es, p0, pw, Cf, sC, (const double *)fro, mask);
        YQ_HIP(hipGetLastError());
    }
}
template void cgs2<float>(c32 *, int64_t, int, int, int64_t, int64_t, c32 *, int64_t, c32 *, int64_t, double *,
                          const int *, int, hipStream_t);
template void cgs2<double>(c64 *, int64_t, int, int, int64_t, int64_t, c64 *, int64_t, c64 *, int64_t, double *,
                           const int *, int, hipStream_t);

// ----------------------------------------------------- pair partial traces --
// One-site density matrices of sites a and a+1 from the 2-site ρ2 of the pair
// (row index 2 s_a + s_{a+1}): ρ_a[x][y] = Σ_z ρ2[2x+z][2y+z],
// ρ_{a+1}[x][y] = Σ_z ρ2[2z+x][2z+y]; every other site of ρ1 is zeroed.
__global__ void ptrace_pair_kernel(const c64 *rho2, int64_t s2, int a, c64 *rho1, int64_t s1, int L) {
    const int t = blockIdx.x;
    const c64 *r2 = rho2 + t * s2;
    c64 *r1 = rho1 + t * s1;
    for (int e = threadIdx.x; e < 4 * L; e += blockDim.x) {
        const int site = e / 4, x = (e % 4) / 2, y = e % 2;
        double2 acc = make_double2(0, 0);
        if (site == a || site == a + 1)
            for (int z = 0; z < 2; ++z) {
                const c64 v = site == a ? r2[(2 * x + z) * 4 + (2 * y + z)] : r2[(2 * z + x) * 4 + (2 * z + y)];
                acc.x += v.re;
                acc.y += v.im;
            }
        r1[e] = cmake<double>(acc.x, acc.y);
    }
}

void ptrace_pair(const c64 *rho2, int64_t s2, int a, c64 *rho1, int64_t s1, int L, int B, hipStream_t s) {
    hipLaunchKernelGGL(ptrace_pair_kernel, dim3(B), dim3(64), 0, s, rho2, s2, a, rho1, s1, L);
    YQ_HIP(hipGetLastError());
}

// ------------------------------------------------------- local operators --
// A[a][s][b] <- Σ_s' M[s][s'] A[a][s'][b]; M given per trajectory through
// op_idx (index into ops, -1 = skip) or shared (op_idx == nullptr -> ops[0]).
template <typename T>
__global__ void local_op_kernel(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *ops, const int *op_idx,
                                const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    int oi = 0;
    if (op_idx) {
        oi = op_idx[t];
        if (oi < 0) return;
    }
    const c64 *M = ops + (int64_t)oi * 4;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)dl * dr) return;
    const int a = (int)(e / dr), b = (int)(e % dr);
    cplx<T> *p0 = A + t * sA + ((int64_t)a * 2 + 0) * dr + b;
    cplx<T> *p1 = A + t * sA + ((int64_t)a * 2 + 1) * dr + b;
    const double2 x0 = ld2(*p0), x1 = ld2(*p1);
    const double2 m00 = make_double2(M[0].re, M[0].im), m01 = make_double2(M[1].re, M[1].im);
    const double2 m10 = make_double2(M[2].re, M[2].im), m11 = make_double2(M[3].re, M[3].im);
    st2(*p0, cadd(cmul(m00, x0), cmul(m01, x1)));
    st2(*p1, cadd(cmul(m10, x0), cmul(m11, x1)));
}

template <typename T>
void local_op(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *ops, const int *op_idx, const int *mask, int B,
              hipStream_t s) {
    const int64_t tot = (int64_t)dl * dr;
    hipLaunchKernelGGL(local_op_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, A, sA, dl, dr, ops,
                       op_idx, mask);
    YQ_HIP(hipGetLastError());
}
template void local_op<float>(c32 *, int64_t, int, int, const c64 *, const int *, const int *, int, hipStream_t);
template void local_op<double>(c64 *, int64_t, int, int, const c64 *, const int *, const int *, int, hipStream_t);

// Θ[a][p][b] <- Σ_p' M[p][p'] Θ[a][p'][b] for a 4x4 M on the combined
// physical index p = 2 s1 + s2.
template <typename T>
__global__ void local_op4_kernel(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *M, const int *op_idx) {
    const int t = blockIdx.y;
    if (op_idx) {  // per-trajectory matrix from a table of 16-entry matrices, < 0: untouched
        if (op_idx[t] < 0) return;
        M += (int64_t)op_idx[t] * 16;
    }
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)dl * dr) return;
    const int a = (int)(e / dr), b = (int)(e % dr);
    double2 x[4];
    for (int p = 0; p < 4; ++p) x[p] = ld2(A[t * sA + ((int64_t)a * 4 + p) * dr + b]);
    for (int p = 0; p < 4; ++p) {
        double2 acc = make_double2(0, 0);
        for (int q = 0; q < 4; ++q) acc = cadd(acc, cmul(make_double2(M[p * 4 + q].re, M[p * 4 + q].im), x[q]));
        st2(A[t * sA + ((int64_t)a * 4 + p) * dr + b], acc);
    }
}
template <typename T>
void local_op4(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *M, int B, hipStream_t s) {
    const int64_t tot = (int64_t)dl * dr;
    hipLaunchKernelGGL(local_op4_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, A, sA, dl, dr, M,
                       nullptr);
    YQ_HIP(hipGetLastError());
}
template void local_op4<float>(c32 *, int64_t, int, int, const c64 *, int, hipStream_t);
template void local_op4<double>(c64 *, int64_t, int, int, const c64 *, int, hipStream_t);
template <typename T>
void local_op4_sel(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *Mtab, const int *op_idx, int B, hipStream_t s) {
    const int64_t tot = (int64_t)dl * dr;
    hipLaunchKernelGGL(local_op4_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, A, sA, dl, dr,
                       Mtab, op_idx);
    YQ_HIP(hipGetLastError());
}
template void local_op4_sel<float>(c32 *, int64_t, int, int, const c64 *, const int *, int, hipStream_t);
template void local_op4_sel<double>(c64 *, int64_t, int, int, const c64 *, const int *, int, hipStream_t);

// ‖A_t‖² in FP64 (one workgroup per trajectory).
template <typename T>
__global__ __launch_bounds__(256) void norm2_kernel(const cplx<T> *A, int64_t sA, int64_t n, double *out, const int *mask) {
    __shared__ double sh[16];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    double s = 0;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        const double2 z = ld2(A[t * sA + e]);
        s += z.x * z.x + z.y * z.y;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[t] = s;
}
template <typename T>
void norm2(const cplx<T> *A, int64_t sA, int64_t n, double *out, const int *mask, int B, hipStream_t s) {
    hipLaunchKernelGGL(norm2_kernel<T>, dim3(B), dim3(256), 0, s, A, sA, n, out, mask);
    YQ_HIP(hipGetLastError());
}
template void norm2<float>(const c32 *, int64_t, int64_t, double *, const int *, int, hipStream_t);
template void norm2<double>(const c64 *, int64_t, int64_t, double *, const int *, int, hipStream_t);

template <typename T>
__global__ void scale_kernel(cplx<T> *A, int64_t sA, int64_t n, const double *f, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    st2(A[t * sA + e], cscale(ld2(A[t * sA + e]), f[t]));
}
template <typename T>
void scale(cplx<T> *A, int64_t sA, int64_t n, const double *f, const int *mask, int B, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel<T>, dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, s, A, sA, n, f, mask);
    YQ_HIP(hipGetLastError());
}
template void scale<float>(c32 *, int64_t, int64_t, const double *, const int *, int, hipStream_t);
template void scale<double>(c64 *, int64_t, int64_t, const double *, const int *, int, hipStream_t);

// Strided 2D copy with precision conversion: Dst[t][r][c] = Src[t][r][c].
template <typename TS, typename TD>
__global__ void copy2d_kernel(const cplx<TS> *S, int64_t sS, int64_t lds, cplx<TD> *D, int64_t sD, int64_t ldd, int rows,
                              int cols, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int r = (int)(e / cols), c = (int)(e % cols);
    st2(D[t * sD + r * ldd + c], ld2(S[t * sS + r * lds + c]));
}
template <typename TS, typename TD>
void copy2d(const cplx<TS> *S, int64_t sS, int64_t lds, cplx<TD> *D, int64_t sD, int64_t ldd, int rows, int cols,
            const int *mask, int B, hipStream_t s) {
    const int64_t tot = (int64_t)rows * cols;
    if (tot <= 0) return;
    hipLaunchKernelGGL((copy2d_kernel<TS, TD>), dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, S, sS, lds, D, sD,
                       ldd, rows, cols, mask);
    YQ_HIP(hipGetLastError());
}
template void copy2d<float, float>(const c32 *, int64_t, int64_t, c32 *, int64_t, int64_t, int, int, const int *, int,
                                   hipStream_t);
template void copy2d<double, double>(const c64 *, int64_t, int64_t, c64 *, int64_t, int64_t, int, int, const int *, int,
                                     hipStream_t);
template void copy2d<float, double>(const c32 *, int64_t, int64_t, c64 *, int64_t, int64_t, int, int, const int *, int,
                                    hipStream_t);
template void copy2d<double, float>(const c64 *, int64_t, int64_t, c32 *, int64_t, int64_t, int, int, const int *, int,
                                    hipStream_t);

// Bond matrix of a gauge step into a Krylov slot: D[r][c] = S[r][c] (or
// S[c][r] with `tr`) for r, c < n, and zero in columns n .. ldd-1 (Krylov
// vectors are even-length element pairs).
template <typename T>
__global__ void copy_bond_kernel(const cplx<T> *S, int64_t sS, int lds, int tr, cplx<T> *D, int64_t sD, int ldd, int n,
                                 const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * ldd) return;
    const int r = (int)(e / ldd), c = (int)(e % ldd);
    double2 v = make_double2(0.0, 0.0);
    if (c < n) v = ld2(S[t * sS + (tr ? (int64_t)c * lds + r : (int64_t)r * lds + c)]);
    st2(D[t * sD + (int64_t)r * ldd + c], v);
}
template <typename T>
void copy_bond(const cplx<T> *S, int64_t sS, int lds, bool tr, cplx<T> *D, int64_t sD, int ldd, int n, const int *mask,
               int B, hipStream_t s) {
    const int64_t tot = (int64_t)n * ldd;
    if (tot <= 0) return;
    hipLaunchKernelGGL(copy_bond_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, S, sS, lds, tr ? 1 : 0,
                       D, sD, ldd, n, mask);
    YQ_HIP(hipGetLastError());
}
template void copy_bond<float>(const c32 *, int64_t, int, bool, c32 *, int64_t, int, int, const int *, int, hipStream_t);
template void copy_bond<double>(const c64 *, int64_t, int, bool, c64 *, int64_t, int, int, const int *, int, hipStream_t);

// Zero column `col` of the rows x ld matrix of every (unmasked) slot.
template <typename T>
__global__ void zero_col_kernel(cplx<T> *Y, int64_t sY, int rows, int ld, int col, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) st2(Y[t * sY + (int64_t)r * ld + col], make_double2(0.0, 0.0));
}
template <typename T>
void zero_col(cplx<T> *Y, int64_t sY, int rows, int ld, int col, const int *mask, int B, hipStream_t s) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(zero_col_kernel<T>, dim3((unsigned)((rows + 255) / 256), B), dim3(256), 0, s, Y, sY, rows, ld, col,
                       mask);
    YQ_HIP(hipGetLastError());
}
template void zero_col<float>(c32 *, int64_t, int, int, int, const int *, int, hipStream_t);
template void zero_col<double>(c64 *, int64_t, int, int, int, const int *, int, hipStream_t);

// Fill with zero / identity-environment patterns.
template <typename T>
__global__ void fill_kernel(cplx<T> *A, int64_t sA, int64_t n) {
    const int t = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) A[t * sA + e].re = A[t * sA + e].im = (T)0;
}
template <typename T>
void fill_zero(cplx<T> *A, int64_t sA, int64_t n, int B, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, s, A, sA, n);
    YQ_HIP(hipGetLastError());
}
template void fill_zero<float>(c32 *, int64_t, int64_t, int, hipStream_t);
template void fill_zero<double>(c64 *, int64_t, int64_t, int, hipStream_t);

// E[i][m0][i] = 1 for i < dim (left env [bra][mpo][ket] or right env
// [mpo][ket][bra] with the identity on mpo index 0); the rest zero.
template <typename T>
__global__ void identity_env_kernel(cplx<T> *E, int64_t sE, int dim, int D, int right) {
    const int t = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)dim * D * dim;
    if (e >= tot) return;
    int i, m, k;
    if (!right) { i = (int)(e / ((int64_t)D * dim)); m = (int)((e / dim) % D); k = (int)(e % dim); }
    else { m = (int)(e / ((int64_t)dim * dim)); i = (int)((e / dim) % dim); k = (int)(e % dim); }
    E[t * sE + e].re = (i == k && m == 0) ? (T)1 : (T)0;
    E[t * sE + e].im = (T)0;
}
template <typename T>
void identity_env(cplx<T> *E, int64_t sE, int dim, int D, int right, int B, hipStream_t s) {
    const int64_t tot = (int64_t)dim * D * dim;
    hipLaunchKernelGGL(identity_env_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, E, sE, dim, D,
                       right);
    YQ_HIP(hipGetLastError());
}
template void identity_env<float>(c32 *, int64_t, int, int, int, int, hipStream_t);
template void identity_env<double>(c64 *, int64_t, int, int, int, int, hipStream_t);

// ρ[t][s][s'] = Σ_{a',b} X[a'][s][b] conj(A[a'][s'][b]) (site reduced matrix,
// d = 2 or 4), one workgroup per trajectory.
template <typename T>
__global__ __launch_bounds__(256) void rho_kernel(const c64 *X, int64_t sX, const cplx<T> *A, int64_t sA, int dl, int d,
                                                  int dr, c64 *rho, int64_t sR, const int *mask) {
    __shared__ double sh[16];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    for (int s = 0; s < d; ++s)
        for (int s2 = 0; s2 < d; ++s2) {
            double2 acc = make_double2(0, 0);
            for (int64_t e = threadIdx.x; e < (int64_t)dl * dr; e += blockDim.x) {
                const int a = (int)(e / dr), b = (int)(e % dr);
                const c64 x = X[t * sX + ((int64_t)a * d + s) * dr + b];
                const double2 y = ld2(A[t * sA + ((int64_t)a * d + s2) * dr + b]);
                acc = cadd(acc, cmulc(y, make_double2(x.re, x.im)));
            }
            acc = block_sum2(acc, sh);
            if (threadIdx.x == 0) {
                c64 o;
                o.re = acc.x;
                o.im = acc.y;
                rho[t * sR + s * d + s2] = o;
            }
        }
}
template <typename T>
void rho(const c64 *X, int64_t sX, const cplx<T> *A, int64_t sA, int dl, int d, int dr, c64 *out, int64_t sR,
         const int *mask, int B, hipStream_t s) {
    hipLaunchKernelGGL(rho_kernel<T>, dim3(B), dim3(256), 0, s, X, sX, A, sA, dl, d, dr, out, sR, mask);
    YQ_HIP(hipGetLastError());
}
template void rho<float>(const c64 *, int64_t, const c32 *, int64_t, int, int, int, c64 *, int64_t, const int *, int,
                         hipStream_t);
template void rho<double>(const c64 *, int64_t, const c64 *, int64_t, int, int, int, c64 *, int64_t, const int *, int,
                          hipStream_t);

// -------------------------------------------------------------- RNG + jumps
// (Philox: rng.h)

// Jump decision (src/StochasticProcess.jl:219-229): dp = 1 - ‖ψ‖², r1 draw.
// state[t]: 0 = no jump (scale factor written), 1 = jump candidate.
__global__ void jump_decide_kernel(const double *norm2, uint64_t seed, uint64_t traj0, uint64_t *rng_k, int has_procs,
                                   int *state, double *fac, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    const double n2 = norm2[t];
    const double dp = 1.0 - n2;
    bool jump = false;
    if (has_procs) {
        const double r1 = philox_uniform_dev(seed, traj0 + t, rng_k[t]);
        rng_k[t] += 1;
        jump = !(r1 >= dp);
    }
    state[t] = jump ? 1 : 0;
    const double nrm = sqrt(n2);
    fac[t] = nrm > 1e-13 ? 1.0 / nrm : 1.0;
}

// Jump choice (src/StochasticProcess.jl:89-212, 233-256) from the site
// reduced matrices ρ_s of the unnormalised state.  cand_* list the
// candidates in the reference's order; weight = dt γ Tr(L†L ρ_site).
__global__ void jump_choose_kernel(const c64 *rho, int64_t sR, int L, int ncand, const int *cand_site,
                                   const c64 *cand_LdL, const double *cand_gamma, double dt, uint64_t seed, uint64_t traj0,
                                   uint64_t *rng_k, int *state, int *chosen, double *wbuf, int B, const c64 *rho2,
                                   int64_t sR2, const c64 *cand_LdL4, const int *cand_kind, const double *lrnu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B || state[t] != 1) return;
    double tot = 0;
    double *w = wbuf + (int64_t)t * ncand;
    for (int c = 0; c < ncand; ++c) {
        if (cand_kind && cand_kind[c] == 2) {  // long-range MPO jump: dt γ ν
            w[c] = dt * cand_gamma[c] * lrnu[(int64_t)t * ncand + c];
            tot += w[c];
            continue;
        }
        // Tr(M ρ) = Σ_{s,s'} M[s'][s] ρ[s][s'], M = L†L: 2x2 on the site's ρ, or
        // 4x4 on the pair's ρ2 (merged-pair weight, src/StochasticProcess.jl:148-160)
        const bool pair = cand_kind && cand_kind[c] == 1;
        const int d = pair ? 4 : 2;
        const c64 *r = pair ? rho2 + t * sR2 + (int64_t)cand_site[c] * 16 : rho + t * sR + (int64_t)cand_site[c] * 4;
        const c64 *M = pair ? cand_LdL4 + (int64_t)c * 16 : cand_LdL + (int64_t)c * 4;
        double v = 0;
        for (int s = 0; s < d; ++s)
            for (int s2 = 0; s2 < d; ++s2) {
                const c64 m = M[s2 * d + s], p = r[s * d + s2];
                v += m.re * p.re - m.im * p.im;
            }
        w[c] = dt * cand_gamma[c] * v;
        tot += w[c];
    }
    if (tot == 0.0) {  // empty distribution: normalise only
        state[t] = 2;
        chosen[t] = -1;
        return;
    }
    const double r2 = philox_uniform_dev(seed, traj0 + t, rng_k[t]);
    rng_k[t] += 1;
    double acc = 0;
    int ch = 0;
    for (int c = 0; c < ncand; ++c) {
        acc += w[c] / tot;
        if (r2 <= acc) { ch = c; break; }
    }
    chosen[t] = ch;
}

void jump_decide(const double *norm2v, uint64_t seed, uint64_t traj0, uint64_t *rng_k, int has_procs, int *state,
                 double *fac, int B, hipStream_t s) {
    hipLaunchKernelGGL(jump_decide_kernel, dim3((B + 255) / 256), dim3(256), 0, s, norm2v, seed, traj0, rng_k, has_procs,
                       state, fac, B);
    YQ_HIP(hipGetLastError());
}
void jump_choose(const c64 *rhov, int64_t sR, int L, int ncand, const int *cand_site, const c64 *cand_LdL,
                 const double *cand_gamma, double dt, uint64_t seed, uint64_t traj0, uint64_t *rng_k, int *state,
                 int *chosen, double *wbuf, int B, hipStream_t s, const c64 *rho2, int64_t sR2, const c64 *cand_LdL4,
                 const int *cand_kind, const double *lrnu) {
    hipLaunchKernelGGL(jump_choose_kernel, dim3((B + 255) / 256), dim3(256), 0, s, rhov, sR, L, ncand, cand_site, cand_LdL,
                       cand_gamma, dt, seed, traj0, rng_k, state, chosen, wbuf, B, rho2, sR2, cand_LdL4, cand_kind, lrnu);
    YQ_HIP(hipGetLastError());
}

__global__ void mask_chosen_kernel(const int *state, const int *chosen, int c, int *mask, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < B) mask[t] = (state[t] == 1 && chosen[t] == c) ? 1 : 0;
}
void mask_chosen(const int *state, const int *chosen, int c, int *mask, int B, hipStream_t s) {
    hipLaunchKernelGGL(mask_chosen_kernel, dim3((B + 255) / 256), dim3(256), 0, s, state, chosen, c, mask, B);
    YQ_HIP(hipGetLastError());
}

__global__ void lr_nu_kernel(const c64 *val, const c64 *rho, int64_t sR, int nsite, double ar, double ai, double br,
                             double bi, int c, int ncand, const int *state, double *lrnu, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B || state[t] != 1) return;
    const c64 *r = rho + t * sR + (int64_t)nsite * 4;
    const double n = r[0].re + r[3].re;
    const c64 cs = val[t];
    const double cn = val[B + t].re;
    const double a2 = ar * ar + ai * ai, b2 = br * br + bi * bi;
    // conj(α) β
    const double xr = ar * br + ai * bi, xi = ar * bi - ai * br;
    const double base = a2 * n + b2 * cn;
    const double on2 = base + 2.0 * (xr * cs.re - xi * cs.im);
    lrnu[(int64_t)t * ncand + c] = on2 > 1e-12 * base ? 1.0 : 0.0;
}
void lr_nu(const c64 *val, const c64 *rho, int64_t sR, int nsite, double ar, double ai, double br, double bi, int c,
           int ncand, const int *state, double *lrnu, int B, hipStream_t s) {
    hipLaunchKernelGGL(lr_nu_kernel, dim3((B + 255) / 256), dim3(256), 0, s, val, rho, sR, nsite, ar, ai, br, bi, c,
                       ncand, state, lrnu, B);
    YQ_HIP(hipGetLastError());
}

template <typename T>
__global__ void mpo_block_kernel(cplx<T> *A, int64_t sA, int ncol, int mode, int nl, int R, const c64 *op, double ar,
                                 double ai, double br, double bi, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nl * R) return;
    const int l = (int)(e / R), r = (int)(e % R);
    cplx<T> *base = A + t * sA;
    const int64_t ls = ncol;  // columns of the site tensor (one physical slice of a row)
    cplx<T> *p0 = base + ((int64_t)l * 2 + 0) * ls + r;
    cplx<T> *p1 = base + ((int64_t)l * 2 + 1) * ls + r;
    const double2 x0 = ld2(*p0), x1 = ld2(*p1);
    if (mode == 1) {
        st2(base[((int64_t)(nl + l) * 2 + 0) * ls + R + r], x0);
        st2(base[((int64_t)(nl + l) * 2 + 1) * ls + R + r], x1);
        return;
    }
    const double2 m00 = make_double2(op[0].re, op[0].im), m01 = make_double2(op[1].re, op[1].im);
    const double2 m10 = make_double2(op[2].re, op[2].im), m11 = make_double2(op[3].re, op[3].im);
    double2 y0 = cadd(cmul(m00, x0), cmul(m01, x1)), y1 = cadd(cmul(m10, x0), cmul(m11, x1));
    if (mode == 0) {
        st2(p0[R], y0);
        st2(p1[R], y1);
        return;
    }
    const double2 be = make_double2(br, bi), al = make_double2(ar, ai);
    st2(base[((int64_t)(nl + l) * 2 + 0) * ls + r], cmul(be, y0));
    st2(base[((int64_t)(nl + l) * 2 + 1) * ls + r], cmul(be, y1));
    st2(*p0, cmul(al, x0));
    st2(*p1, cmul(al, x1));
}
template <typename T>
void mpo_block(cplx<T> *A, int64_t sA, int ncol, int mode, int nl, int R, const c64 *op, double ar, double ai,
               double br, double bi, const int *mask, int B, hipStream_t s) {
    const int64_t tot = (int64_t)nl * R;
    if (tot <= 0) return;
    hipLaunchKernelGGL(mpo_block_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, A, sA, ncol, mode, nl,
                       R, op, ar, ai, br, bi, mask);
    YQ_HIP(hipGetLastError());
}
template void mpo_block<float>(c32 *, int64_t, int, int, int, int, const c64 *, double, double, double, double,
                               const int *, int, hipStream_t);
template void mpo_block<double>(c64 *, int64_t, int, int, int, int, const c64 *, double, double, double, double,
                                const int *, int, hipStream_t);

// Per-site op index for the chosen jumps: opidx[t] = op of candidate
// chosen[t] if it acts on `site` (first or second factor), else -1.
__global__ void jump_opidx_kernel(const int *state, const int *chosen, const int *cand_site, const int *cand_op,
                                  const int *cand_site2, const int *cand_op2, int site, int *opidx, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    int o = -1;
    if (state[t] == 1 && chosen[t] >= 0) {
        const int c = chosen[t];
        if (cand_site[c] == site) o = cand_op[c];
        else if (cand_site2[c] == site) o = cand_op2[c];
    }
    opidx[t] = o;
}
// Trajectories whose chosen jump is a merged-pair candidate (cand_pair[c] >= 0:
// its 4x4 operator's index) on pair (p, p+1) (mode 0; opidx = that index,
// else -1) or on a pair beyond p (mode 1; opidx untouched).
__global__ void jump_pair_mask_kernel(const int *state, const int *chosen, const int *cand_site, const int *cand_pair,
                                      int p, int mode, int *mask, int *opidx, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    int m = 0, o = -1;
    if (state[t] == 1 && chosen[t] >= 0 && cand_pair[chosen[t]] >= 0) {
        const int site = cand_site[chosen[t]];
        if (mode == 0 && site == p) {
            m = 1;
            o = cand_pair[chosen[t]];
        } else if (mode == 1 && site > p) {
            m = 1;
        }
    }
    mask[t] = m;
    if (mode == 0) opidx[t] = o;
}
void jump_pair_mask(const int *state, const int *chosen, const int *cand_site, const int *cand_pair, int p, int mode,
                    int *mask, int *opidx, int B, hipStream_t s) {
    hipLaunchKernelGGL(jump_pair_mask_kernel, dim3((B + 255) / 256), dim3(256), 0, s, state, chosen, cand_site,
                       cand_pair, p, mode, mask, opidx, B);
    YQ_HIP(hipGetLastError());
}

void jump_opidx(const int *state, const int *chosen, const int *cand_site, const int *cand_op, const int *cand_site2,
                const int *cand_op2, int site, int *opidx, int B, hipStream_t s) {
    hipLaunchKernelGGL(jump_opidx_kernel, dim3((B + 255) / 256), dim3(256), 0, s, state, chosen, cand_site, cand_op,
                       cand_site2, cand_op2, site, opidx, B);
    YQ_HIP(hipGetLastError());
}

// mask[t] = (state[t] == v) or, with v < 0, state[t] != 0.
__global__ void mask_from_state_kernel(const int *state, int v, int *mask, int *count, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    const int m = (v >= 0) ? (state[t] == v) : (state[t] != 0);
    mask[t] = m;
    if (m && count) atomicAdd(count, 1);
}
void mask_from_state(const int *state, int v, int *mask, int *count, int B, hipStream_t s) {
    hipLaunchKernelGGL(mask_from_state_kernel, dim3((B + 255) / 256), dim3(256), 0, s, state, v, mask, count, B);
    YQ_HIP(hipGetLastError());
}

// Observables: out[q][col][t0+t] = Re Tr(O_q ρ_{site_q}) (2x2 or 4x4).
__global__ void observe_kernel(const c64 *rho1, int64_t sR1, const c64 *rho2, int64_t sR2, int nobs, const int *obs_site,
                               const int *obs_n, const c64 *obs_op, int col, int ncol, int64_t ntraj, int64_t t0,
                               double *out, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    for (int q = 0; q < nobs; ++q) {
        const int d = obs_n[q] == 1 ? 2 : 4;
        const c64 *r = (d == 2) ? rho1 + t * sR1 + (int64_t)obs_site[q] * 4 : rho2 + t * sR2 + (int64_t)obs_site[q] * 16;
        const c64 *O = obs_op + (int64_t)q * 16;
        double v = 0;
        for (int s = 0; s < d; ++s)
            for (int s2 = 0; s2 < d; ++s2) {
                const c64 o = O[s2 * d + s], p = r[s * d + s2];
                v += o.re * p.re - o.im * p.im;
            }
        out[((int64_t)q * ncol + col) * ntraj + t0 + t] = v;
    }
}
void observe(const c64 *rho1, int64_t sR1, const c64 *rho2, int64_t sR2, int nobs, const int *obs_site, const int *obs_n,
             const c64 *obs_op, int col, int ncol, int64_t ntraj, int64_t t0, double *out, int B, hipStream_t s) {
    hipLaunchKernelGGL(observe_kernel, dim3((B + 255) / 256), dim3(256), 0, s, rho1, sR1, rho2, sR2, nobs, obs_site, obs_n,
                       obs_op, col, ncol, ntraj, t0, out, B);
    YQ_HIP(hipGetLastError());
}

// dims[t][b] bookkeeping of each trajectory's true bond dimension.
__global__ void set_dims_kernel(int *dims, int ld, int b, const int *keep, int val, const int *mask, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    if (mask && !mask[t]) return;
    dims[(int64_t)t * ld + b] = keep ? keep[t] : val;
}
void set_dims(int *dims, int ld, int b, const int *keep, int val, const int *mask, int B, hipStream_t s) {
    hipLaunchKernelGGL(set_dims_kernel, dim3((B + 255) / 256), dim3(256), 0, s, dims, ld, b, keep, val, mask, B);
    YQ_HIP(hipGetLastError());
}

__global__ void fill_int_kernel(int *p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void fill_int(int *p, int v, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    YQ_HIP(hipGetLastError());
}

// One workgroup per trajectory: any non-finite diagonal entry flags the
// trajectory and replaces its matrix by e0 e0^T (finite, rank one, so the
// split keeps at most two columns and the wave's working dimension does not
// grow for it).
__global__ void guard_gram_kernel(c64 *G, int64_t sG, int n, int *fail, const int *mask) {
    __shared__ int bad;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    c64 *g = G + (int64_t)t * sG;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const c64 v = g[(int64_t)i * n + i];
        if (!isfinite(v.re) || !isfinite(v.im)) bad = 1;
    }
    __syncthreads();
    if (!bad) return;
    if (threadIdx.x == 0) fail[t] = 1;
    const int64_t nn = (int64_t)n * n;
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) g[e] = cmake(e == 0 ? 1.0 : 0.0, 0.0);
}

void guard_gram(c64 *G, int64_t sG, int n, int *fail, const int *mask, int B, hipStream_t s) {
    if (B <= 0 || n <= 0) return;
    hipLaunchKernelGGL(guard_gram_kernel, dim3(B), dim3(256), 0, s, G, sG, n, fail, mask);
    YQ_HIP(hipGetLastError());
}

template <typename T>
__global__ void poison_kernel(cplx<T> *A) {
    A->re = (T)NAN;
    A->im = (T)NAN;
}

template <typename T>
void poison_slot(cplx<T> *A, int64_t sA, int t, hipStream_t s) {
    hipLaunchKernelGGL(poison_kernel<T>, dim3(1), dim3(1), 0, s, A + (int64_t)t * sA);
    YQ_HIP(hipGetLastError());
}
template void poison_slot<float>(c32 *, int64_t, int, hipStream_t);
template void poison_slot<double>(c64 *, int64_t, int, hipStream_t);

// one wavefront: c = exp(-i tau T) e0 of the Lanczos stopping rule, for the
// test hook below (alpha: k values, beta: k values, beta[0] unused)
__global__ __launch_bounds__(64) void tridiag_exp_probe_kernel(const double *alpha, const double *beta, int k,
                                                                double tau, double2 *out) {
    const int lane = threadIdx.x;
    const double2 v = tridiag_exp_lane(alpha, beta, k, tau, lane, alpha[k - 1]);
    if (lane < k) out[lane] = v;
}

}  // namespace yq

// Test hook (not part of the public header): the device tridiagonal
// exponential of the Lanczos stopping rule, c = exp(-i tau T) e0 for T with
// diagonal alpha[0..k) and couplings beta[1..k) (host arrays, 1 <= k <= 64);
// out: k complex values as interleaved doubles.
extern "C" int yaqs_debug_tridiag_exp(int k, const double *alpha, const double *beta, double tau, double *out) {
    using namespace yq;
    if (k < 1 || k > 64) return 1;
    try {
        double *da, *db;
        double2 *dout;
        YQ_HIP(hipMalloc(&da, k * sizeof(double)));
        YQ_HIP(hipMalloc(&db, (k + 1) * sizeof(double)));
        YQ_HIP(hipMalloc(&dout, k * sizeof(double2)));
        YQ_HIP(hipMemcpy(da, alpha, k * sizeof(double), hipMemcpyHostToDevice));
        YQ_HIP(hipMemset(db, 0, (k + 1) * sizeof(double)));
        YQ_HIP(hipMemcpy(db, beta, k * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(tridiag_exp_probe_kernel, dim3(1), dim3(64), 0, 0, da, db, k, tau, dout);
        YQ_HIP(hipGetLastError());
        YQ_HIP(hipDeviceSynchronize());
        YQ_HIP(hipMemcpy(out, dout, k * sizeof(double2), hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dout);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}

// Test hook (not part of the public header): CGS2 of `count` matrices of
// nvec vectors of length len (host arrays of interleaved doubles, one matrix
// after the other), as rows (lq = 1: nvec x len row-major) or as columns
// (lq = 0: len x nvec row-major); A is overwritten with Q, Cf receives the
// nvec x nvec coefficients.  prec 0 = FP64, 1 = FP32.
extern "C" int yaqs_debug_cgs2(int prec, int nvec, int len, int lq, int count, double *A, double *Cf) {
    using namespace yq;
    try {
        const int64_t na = (int64_t)nvec * len, nc = (int64_t)nvec * nvec;
        const int64_t ns = cgs2_scratch_elems(nvec, len);
        auto run = [&](auto *tag) {
            typedef std::remove_pointer_t<decltype(tag)> T;
            std::vector<cplx<T>> ha(na * count), hc(nc * count);
            for (int64_t i = 0; i < na * count; ++i) ha[i] = cmake<T>((T)A[2 * i], (T)A[2 * i + 1]);
            cplx<T> *da, *dc, *dsc;
            double *df;
            YQ_HIP(hipMalloc(&da, na * count * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&dc, nc * count * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&dsc, ns * count * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&df, count * sizeof(double)));
            YQ_HIP(hipMemcpy(da, ha.data(), na * count * sizeof(cplx<T>), hipMemcpyHostToDevice));
            if (lq) cgs2<T>(da, na, nvec, len, len, 1, dc, nc, dsc, ns, df, nullptr, count, 0);
            else cgs2<T>(da, na, nvec, len, 1, nvec, dc, nc, dsc, ns, df, nullptr, count, 0);
            YQ_HIP(hipDeviceSynchronize());
            YQ_HIP(hipMemcpy(ha.data(), da, na * count * sizeof(cplx<T>), hipMemcpyDeviceToHost));
            YQ_HIP(hipMemcpy(hc.data(), dc, nc * count * sizeof(cplx<T>), hipMemcpyDeviceToHost));
            (void)hipFree(da);
            (void)hipFree(dc);
            (void)hipFree(dsc);
            (void)hipFree(df);
            for (int64_t i = 0; i < na * count; ++i) {
                A[2 * i] = (double)ha[i].re;
                A[2 * i + 1] = (double)ha[i].im;
            }
            for (int64_t i = 0; i < nc * count; ++i) {
                Cf[2 * i] = (double)hc[i].re;
                Cf[2 * i + 1] = (double)hc[i].im;
            }
        };
        if (prec == 1) run((float *)nullptr);
        else run((double *)nullptr);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}

// Calibration hook (not part of the public header): copies `bytes` (a
// multiple of 16) between two device buffers with one plain 16-byte-per-lane
// copy kernel, `reps` times, so that hardware traffic counters (FETCH_SIZE,
// WRITE_SIZE) can be scaled against a launch of known traffic.
__global__ __launch_bounds__(256) void copy16_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int yaqs_debug_copy_bytes(int64_t bytes, int reps) {
    using namespace yq;
    if (bytes <= 0 || (bytes & 15) || reps < 1) return 1;
    try {
        double2 *a = nullptr, *b = nullptr;
        YQ_HIP(hipMalloc(&a, bytes));
        YQ_HIP(hipMalloc(&b, bytes));
        YQ_HIP(hipMemset(a, 0, bytes));
        const int64_t n = bytes / 16;
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(copy16_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
        YQ_HIP(hipGetLastError());
        YQ_HIP(hipDeviceSynchronize());
        (void)hipFree(a);
        (void)hipFree(b);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}

namespace yq {

// ------------------------------------------------ split: deflated tail --
// Kept eigenvalues of a Gram matrix below tau * lambda_max (sorted
// descending) carry singular values below sqrt(tau) of the largest, whose
// eigenvectors the Gram route cannot resolve (their eigenvalues sit at the
// rounding level of G).  dcount[t] = kept eigenvalues above the cut, flag[t]
// = 1 when some kept ones are below it (and at least one above), counted in
// *count.
__global__ void tail_flags_kernel(const double *evals, int64_t sE, const int *keep, double tau, int *dcount, int *flag,
                                  int *count, int B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    const double *ev = evals + t * sE;
    const int k = keep[t];
    int d = 0;
    while (d < k && ev[d] > tau * ev[0]) ++d;
    const int f = (d >= 1 && d < k) ? 1 : 0;
    dcount[t] = d;
    flag[t] = f;
    if (f) atomicAdd(count, 1);
}
void tail_flags(const double *evals, int64_t sE, const int *keep, double tau, int *dcount, int *flag, int *count, int B,
                hipStream_t s) {
    hipLaunchKernelGGL(tail_flags_kernel, dim3((B + 255) / 256), dim3(256), 0, s, evals, sE, keep, tau, dcount, flag,
                       count, B);
    YQ_HIP(hipGetLastError());
}

// D = A - C (n elements per slot, masked)
template <typename T>
__global__ void sub_kernel(const cplx<T> *A, int64_t sA, const cplx<T> *C, int64_t sC, cplx<T> *D, int64_t sD, int64_t n,
                           const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) st2(D[t * sD + e], csub(ld2(A[t * sA + e]), ld2(C[t * sC + e])));
}
template <typename T>
void sub(const cplx<T> *A, int64_t sA, const cplx<T> *C, int64_t sC, cplx<T> *D, int64_t sD, int64_t n, const int *mask,
         int B, hipStream_t s) {
    hipLaunchKernelGGL(sub_kernel<T>, dim3((unsigned)((n + 255) / 256), B), dim3(256), 0, s, A, sA, C, sC, D, sD, n, mask);
    YQ_HIP(hipGetLastError());
}
template void sub<float>(const c32 *, int64_t, const c32 *, int64_t, c32 *, int64_t, int64_t, const int *, int, hipStream_t);
template void sub<double>(const c64 *, int64_t, const c64 *, int64_t, c64 *, int64_t, int64_t, const int *, int,
                          hipStream_t);

// Columns (mode 0) or rows (mode 1, conjugated) k in [dcount, keep) of Out
// from the leading eigenvectors of U (sorted by perm): the isometry written
// by write_iso with its resolved part kept and its tail replaced.
template <typename T>
__global__ void tail_fix_kernel(cplx<T> *Out, int64_t sO, int n, int kw, int mode, const int *dcount, const int *keep,
                                const c64 *U, int64_t sU, const int *perm, int64_t sE, const int *mask) {
    const int t = blockIdx.y;
    if (mask && !mask[t]) return;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * kw) return;
    int r, k;
    if (mode == 0) { r = (int)(e / kw); k = (int)(e % kw); }
    else { k = (int)(e / n); r = (int)(e % n); }
    const int d = dcount[t];
    if (k < d || k >= keep[t]) return;
    const c64 u = U[t * sU + (int64_t)r * n + perm[t * sE + (k - d)]];
    st2(Out[t * sO + e], make_double2(u.re, mode == 0 ? u.im : -u.im));
}
template <typename T>
void tail_fix(cplx<T> *Out, int64_t sO, int n, int kw, int mode, const int *dcount, const int *keep, const c64 *U,
              int64_t sU, const int *perm, int64_t sE, const int *mask, int B, hipStream_t s) {
    const int64_t tot = (int64_t)n * kw;
    hipLaunchKernelGGL(tail_fix_kernel<T>, dim3((unsigned)((tot + 255) / 256), B), dim3(256), 0, s, Out, sO, n, kw, mode,
                       dcount, keep, U, sU, perm, sE, mask);
    YQ_HIP(hipGetLastError());
}
template void tail_fix<float>(c32 *, int64_t, int, int, int, const int *, const int *, const c64 *, int64_t, const int *,
                              int64_t, const int *, int, hipStream_t);
template void tail_fix<double>(c64 *, int64_t, int, int, int, const int *, const int *, const c64 *, int64_t, const int *,
                               int64_t, const int *, int, hipStream_t);

}  // namespace yq
