// Batched complex GEMM on the CDNA4 matrix cores (gfx950).
//
// Complex products are split into real MFMAs on LDS-staged tiles:
//   Re C += Re A Re B - Im A Im B,  Im C += Re A Im B + Im A Re B.
// A 256-thread workgroup owns a 64x64 tile of C; each of its four
// wavefronts owns a 32x32 quarter as 2x2 blocks of 16x16 MFMA outputs
// (v_mfma_f32_16x16x4f32 for complex-FP32, v_mfma_f64_16x16x4f64 for
// complex-FP64).  Tiles of A and B are staged through LDS split into real
// and imaginary planes, K-major, so the 16 lanes that share a k read 16
// consecutive LDS words.
#include <hip/hip_runtime.h>

#include "kernels.h"

#include <cstdlib>
#include <type_traits>
#include <vector>

namespace yq {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Operand layout (both): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16].
// Output register r of lane l holds D[row][l % 16] with
//   f32: row = 4 (l / 16) + r        f64: row = 4 r + l / 16
// (measured on gfx950 with one-hot operands).
template <>
struct Mfma<float> {
    typedef f32x4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk * 4 + r; }
};
template <>
struct Mfma<double> {
    typedef f64x4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return r * 4 + lk; }
};

template <typename T>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmDesc d, const cplx<T> *__restrict__ A,
                                                        const cplx<T> *__restrict__ B, cplx<T> *__restrict__ C) {
    constexpr int BM = 64, BN = 64, BK = 16, PAD = 4;
    typedef typename Mfma<T>::V V;
    const int t = blockIdx.z, jb = blockIdx.y;
    if (d.active && !d.active[t]) return;
    const int tilesN = (d.N + BN - 1) / BN;
    const int m0 = (blockIdx.x / tilesN) * BM, n0 = (blockIdx.x % tilesN) * BN;
    const cplx<T> *Ab = A + t * d.sA_outer + jb * d.sA_inner;
    const cplx<T> *Bb = B + t * d.sB_outer + jb * d.sB_inner;
    cplx<T> *Cb = C + t * d.sC_outer + jb * d.sC_inner;
    __shared__ T As_r[BK][BM + PAD], As_i[BK][BM + PAD], Bs_r[BK][BN + PAD], Bs_i[BK][BN + PAD];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    V accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) accr[i][j][r] = acci[i][j][r] = (T)0;
    for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + 256 * r;
            int row, kk;
            if (d.opA == OP_N) { row = e / BK; kk = e % BK; }
            else { kk = e / BM; row = e % BM; }
            const int gi = m0 + row, gk = k0 + kk;
            T vr = 0, vi = 0;
            if (gi < d.M && gk < d.K) {
                const cplx<T> z = (d.opA == OP_N) ? Ab[(int64_t)gi * d.lda + gk] : Ab[(int64_t)gk * d.lda + gi];
                vr = z.re;
                vi = (d.opA == OP_C) ? -z.im : z.im;
            }
            As_r[kk][row] = vr;
            As_i[kk][row] = vi;
            int col;
            if (d.opB == OP_N) { kk = e / BN; col = e % BN; }
            else { col = e / BK; kk = e % BK; }
            const int gj = n0 + col, gk2 = k0 + kk;
            T wr = 0, wi = 0;
            if (gj < d.N && gk2 < d.K) {
                const cplx<T> z = (d.opB == OP_N) ? Bb[(int64_t)gk2 * d.ldb + gj] : Bb[(int64_t)gj * d.ldb + gk2];
                wr = z.re;
                wi = (d.opB == OP_C) ? -z.im : z.im;
            }
            Bs_r[kk][col] = wr;
            Bs_i[kk][col] = wi;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kk = ks * 4 + lk;
            T ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ar[i] = As_r[kk][wm + i * 16 + li];
                ai[i] = As_i[kk][wm + i * 16 + li];
                nai[i] = -ai[i];
                br[i] = Bs_r[kk][wn + i * 16 + li];
                bi[i] = Bs_i[kk][wn + i * 16 + li];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accr[i][j] = Mfma<T>::mma(ar[i], br[j], accr[i][j]);
                    accr[i][j] = Mfma<T>::mma(nai[i], bi[j], accr[i][j]);
                    acci[i][j] = Mfma<T>::mma(ar[i], bi[j], acci[i][j]);
                    acci[i][j] = Mfma<T>::mma(ai[i], br[j], acci[i][j]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                const int gj = n0 + wn + j * 16 + li;
                if (gi < d.M && gj < d.N) {
                    cplx<T> &c = Cb[(int64_t)gi * d.ldc + gj];
                    if (d.accumulate) {
                        c.re += accr[i][j][r];
                        c.im += acci[i][j][r];
                    } else {
                        c.re = accr[i][j][r];
                        c.im = acci[i][j][r];
                    }
                }
            }
}

static int mfma_mode() {
    static int mode = -1;
    if (mode < 0) {
        const char *e = getenv("YAQS_GEMM");
        mode = (e && e[0] == 'g') ? 0 : 1;  // "generic" disables the matrix-core path
    }
    return mode;
}

template <typename T>
bool gemm_mfma(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C, hipStream_t s) {
    if (!mfma_mode()) return false;
    if (d.M <= 0 || d.N <= 0 || d.K <= 0 || d.outer <= 0 || d.inner <= 0) return false;
    const int tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64);
    dim3 grid(tiles, d.inner, d.outer);
    hipLaunchKernelGGL(gemm_mfma_kernel<T>, grid, dim3(256), 0, s, d, A, B, C);
    YQ_HIP(hipGetLastError());
    return true;
}
template bool gemm_mfma<float>(const GemmDesc &, const c32 *, const c32 *, c32 *, hipStream_t);
template bool gemm_mfma<double>(const GemmDesc &, const c64 *, const c64 *, c64 *, hipStream_t);

}  // namespace yq

// Test hook (not part of the public header): one complex GEMM on the device
// through either the matrix-core or the generic kernel, for the GEMM unit
// tests.  A, B, C are host arrays of interleaved doubles, row-major, with
// op(A) M x K and op(B) K x N; prec 0 = FP64, 1 = FP32.
extern "C" int yaqs_debug_gemm(int prec, int M, int N, int K, int opA, int opB, const double *A, const double *B,
                               double *Cout, int use_mfma) {
    using namespace yq;
    try {
        const int64_t na = (int64_t)M * K, nbb = (int64_t)K * N, nc = (int64_t)M * N;
        GemmDesc d;
        d.M = M; d.N = N; d.K = K; d.opA = opA; d.opB = opB;
        d.lda = opA == OP_N ? K : M;
        d.ldb = opB == OP_N ? N : K;
        d.ldc = N;
        d.sA_outer = d.sB_outer = d.sC_outer = 0;
        d.sA_inner = d.sB_inner = d.sC_inner = 0;
        d.inner = 1; d.outer = 1; d.accumulate = 0; d.active = nullptr;
        auto run = [&](auto *tag) {
            typedef std::remove_pointer_t<decltype(tag)> T;
            std::vector<cplx<T>> ha(na), hb(nbb), hc(nc);
            for (int64_t i = 0; i < na; ++i) ha[i] = cmake<T>((T)A[2 * i], (T)A[2 * i + 1]);
            for (int64_t i = 0; i < nbb; ++i) hb[i] = cmake<T>((T)B[2 * i], (T)B[2 * i + 1]);
            cplx<T> *da, *db, *dc;
            YQ_HIP(hipMalloc(&da, na * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&db, nbb * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&dc, nc * sizeof(cplx<T>)));
            YQ_HIP(hipMemcpy(da, ha.data(), na * sizeof(cplx<T>), hipMemcpyHostToDevice));
            YQ_HIP(hipMemcpy(db, hb.data(), nbb * sizeof(cplx<T>), hipMemcpyHostToDevice));
            if (!(use_mfma && gemm_mfma<T>(d, da, db, dc, 0))) gemm<T, T, T, T>(d, da, db, dc, 0);
            YQ_HIP(hipDeviceSynchronize());
            YQ_HIP(hipMemcpy(hc.data(), dc, nc * sizeof(cplx<T>), hipMemcpyDeviceToHost));
            (void)hipFree(da);
            (void)hipFree(db);
            (void)hipFree(dc);
            for (int64_t i = 0; i < nc; ++i) {
                Cout[2 * i] = (double)hc[i].re;
                Cout[2 * i + 1] = (double)hc[i].im;
            }
        };
        if (prec == 1) run((float *)nullptr);
        else run((double *)nullptr);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}
