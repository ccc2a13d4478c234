// Batched complex GEMM on the CDNA4 matrix cores (gfx950).
//
// Complex products are split into real MFMAs on LDS-staged tiles.  FP32
// takes three real products per complex one (the 3M method):
//   P1 = Re A Re B, P2 = Im A Im B, P3 = (Re A + Im A)(Re B + Im B),
//   Re C = P1 - P2, Im C = P3 - P1 - P2,
// FP64 the four of the definition:
//   Re C += Re A Re B - Im A Im B,  Im C += Re A Im B + Im A Re B.
// A 256-thread workgroup owns a BM x BN tile of C (BM, BN in {32, 64}, the
// shape chosen per GEMM shape by the engine's tuner); each of its four
// wavefronts owns a quarter as (BM/32) x (BN/32) blocks of 16x16 MFMA outputs
// (v_mfma_f32_16x16x4f32 for complex-FP32, v_mfma_f64_16x16x4f64 for
// complex-FP64).  Tiles of A and B are staged through LDS split into real
// and imaginary planes; the op (N, T, C) of each operand is a template
// parameter.
#include <hip/hip_runtime.h>

#include "kernels.h"

#include <cstdint>
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

// Arguments of the MPO step fused into a GEMM epilogue; W's non-zeros are
// staged in LDS (at most W_NNZ_MAX, else the unfused kernels run).  The
// epilogue accumulates in the contraction precision T with W rounded to T,
// in wstep_kernel's order, so fused and unfused results are identical.
constexpr int W_NNZ_MAX = 256;
struct WFuse {
    int P = 0, Dl = 1, dph = 1, Dr = 1, Q = 0, pb = 0, qb = 0;
    int64_t sY = 0;
    const int *rowptr = nullptr, *col = nullptr;
    const c64 *val = nullptr;
    void *Y = nullptr;
};

// Element (m, k) of op(A) and (k, n) of op(B) in row-major storage.
template <int OP>
__device__ __forceinline__ int64_t a_index(int m, int k, int64_t ld) {
    return OP == OP_N ? (int64_t)m * ld + k : (int64_t)k * ld + m;
}
template <int OP>
__device__ __forceinline__ int64_t b_index(int k, int n, int64_t ld) {
    return OP == OP_N ? (int64_t)k * ld + n : (int64_t)n * ld + k;
}

// Tile BM x BN of C per 256-thread workgroup (2 x 2 wavefronts, each
// (BM/2) x (BN/2) as 16 x 16 MFMA blocks), K in steps of BK.  Every output
// element sees the same sequence of MFMAs whatever the tile shape, so all
// tile shapes give bit-identical results.  LDS rows hold
// one row of op(A) (or one column of op(B)) with k permuted as
// (k % 4) * (BK / 4) + k / 4, so the BK / 4 values a lane feeds to the MFMAs
// of consecutive k steps are contiguous.  The next k tile is loaded into
// registers while the current one is multiplied (two LDS buffers, one
// barrier per tile); loads outside the matrix read element 0 and are
// zeroed at the LDS store, so there is no branch between the loads.
//
// FW: the MPO step is fused into the epilogue (WFuse).  The GEMM is then
// T1[(p, m)][(s, q)] = A B with rows (p, m), m < Dl, and columns (s, q),
// s < dph, q < Q; a tile of BM = 16 Dl rows covers all m of pb = 16
// consecutive p and all s of qb = 64 / dph consecutive q (no padded rows),
// stages T1 in LDS and writes
// Y[p][(s', n)][q] = sum W[(s', n)][(m, s)] T1[(p, m)][(s, q)] (W in CSR,
// rows s' Dr + n, columns m dph + s) straight to memory.
constexpr int GEMM_XCDS = 8;  // accelerator complex dies of an MI355X
inline unsigned gemm_blocks(int ntile, int inner, int outer) {
    return (unsigned)ntile * inner * ((outer + GEMM_XCDS - 1) / GEMM_XCDS * GEMM_XCDS);
}

// TI: storage type of A and B (float with T = double: complex-FP32 operands
// converted to FP64 as they are staged, FP64 products and sums, as the
// Gram matrices of the bond splits take them).  With d.lower set, the C
// tiles strictly above the diagonal are not computed.
template <typename T, int OPA, int OPB, int BM, int BN, int BK, bool FW, bool DL = false, typename TI = T>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmDesc d, const cplx<TI> *__restrict__ A,
                                                        const cplx<TI> *__restrict__ B, cplx<T> *__restrict__ C,
                                                        WFuse w, int ntile) {
    constexpr int LDK = BK + 4, KQ = BK / 4;  // k values per lane and tile
    // 16 x 16 blocks per wave: 2 x 2 wavefronts of (BM/2) x (BN/2), or for
    // the fused kernel 1 x 4 wavefronts of BM x 16 (BM = 16 Dl rows)
    constexpr int IM = FW ? BM / 16 : BM / 32, JN = FW ? 1 : BN / 32;
    typedef typename Mfma<T>::V V;
    // XCD-aware block order: workgroups are dealt round-robin to the XCDs,
    // so block b runs on XCD b % GEMM_XCDS; all tiles of trajectory t are
    // given to blocks on XCD t % GEMM_XCDS, where they share that XCD's L2
    // (the A and B operands of a trajectory are fetched from HBM once).
    const int units = ntile * d.inner;
    const int xcd = blockIdx.x % GEMM_XCDS, kx = blockIdx.x / GEMM_XCDS;
    const int t = (kx / units) * GEMM_XCDS + xcd, u = kx % units;
    const int jb = u / ntile, tile = u % ntile;
    if (t >= d.outer) return;
    if (d.active && !d.active[t]) return;
    int m0, n0, p0 = 0, q0 = 0, rows_ok = BM;
    if (FW) {
        const int tilesQ = (w.Q + w.qb - 1) / w.qb;
        p0 = (tile / tilesQ) * w.pb;
        q0 = (tile % tilesQ) * w.qb;
        m0 = p0 * w.Dl;
        n0 = 0;
        rows_ok = w.pb * w.Dl;
    } else {
        const int tilesN = (d.N + BN - 1) / BN;
        m0 = (tile / tilesN) * BM;
        n0 = (tile % tilesN) * BN;
    }
    static_assert(std::is_same<TI, T>::value || (!FW && !DL), "converted operands: register-staged, unfused");
    if (!FW && d.lower && n0 >= m0 + BM) return;
    const cplx<TI> *Ab = A + t * d.sA_outer + jb * d.sA_inner;
    const cplx<TI> *Bb = B + t * d.sB_outer + jb * d.sB_inner;
    constexpr int SM_AB = 2 * 2 * (BM + BN) * LDK;
    // operand buffers; the fused epilogue reuses them for T1 and appends W
    constexpr int SM_W = FW ? ((65 + W_NNZ_MAX) * (int)sizeof(int) + W_NNZ_MAX * (int)sizeof(c64)) / (int)sizeof(T) + 4 : 0;
    constexpr int SM_T1W = FW ? BM * (BN + 1) * 2 + SM_W : 0;
    // DL: k-tile stages of the operands as stored (complex interleaved);
    // three stages of 16 k, or four of 8 k (the fused kernel: 5 wavefronts'
    // worth of LDS less, so three workgroups fit a CU)
    constexpr int DL_NS = BK == 16 ? 3 : 4;
    constexpr int DL_SA = BM * BK * 8, DL_SB = BK * 512, DL_ST = DL_SA + DL_SB;
    constexpr int SM_DL = DL ? DL_NS * DL_ST / (int)sizeof(T) : 0;
    constexpr int SM_1 = SM_AB > SM_T1W ? SM_AB : SM_T1W;
    __shared__ __align__(1024) T smem[SM_1 > SM_DL ? SM_1 : SM_DL];
    T(*As)[2][BM][LDK] = reinterpret_cast<T(*)[2][BM][LDK]>(smem);            // [buffer][re/im][m][k']
    T(*Bs)[2][BN][LDK] = reinterpret_cast<T(*)[2][BN][LDK]>(smem + 2 * 2 * BM * LDK);  // [buffer][re/im][n][k']
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = FW ? 0 : (wave >> 1) * (BM / 2), wn = FW ? wave * 16 : (wave & 1) * (BN / 2);
    const int li = lane & 15, lk = lane >> 4;
    // Two load mappings, chosen per operand by which index is contiguous in
    // memory.  Row-contiguous (op(A) with A stored K x M, op(B) = B): thread t
    // owns row (t % BM) and k group g = t / BM, i.e. the values
    // k = g + (256 / BM) q (for BM = 64 the four that sit contiguously in its
    // permuted LDS row, one conflict-free 16-byte store per plane), loads
    // coalesced along the rows.  K-contiguous (op(A) = A, op(B) with B stored
    // N x K): element e = t + 256 r is (row e / BK, k e % BK), loads
    // coalesced along k and scalar LDS stores.
    static_assert(FW || ((BM == 32 || BM == 64) && (BN == 32 || BN == 64)), "tile shape outside the load mappings");
    static_assert(!FW || (BM % 16 == 0 && BM <= 80 && BN == 64 && OPA == OP_N && OPB == OP_N),
                  "the fused kernel takes op N operands, 64 columns and 16 Dl rows");
    static_assert(BK == 16 || (DL && BK == 8), "k tile");
    // loads per thread and k tile of the register-staged path (DL: unused)
    constexpr int NLA = DL ? 1 : BM * BK / 256, NLB = DL ? 1 : BN * BK / 256;
    constexpr int GA = 256 / BM, GB = 256 / BN;              // k groups of the row mapping
    constexpr bool ROWA = OPA != OP_N, ROWB = OPB == OP_N;
    int ar_[NLA], ak_[NLA], br_[NLB], bk_[NLB];
#pragma unroll
    for (int q = 0; q < NLA; ++q) {
        const int e = tid + 256 * q;
        ar_[q] = ROWA ? tid % BM : e / BK;
        ak_[q] = ROWA ? tid / BM + GA * q : e % BK;
    }
#pragma unroll
    for (int q = 0; q < NLB; ++q) {
        const int e = tid + 256 * q;
        br_[q] = ROWB ? tid % BN : e / BK;
        bk_[q] = ROWB ? tid / BN + GB * q : e % BK;
    }
    // global column of each B element this thread loads (FW: tile column
    // j -> (s = j / qb, q = q0 + j % qb) -> s Q + q)
    int bgn[NLB];
    bool bcol_ok[NLB], arow_ok[NLA];
#pragma unroll
    for (int q = 0; q < NLB; ++q) {
        if (FW) {
            const int sj = br_[q] / w.qb, qq = q0 + br_[q] % w.qb;
            bgn[q] = sj * w.Q + qq;
            bcol_ok[q] = sj < w.dph && qq < w.Q;
        } else {
            bgn[q] = n0 + br_[q];
            bcol_ok[q] = bgn[q] < d.N;
        }
    }
#pragma unroll
    for (int q = 0; q < NLA; ++q) arow_ok[q] = ar_[q] < rows_ok && m0 + ar_[q] < d.M;
    cplx<TI> ra[NLA], rb[NLB];
    unsigned oka = 0, okb = 0;  // in-bounds flags of the pending loads, applied at the LDS store
    auto load = [&](int k0) {
        oka = okb = 0;
#pragma unroll
        for (int q = 0; q < NLA; ++q) {
            const int gm = m0 + ar_[q], gk = k0 + ak_[q];
            const bool ok = arow_ok[q] && gk < d.K;
            ra[q] = Ab[ok ? a_index<OPA>(gm, gk, d.lda) : 0];
            oka |= (ok ? 1u : 0u) << q;
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const int gn = bgn[q], gk2 = k0 + bk_[q];
            const bool ok2 = bcol_ok[q] && gk2 < d.K;
            rb[q] = Bb[ok2 ? b_index<OPB>(gk2, gn, d.ldb) : 0];
            okb |= (ok2 ? 1u : 0u) << q;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < NLA; ++q) {
            const bool ok = (oka >> q) & 1u;
            const int ka = (ak_[q] & 3) * KQ + (ak_[q] >> 2);
            As[buf][0][ar_[q]][ka] = ok ? (T)ra[q].re : (T)0;
            As[buf][1][ar_[q]][ka] = ok ? (T)(OPA == OP_C ? -ra[q].im : ra[q].im) : (T)0;
        }
#pragma unroll
        for (int q = 0; q < NLB; ++q) {
            const bool ok2 = (okb >> q) & 1u;
            const int kb = (bk_[q] & 3) * KQ + (bk_[q] >> 2);
            Bs[buf][0][br_[q]][kb] = ok2 ? (T)rb[q].re : (T)0;
            Bs[buf][1][br_[q]][kb] = ok2 ? (T)(OPB == OP_C ? -rb[q].im : rb[q].im) : (T)0;
        }
    };
    // complex-FP32: three real products per complex one (P1 = Re A Re B,
    // P2 = Im A Im B, P3 = (Re A + Im A)(Re B + Im B); Re C = P1 - P2,
    // Im C = P3 - P1 - P2); complex-FP64: the four real products
    constexpr bool G3M = std::is_same<T, float>::value;
    V accr[IM][JN], acci[IM][JN], accs[IM][JN];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) accr[i][j][r] = acci[i][j][r] = accs[i][j][r] = (T)0;
    if constexpr (DL) {
        // Direct-to-LDS k tiles (complex-FP32, op N operands, whole tiles):
        // each wavefront load instruction (global_load_lds_dwordx4) moves
        // 64 x 16 bytes, two consecutive complex elements per lane, into a
        // 1 KB block of the stage; no registers hold the tiles, so DL_NS - 1
        // k tiles are in flight while one is multiplied.  A stage holds A as
        // BM rows of CPR = BK/2 16-byte chunks (chunk kc of row m at position
        // kc ^ fa(m), fa(m) = m % 8 for 8 chunks a row, (m / 4) % 4 for 4)
        // and B as BK k rows of 32 chunks (chunk nc of row k at position
        // nc ^ 8 (k % 4)), so the lanes of an MFMA operand read hit distinct
        // banks.  The MFMA sequence per output element is the one of the
        // register-staged path (same bits).
        static_assert(std::is_same<T, float>::value && OPA == OP_N && OPB == OP_N && BN == 64,
                      "direct-to-LDS tiles: complex-FP32, op N, 64 columns");
        constexpr int NS = DL_NS, CPR = BK / 2, RPB = 1024 / (BK * 8);  // chunks a row, rows a block
        constexpr int NBA = BM / RPB, NBB = BK / 2;        // 1 KB blocks of A and B per stage
        constexpr int NA = (NBA + 3) / 4, NB = NBB / 4;    // load instructions per wave (at most)
        constexpr int NMIN = NBA / 4 + NB;                 // per tile, fewest of any wave
        static_assert(BM % RPB == 0 && NBB % 4 == 0, "whole 1 KB blocks, B blocks per wave");
        auto fa = [](int m) { return CPR == 8 ? (m & 7) : ((m >> 2) & 3); };
        char *sm = reinterpret_cast<char *>(smem);
        int64_t aoff[NA], boff[NB];
#pragma unroll
        for (int v = 0; v < NA; ++v) {
            const int m = RPB * (wave + 4 * v) + lane / CPR, kc = (lane % CPR) ^ fa(m);
            aoff[v] = (int64_t)(m0 + m) * d.lda + 2 * kc;
        }
#pragma unroll
        for (int v = 0; v < NB; ++v) {
            const int k = 2 * (wave + 4 * v) + (lane >> 5), nc = (lane & 31) ^ (8 * (k & 3));
            int col = n0 + 2 * nc;
            if (FW) col = (2 * nc / w.qb) * w.Q + q0 + (2 * nc) % w.qb;
            boff[v] = (int64_t)k * d.ldb + col;
        }
        // The loads are issued by inline asm: the compiler would otherwise
        // treat them as LDS writes in flight and wait for all of them before
        // every LDS read.  The explicit vmcnt wait and the barrier of each
        // step order them before the reads of their stage.  M0 carries the
        // LDS address (the compiler keeps nothing in M0 in this kernel: LDS
        // instructions on gfx950 do not read it); the instruction writes no
        // register.
        auto dma = [&](const cplx<TI> *g, const char *l) {
            const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)l);
            asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(la), "v"(g)
                         : "memory");
        };
        auto issue = [&](int kt, int st) {
            const int k0 = kt * BK;
            char *sa = sm + st * DL_ST, *sb = sa + DL_SA;
#pragma unroll
            for (int v = 0; v < NA; ++v)
                if (wave + 4 * v < NBA) dma(Ab + aoff[v] + k0, sa + (wave + 4 * v) * 1024);
#pragma unroll
            for (int v = 0; v < NB; ++v) dma(Bb + boff[v] + (int64_t)k0 * d.ldb, sb + (wave + 4 * v) * 1024);
        };
        auto compute = [&](int st) {
            const char *sa = sm + st * DL_ST, *sb = sa + DL_SA;
            T ar[IM][KQ], ai[IM][KQ], br[JN][KQ], bi[JN][KQ], as[IM][KQ], bs[JN][KQ];
#pragma unroll
            for (int i = 0; i < IM; ++i) {
                const int m = wm + 16 * i + li;
#pragma unroll
                for (int q = 0; q < KQ; ++q) {
                    const int kc = 2 * q + (lk >> 1);
                    const float2 v = *reinterpret_cast<const float2 *>(sa + m * (BK * 8) + ((kc ^ fa(m)) << 4) +
                                                                       ((lk & 1) << 3));
                    ar[i][q] = v.x;
                    ai[i][q] = v.y;
                    as[i][q] = v.x + v.y;
                }
            }
#pragma unroll
            for (int j = 0; j < JN; ++j) {
                const int n = wn + 16 * j + li;
#pragma unroll
                for (int q = 0; q < KQ; ++q) {
                    const int k = 4 * q + lk;
                    const float2 v = *reinterpret_cast<const float2 *>(sb + k * 512 + (((n >> 1) ^ (8 * (k & 3))) << 4) +
                                                                       ((n & 1) << 3));
                    br[j][q] = v.x;
                    bi[j][q] = v.y;
                    bs[j][q] = v.x + v.y;
                }
            }
#pragma unroll
            for (int q = 0; q < KQ; ++q)
#pragma unroll
                for (int i = 0; i < IM; ++i)
#pragma unroll
                    for (int j = 0; j < JN; ++j) {
                        accr[i][j] = Mfma<T>::mma(ar[i][q], br[j][q], accr[i][j]);
                        acci[i][j] = Mfma<T>::mma(ai[i][q], bi[j][q], acci[i][j]);
                        accs[i][j] = Mfma<T>::mma(as[i][q], bs[j][q], accs[i][j]);
                    }
        };
        const int nk = d.K / BK;
        for (int kt = 0; kt < NS - 1 && kt < nk; ++kt) issue(kt, kt);
        int st = 0;
        for (int kt = 0; kt < nk; ++kt) {
            // tile kt has landed (the loads of the later tiles already issued,
            // NMIN or more per tile, may be in flight) and this wave's LDS
            // reads of the previous step are done
            const int later = min(NS - 2, nk - 1 - kt);
            if (later >= 2) __builtin_amdgcn_s_waitcnt((2 * NMIN) | (0x7 << 4));
            else if (later == 1) __builtin_amdgcn_s_waitcnt(NMIN | (0x7 << 4));
            else __builtin_amdgcn_s_waitcnt(0x7 << 4);
            __builtin_amdgcn_s_barrier();
            if (kt + NS - 1 < nk) issue(kt + NS - 1, st == 0 ? NS - 1 : st - 1);
            compute(st);
            st = st == NS - 1 ? 0 : st + 1;
        }
        if (FW) __syncthreads();  // the epilogue reuses the stages
    } else {
        const int nk = (d.K + BK - 1) / BK;
        load(0);
        store(0);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int buf = kt & 1;
            if (kt + 1 < nk) load((kt + 1) * BK);
            T ar[IM][KQ], ai[IM][KQ], br[JN][KQ], bi[JN][KQ];
    #pragma unroll
            for (int i = 0; i < IM; ++i)
    #pragma unroll
                for (int q = 0; q < KQ; ++q) {
                    ar[i][q] = As[buf][0][wm + i * 16 + li][lk * KQ + q];
                    ai[i][q] = As[buf][1][wm + i * 16 + li][lk * KQ + q];
                }
    #pragma unroll
            for (int j = 0; j < JN; ++j)
    #pragma unroll
                for (int q = 0; q < KQ; ++q) {
                    br[j][q] = Bs[buf][0][wn + j * 16 + li][lk * KQ + q];
                    bi[j][q] = Bs[buf][1][wn + j * 16 + li][lk * KQ + q];
                }
            if constexpr (G3M) {
                T as[IM][KQ], bs[JN][KQ];
    #pragma unroll
                for (int i = 0; i < IM; ++i)
    #pragma unroll
                    for (int q = 0; q < KQ; ++q) as[i][q] = ar[i][q] + ai[i][q];
    #pragma unroll
                for (int j = 0; j < JN; ++j)
    #pragma unroll
                    for (int q = 0; q < KQ; ++q) bs[j][q] = br[j][q] + bi[j][q];
    #pragma unroll
                for (int q = 0; q < KQ; ++q)
    #pragma unroll
                    for (int i = 0; i < IM; ++i)
    #pragma unroll
                        for (int j = 0; j < JN; ++j) {
                            accr[i][j] = Mfma<T>::mma(ar[i][q], br[j][q], accr[i][j]);
                            acci[i][j] = Mfma<T>::mma(ai[i][q], bi[j][q], acci[i][j]);
                            accs[i][j] = Mfma<T>::mma(as[i][q], bs[j][q], accs[i][j]);
                        }
            } else {
    #pragma unroll
                for (int q = 0; q < KQ; ++q)
    #pragma unroll
                    for (int i = 0; i < IM; ++i)
    #pragma unroll
                        for (int j = 0; j < JN; ++j) {
                            accr[i][j] = Mfma<T>::mma(ar[i][q], br[j][q], accr[i][j]);
                            acci[i][j] = Mfma<T>::mma(ar[i][q], bi[j][q], acci[i][j]);
                        }
    #pragma unroll
                for (int q = 0; q < KQ; ++q)
    #pragma unroll
                    for (int i = 0; i < IM; ++i)
    #pragma unroll
                        for (int j = 0; j < JN; ++j) {
                            accr[i][j] = Mfma<T>::mma(-ai[i][q], bi[j][q], accr[i][j]);
                            acci[i][j] = Mfma<T>::mma(ai[i][q], br[j][q], acci[i][j]);
                        }
            }
            if (kt + 1 < nk) store(buf ^ 1);
            __syncthreads();
        }
    }
    if constexpr (G3M) {
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
            for (int j = 0; j < JN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const T p1 = accr[i][j][r], p2 = acci[i][j][r];
                    accr[i][j][r] = p1 - p2;
                    acci[i][j][r] = (accs[i][j][r] - p1) - p2;
                }
    }
    if (FW) {
        // stage T1 in LDS (the operand buffers are free after the last barrier)
        static_assert(SM_T1W <= (int)(sizeof(smem) / sizeof(T)), "T1 and W stage exceed the shared buffer");
        cplx<T>(*T1s)[BN + 1] = reinterpret_cast<cplx<T>(*)[BN + 1]>(smem);
#pragma unroll
        for (int i = 0; i < IM; ++i)
#pragma unroll
            for (int j = 0; j < JN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    T1s[wm + i * 16 + Mfma<T>::row(lk, r)][wn + j * 16 + li] = cmake<T>(accr[i][j][r], acci[i][j][r]);
        __syncthreads();
        // W in CSR staged next to T1 (at most 64 rows and W_NNZ_MAX non-zeros)
        const int J = w.dph * w.Dr;
        static_assert(BM * (BN + 1) * 2 * sizeof(T) % 16 == 0, "W stage alignment");
        c64 *wval = reinterpret_cast<c64 *>(smem + BM * (BN + 1) * 2);
        int *wrp = reinterpret_cast<int *>(wval + W_NNZ_MAX);
        int *wcol = wrp + 65;
        const int nnz = w.rowptr[J];
        for (int i = tid; i <= J; i += 256) wrp[i] = w.rowptr[i];
        for (int i = tid; i < nnz; i += 256) {
            const int c = w.col[i];
            wcol[i] = (c / w.dph) * (BN + 1) + (c % w.dph) * w.qb;  // offset of (m, s) in the T1 stage
            wval[i] = w.val[i];
        }
        __syncthreads();
        // qb lanes per (group of PG rows p, row of W): q fastest for coalesced
        // stores; each W entry read from LDS serves PG independent rows
        constexpr int PG = 4;
        const int qb = w.qb, qq = tid % qb, slot = tid / qb, nslot = 256 / qb;
        const int q = q0 + qq;
        cplx<T> *Yb = reinterpret_cast<cplx<T> *>(w.Y) + t * w.sY;
        const cplx<T> *T1f = &T1s[0][0];
        const int ngp = (w.pb + PG - 1) / PG;  // pb = 16: no partial group
        const int rs = w.Dl * (BN + 1);        // T1 stage stride of p
        if (q < w.Q) {
            for (int u = slot; u < ngp * J; u += nslot) {
                const int pg = u / J, jj = u - pg * J;
                const cplx<T> *base = T1f + pg * PG * rs + qq;
                T ax[PG], ay[PG];  // the contraction precision (as wstep_kernel)
#pragma unroll
                for (int r = 0; r < PG; ++r) ax[r] = ay[r] = (T)0;
                for (int z = wrp[jj]; z < wrp[jj + 1]; ++z) {
                    const c64 wv = wval[z];
                    const T wr = (T)wv.re, wi = (T)wv.im;
                    const int off = wcol[z];
#pragma unroll
                    for (int r = 0; r < PG; ++r) {
                        const cplx<T> x = base[r * rs + off];
                        // same operation order as wstep_kernel
                        ax[r] = fma_t(wr, x.re, ax[r]);
                        ax[r] = fma_t(-wi, x.im, ax[r]);
                        ay[r] = fma_t(wr, x.im, ay[r]);
                        ay[r] = fma_t(wi, x.re, ay[r]);
                    }
                }
#pragma unroll
                for (int r = 0; r < PG; ++r) {
                    const int p = p0 + pg * PG + r;
                    if (p < w.P) Yb[((int64_t)p * J + jj) * w.Q + q] = cmake<T>((T)ax[r], (T)ay[r]);
                }
            }
        }
        return;
    }
    cplx<T> *Cb = C + t * d.sC_outer + jb * d.sC_inner;
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
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

template <typename T, int OA, int OB, int BM, int BN, bool DL = false, int BK = 16>
static void launch_tile(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C, hipStream_t s) {
    const int tiles = ((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
    hipLaunchKernelGGL((gemm_mfma_kernel<T, OA, OB, BM, BN, BK, false, DL>), dim3(gemm_blocks(tiles, d.inner, d.outer)),
                       dim3(256), 0, s, d, A, B, C, WFuse{}, tiles);
    YQ_HIP(hipGetLastError());
}

// The direct-to-LDS k tiles need complex-FP32 op N operands, whole BM x 64
// C tiles and k tiles of 16, and 16-byte aligned element pairs (even
// leading dimensions and batch strides, 16-byte aligned base pointers).
template <typename T>
static bool dl_fits(const GemmDesc &d, const void *A, const void *B, int BM) {
    if (!std::is_same<T, float>::value || d.opA != OP_N || d.opB != OP_N) return false;
    if (d.M % BM || d.N % 64 || d.K % 16) return false;
    if ((d.lda | d.ldb | d.sA_outer | d.sB_outer | d.sA_inner | d.sB_inner) & 1) return false;
    return (((uintptr_t)A | (uintptr_t)B) & 15) == 0;
}

template <typename T, int OA, int OB>
static void launch_shape(int tile, const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C,
                         hipStream_t s) {
    if constexpr (OB == OP_N || OB == OP_C) {
        switch (tile) {
            case 1: return launch_tile<T, OA, OB, 32, 64>(d, A, B, C, s);
            case 2: return launch_tile<T, OA, OB, 64, 32>(d, A, B, C, s);
            case 3: return launch_tile<T, OA, OB, 32, 32>(d, A, B, C, s);
            case 4:
            case 5:
                if constexpr (std::is_same<T, float>::value && OA == OP_N && OB == OP_N) {
                    if (dl_fits<T>(d, A, B, 64)) {
                        if (tile == 4) return launch_tile<T, OA, OB, 64, 64, true>(d, A, B, C, s);
                        return launch_tile<T, OA, OB, 64, 64, true, 8>(d, A, B, C, s);
                    }
                }
                break;
            default: break;
        }
    }
    launch_tile<T, OA, OB, 64, 64>(d, A, B, C, s);
}

template <typename T>
bool gemm_mfma(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C, hipStream_t s, int tile) {
    if (!mfma_mode()) return false;
    if (d.M <= 0 || d.N <= 0 || d.K <= 0 || d.outer <= 0 || d.inner <= 0) return false;
#define YQ_GEMM_CASE(OA, OB)                                    \
    if (d.opA == OA && d.opB == OB) {                          \
        launch_shape<T, OA, OB>(tile, d, A, B, C, s);          \
        return true;                                           \
    }
    YQ_GEMM_CASE(OP_N, OP_N)
    YQ_GEMM_CASE(OP_N, OP_T)
    YQ_GEMM_CASE(OP_N, OP_C)
    YQ_GEMM_CASE(OP_T, OP_N)
    YQ_GEMM_CASE(OP_T, OP_T)
    YQ_GEMM_CASE(OP_T, OP_C)
    YQ_GEMM_CASE(OP_C, OP_N)
    YQ_GEMM_CASE(OP_C, OP_T)
    YQ_GEMM_CASE(OP_C, OP_C)
#undef YQ_GEMM_CASE
    return false;
}
// FP64 Gram matrices of complex-FP32 operands (A Y^H or Y^H A of the bond
// split): FP64 matrix cores on operands converted as they are staged.
bool gemm_mfma_gram(const GemmDesc &d, const c32 *A, const c32 *B, c64 *C, hipStream_t s) {
    if (!mfma_mode()) return false;
    if (d.M <= 0 || d.N <= 0 || d.K <= 0 || d.outer <= 0 || d.inner <= 0) return false;
    const int tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64);
    const dim3 grid(gemm_blocks(tiles, d.inner, d.outer));
    if (d.opA == OP_N && d.opB == OP_C)
        hipLaunchKernelGGL((gemm_mfma_kernel<double, OP_N, OP_C, 64, 64, 16, false, false, float>), grid, dim3(256), 0,
                           s, d, A, B, C, WFuse{}, tiles);
    else if (d.opA == OP_C && d.opB == OP_N)
        hipLaunchKernelGGL((gemm_mfma_kernel<double, OP_C, OP_N, 64, 64, 16, false, false, float>), grid, dim3(256), 0,
                           s, d, A, B, C, WFuse{}, tiles);
    else
        return false;
    YQ_HIP(hipGetLastError());
    return true;
}

bool gemm_w_covers(const GemmDesc &d, int P, int Dl, int dph, int Dr, int Q, int nnz) {
    if (!mfma_mode()) return false;
    if (nnz > W_NNZ_MAX) return false;  // W staged in LDS
    if (d.opA != OP_N || d.opB != OP_N || d.inner != 1 || d.accumulate) return false;
    if (Dl < 1 || Dl > 5 || (dph != 2 && dph != 4) || d.M != P * Dl || d.N != dph * Q) return false;
    if (dph * Dr > 64) return false;  // W rows staged in LDS
    if (P <= 0 || Q <= 0 || d.K <= 0 || d.outer <= 0) return false;
    return true;
}

template <typename T>
bool gemm_dl_fits(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B) {
    return mfma_mode() && d.M > 0 && d.N > 0 && d.K > 0 && dl_fits<T>(d, A, B, 64);
}

template <typename T>
bool gemm_w_dl_fits(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, int Dl) {
    return Dl >= 1 && Dl <= 5 && dl_fits<T>(d, A, B, 16 * Dl);
}

template <typename T>
bool gemm_w_mfma(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, int P, int Dl, int dph, int Dr, int Q,
                 cplx<T> *Y, int64_t sY, const int *rowptr, const int *col, const c64 *val, int nnz, hipStream_t s,
                 bool dl) {
    if (!gemm_w_covers(d, P, Dl, dph, Dr, Q, nnz)) return false;
    dl = dl && gemm_w_dl_fits<T>(d, A, B, Dl);
    WFuse w;
    w.P = P;
    w.Dl = Dl;
    w.dph = dph;
    w.Dr = Dr;
    w.Q = Q;
    w.pb = 16;
    w.qb = 64 / dph;
    w.sY = sY;
    w.rowptr = rowptr;
    w.col = col;
    w.val = val;
    w.Y = Y;
    const int tiles = ((P + w.pb - 1) / w.pb) * ((Q + w.qb - 1) / w.qb);
    const dim3 grid(gemm_blocks(tiles, 1, d.outer));
#define YQ_GEMM_W(DL)                                                                                          \
    case DL:                                                                                                   \
        if constexpr (std::is_same<T, float>::value) {                                                         \
            if (dl) {                                                                                          \
                hipLaunchKernelGGL((gemm_mfma_kernel<T, OP_N, OP_N, 16 * DL, 64, 8, true, true>), grid, dim3(256),  \
                                   0, s, d, A, B, (cplx<T> *)nullptr, w, tiles);                               \
                break;                                                                                         \
            }                                                                                                  \
        }                                                                                                      \
        hipLaunchKernelGGL((gemm_mfma_kernel<T, OP_N, OP_N, 16 * DL, 64, 16, true>), grid, dim3(256), 0, s, d, A,  \
                           B, (cplx<T> *)nullptr, w, tiles);                                                  \
        break;
    switch (Dl) {
        YQ_GEMM_W(1)
        YQ_GEMM_W(2)
        YQ_GEMM_W(3)
        YQ_GEMM_W(4)
        YQ_GEMM_W(5)
        default: return false;
    }
#undef YQ_GEMM_W
    YQ_HIP(hipGetLastError());
    return true;
}
template bool gemm_w_mfma<float>(const GemmDesc &, const c32 *, const c32 *, int, int, int, int, int, c32 *, int64_t,
                                 const int *, const int *, const c64 *, int, hipStream_t, bool);
template bool gemm_w_mfma<double>(const GemmDesc &, const c64 *, const c64 *, int, int, int, int, int, c64 *, int64_t,
                                  const int *, const int *, const c64 *, int, hipStream_t, bool);
template bool gemm_dl_fits<float>(const GemmDesc &, const c32 *, const c32 *);
template bool gemm_dl_fits<double>(const GemmDesc &, const c64 *, const c64 *);
template bool gemm_w_dl_fits<float>(const GemmDesc &, const c32 *, const c32 *, int);
template bool gemm_w_dl_fits<double>(const GemmDesc &, const c64 *, const c64 *, int);
template bool gemm_mfma<float>(const GemmDesc &, const c32 *, const c32 *, c32 *, hipStream_t, int);
template bool gemm_mfma<double>(const GemmDesc &, const c64 *, const c64 *, c64 *, hipStream_t, int);

}  // namespace yq

// Test hooks (not part of the public header).
//
// yaqs_debug_gemm_tile: one complex GEMM on the device through the
// matrix-core kernel with C tile `tile` (0 .. GEMM_TILES-1), or the generic
// kernel for tile < 0.  A, B, C are host arrays of interleaved doubles,
// row-major, with op(A) M x K and op(B) K x N; prec 0 = FP64, 1 = FP32.
extern "C" int yaqs_debug_gemm_tile(int prec, int M, int N, int K, int opA, int opB, const double *A, const double *B,
                                    double *Cout, int tile) {
    using namespace yq;
    try {
        if (tile >= GEMM_TILES) return 1;
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
            if (!(tile >= 0 && gemm_mfma<T>(d, da, db, dc, 0, tile))) gemm<T, T, T, T>(d, da, db, dc, 0);
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

extern "C" int yaqs_debug_gemm(int prec, int M, int N, int K, int opA, int opB, const double *A, const double *B,
                               double *Cout, int use_mfma) {
    return yaqs_debug_gemm_tile(prec, M, N, K, opA, opB, A, B, Cout, use_mfma ? 0 : -1);
}

// yaqs_debug_gram: FP64 Gram matrices of `count` complex-FP32 matrices Y
// (M x K row-major, given in FP64 and rounded to FP32): G = Y Y^H (form 0)
// or Y^H Y (form 1; Y is then K x M), through the FP64 matrix cores
// (mfma = 1) or the FP64 generic kernel (mfma = 0).  With lower set the
// tiles above the diagonal are skipped; G is cleared first.  Returns 0, or
// 1 when the matrix-core kernel did not launch.
extern "C" int yaqs_debug_gram(int M, int K, int form, int lower, int count, const double *Y, double *Gout,
                               int mfma) {
    using namespace yq;
    try {
        const int64_t ny = (int64_t)M * K, ng = (int64_t)M * M;
        GemmDesc d;
        d.M = M; d.N = M; d.K = K;
        d.opA = form == 0 ? OP_N : OP_C;
        d.opB = form == 0 ? OP_C : OP_N;
        d.lda = d.ldb = form == 0 ? K : M;
        d.ldc = M;
        d.sA_outer = d.sB_outer = ny;
        d.sC_outer = ng;
        d.sA_inner = d.sB_inner = d.sC_inner = 0;
        d.inner = 1; d.outer = count; d.accumulate = 0; d.active = nullptr; d.lower = lower;
        std::vector<c32> hy(ny * count);
        for (int64_t i = 0; i < ny * count; ++i) hy[i] = cmake<float>((float)Y[2 * i], (float)Y[2 * i + 1]);
        c32 *dy;
        c64 *dg;
        YQ_HIP(hipMalloc(&dy, ny * count * sizeof(c32)));
        YQ_HIP(hipMalloc(&dg, ng * count * sizeof(c64)));
        YQ_HIP(hipMemcpy(dy, hy.data(), ny * count * sizeof(c32), hipMemcpyHostToDevice));
        YQ_HIP(hipMemset(dg, 0, ng * count * sizeof(c64)));
        bool ok = true;
        if (mfma) ok = gemm_mfma_gram(d, dy, dy, dg, 0);
        else gemm<float, float, double, double>(d, dy, dy, dg, 0);
        YQ_HIP(hipDeviceSynchronize());
        YQ_HIP(hipMemcpy(Gout, dg, ng * count * sizeof(c64), hipMemcpyDeviceToHost));
        (void)hipFree(dy);
        (void)hipFree(dg);
        return ok ? 0 : 1;
    } catch (const std::exception &) {
        return 3;
    }
}

// yaqs_debug_gemm_w: T1 = A B (A: (P Dl) x K, B: K x (dph Q), row-major)
// followed by the MPO step with W in CSR (rows J = dph Dr, columns
// I = Dl dph, index m dph + s), through the fused kernel (Yf) and through
// the unfused GEMM + wstep (Yu), and, when Yd is not null, through the
// fused kernel with k tiles loaded straight into LDS (Yd); Y is P x J x Q.
// Returns 0, or 1 when the fused kernel does not cover the shape, plus 4
// when Yd was asked for and the direct-to-LDS kernel does not apply.
extern "C" int yaqs_debug_gemm_w(int prec, int P, int Dl, int dph, int Dr, int Q, int K, const double *A,
                                 const double *B, const int *rowptr, const int *col, const double *val, double *Yf,
                                 double *Yu, double *Yd) {
    using namespace yq;
    try {
        const int M = P * Dl, N = dph * Q, J = dph * Dr, nnz = rowptr[J];
        const int64_t na = (int64_t)M * K, nbb = (int64_t)K * N, nc = (int64_t)M * N, ny = (int64_t)P * J * Q;
        GemmDesc d;
        d.M = M; d.N = N; d.K = K; d.opA = OP_N; d.opB = OP_N;
        d.lda = K; d.ldb = N; d.ldc = N;
        d.sA_outer = d.sB_outer = d.sC_outer = 0;
        d.sA_inner = d.sB_inner = d.sC_inner = 0;
        d.inner = 1; d.outer = 1; d.accumulate = 0; d.active = nullptr;
        int rc = 0;
        auto run = [&](auto *tag) {
            typedef std::remove_pointer_t<decltype(tag)> T;
            std::vector<cplx<T>> ha(na), hb(nbb), hy(ny);
            for (int64_t i = 0; i < na; ++i) ha[i] = cmake<T>((T)A[2 * i], (T)A[2 * i + 1]);
            for (int64_t i = 0; i < nbb; ++i) hb[i] = cmake<T>((T)B[2 * i], (T)B[2 * i + 1]);
            cplx<T> *da, *db, *dt, *dy;
            int *drp, *dcol;
            c64 *dval;
            YQ_HIP(hipMalloc(&da, na * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&db, nbb * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&dt, nc * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&dy, ny * sizeof(cplx<T>)));
            YQ_HIP(hipMalloc(&drp, (J + 1) * sizeof(int)));
            YQ_HIP(hipMalloc(&dcol, (nnz + 1) * sizeof(int)));
            YQ_HIP(hipMalloc(&dval, (nnz + 1) * sizeof(c64)));
            YQ_HIP(hipMemcpy(da, ha.data(), na * sizeof(cplx<T>), hipMemcpyHostToDevice));
            YQ_HIP(hipMemcpy(db, hb.data(), nbb * sizeof(cplx<T>), hipMemcpyHostToDevice));
            YQ_HIP(hipMemcpy(drp, rowptr, (J + 1) * sizeof(int), hipMemcpyHostToDevice));
            YQ_HIP(hipMemcpy(dcol, col, nnz * sizeof(int), hipMemcpyHostToDevice));
            YQ_HIP(hipMemcpy(dval, val, nnz * sizeof(c64), hipMemcpyHostToDevice));
            auto fetch = [&](double *out) {
                YQ_HIP(hipDeviceSynchronize());
                YQ_HIP(hipMemcpy(hy.data(), dy, ny * sizeof(cplx<T>), hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < ny; ++i) {
                    out[2 * i] = (double)hy[i].re;
                    out[2 * i + 1] = (double)hy[i].im;
                }
            };
            if (gemm_w_mfma<T>(d, da, db, P, Dl, dph, Dr, Q, dy, 0, drp, dcol, dval, nnz, 0)) fetch(Yf);
            else rc = 1;
            if (Yd) {
                if (rc == 0 && gemm_w_dl_fits<T>(d, da, db, Dl) &&
                    gemm_w_mfma<T>(d, da, db, P, Dl, dph, Dr, Q, dy, 0, drp, dcol, dval, nnz, 0, true))
                    fetch(Yd);
                else rc |= 4;
            }
            if (!gemm_mfma<T>(d, da, db, dt, 0, 0)) gemm<T, T, T, T>(d, da, db, dt, 0);
            wstep<T>(dt, dy, P, Dl * dph, J, Q, 0, 0, drp, dcol, dval, nullptr, 1, 0);
            fetch(Yu);
            (void)hipFree(da);
            (void)hipFree(db);
            (void)hipFree(dt);
            (void)hipFree(dy);
            (void)hipFree(drp);
            (void)hipFree(dcol);
            (void)hipFree(dval);
        };
        if (prec == 1) run((float *)nullptr);
        else run((double *)nullptr);
        return rc;
    } catch (const std::exception &) {
        return 3;
    }
}
