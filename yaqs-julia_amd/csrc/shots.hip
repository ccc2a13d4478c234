// Shot sampling of an MPS (single_shot_measure / measure_shots,
// src/MPS.jl:683-749), batched over shots.
//
// The reference moves the centre to site 1 and then, per shot and site,
// takes the outcome probabilities as the slice norms of the current site
// tensor, draws, projects, renormalises and contracts the projected slice
// into the next site.  Here the MPS is brought to right-canonical form once
// on the device (an LQ sweep L..2 with the CGS2 gauge step), so for every
// shot the probabilities of site i given its earlier outcomes are
// p_s = ||v A_i[:, s, :]||^2 with v the shot's normalised left vector.  Per
// site one FP64 GEMM forms W = V A_i for a chunk of shots (V: shots x dl,
// A_i: dl x 2 dr), and one kernel draws, records the bit and writes the next
// left vectors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "kernels.h"
#include "reduce.h"
#include "rng.h"
#include "shots.h"

namespace yq {

namespace {

struct DMem {
    void *p = nullptr;
    explicit DMem(size_t n) {
        if (n) YQ_HIP(hipMalloc(&p, n));
    }
    ~DMem() {
        if (p) (void)hipFree(p);
    }
    DMem(const DMem &) = delete;
    DMem &operator=(const DMem &) = delete;
    template <class X>
    X *as() const {
        return static_cast<X *>(p);
    }
};

GemmDesc desc(int M, int N, int K, int64_t lda, int64_t ldb, int64_t ldc) {
    GemmDesc d;
    d.M = M; d.N = N; d.K = K; d.opA = OP_N; d.opB = OP_N;
    d.lda = lda; d.ldb = ldb; d.ldc = ldc;
    d.sA_outer = d.sB_outer = d.sC_outer = 0;
    d.sA_inner = d.sB_inner = d.sC_inner = 0;
    d.inner = 1; d.outer = 1; d.accumulate = 0; d.active = nullptr;
    return d;
}

constexpr int SHOT_WAVES = 4;  // shots per 256-thread block, one wavefront each

// One wavefront per shot (t is uniform in a wavefront, so whole wavefronts
// leave together).  W[t] holds the shot's two slices [s][b], b < dr.  The
// outcome is the first s with r <= p_0 + ... + p_s, else the last one, as
// the reference's loop leaves `chosen`; the chosen slice is divided by
// sqrt(p_chosen) unless that is below 1e-12 (src/MPS.jl:716-719).
__global__ __launch_bounds__(SHOT_WAVES * 64) void shot_step_kernel(const c64 *__restrict__ W, int dr,
                                                                     c64 *__restrict__ V, int64_t ldv, int site,
                                                                     int last, uint64_t seed, uint64_t shot0,
                                                                     int64_t nshots, int64_t *__restrict__ bits) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * SHOT_WAVES + (threadIdx.x >> 6);
    if (t >= nshots) return;
    const c64 *w = W + t * 2 * dr;
    double p[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        double acc = 0.0;
        for (int b = lane; b < dr; b += 64) {
            const c64 z = w[s * dr + b];
            acc += z.re * z.re + z.im * z.im;
        }
        p[s] = wave_sum(acc);
    }
    const double r = philox_uniform_dev(seed, shot0 + (uint64_t)t, (uint64_t)site);
    // at site 0 the weights sum to <ψ|ψ>: the draw uses them normalised (the
    // reference's raw weights presume a normalised state); later sites sum
    // to 1 through the renormalised slices
    double q0 = p[0];
    if (site == 0 && p[0] + p[1] > 0.0) q0 = p[0] / (p[0] + p[1]);
    const int chosen = r <= q0 ? 0 : 1;
    if (lane == 0) bits[t] |= (int64_t)chosen << site;
    if (last) return;
    const double nrm = sqrt(p[chosen]);
    const double f = nrm > 1e-12 ? 1.0 / nrm : 1.0;
    for (int b = lane; b < dr; b += 64) {
        const c64 z = w[chosen * dr + b];
        V[t * ldv + b] = cmake(z.re * f, z.im * f);
    }
}

__global__ void shot_init_kernel(c64 *V, int64_t ldv, int64_t n, int64_t *bits) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    V[t * ldv] = cmake(1.0, 0.0);
    bits[t] = 0;
}

}  // namespace

void measure_shots(int L, const std::vector<std::vector<c64>> &A, const std::vector<int> &dl,
                   const std::vector<int> &dr, int64_t shots, uint64_t seed, int64_t *bits_out, hipStream_t s) {
    if (L < 1 || L > 63) throw std::invalid_argument("measure_shots: L must be 1..63 (bits of an Int64)");
    if (shots < 0) throw std::invalid_argument("measure_shots: shots < 0");
    if (dl[0] != 1 || dr[L - 1] != 1) throw std::invalid_argument("measure_shots: boundary bonds must be 1");
    int chi = 1;
    std::vector<int64_t> off(L + 1, 0);
    for (int i = 0; i < L; ++i) {
        if (i > 0 && dl[i] != dr[i - 1]) throw std::invalid_argument("measure_shots: bond dimensions do not match");
        chi = std::max(chi, std::max(dl[i], dr[i]));
        off[i + 1] = off[i] + (int64_t)dl[i] * 2 * dr[i];
    }
    // site tensors on the device, right-canonical after an LQ sweep L..2
    DMem dA((size_t)off[L] * sizeof(c64));
    for (int i = 0; i < L; ++i)
        YQ_HIP(hipMemcpyAsync(dA.as<c64>() + off[i], A[i].data(), A[i].size() * sizeof(c64), hipMemcpyHostToDevice, s));
    int64_t scr = 1;
    for (int i = 1; i < L; ++i) scr = std::max<int64_t>(scr, cgs2_scratch_elems(dl[i], 2 * dr[i]));
    DMem dCf((size_t)chi * chi * sizeof(c64)), dS((size_t)scr * sizeof(c64)), dF(sizeof(double)),
        dT((size_t)2 * chi * chi * sizeof(c64));
    for (int i = L - 1; i >= 1; --i) {
        c64 *Ai = dA.as<c64>() + off[i], *Ap = dA.as<c64>() + off[i - 1];
        const int c0 = dl[i], c1 = dr[i], cm = dl[i - 1];
        cgs2<double>(Ai, 0, c0, 2 * c1, 2 * c1, 1, dCf.as<c64>(), 0, dS.as<c64>(), 0, dF.as<double>(), nullptr, 1, s);
        gemm<double, double, double, double>(desc(2 * cm, c0, c0, c0, c0, c0), Ap, dCf.as<c64>(), dT.as<c64>(), s);
        YQ_HIP(hipMemcpyAsync(Ap, dT.as<c64>(), (size_t)2 * cm * c0 * sizeof(c64), hipMemcpyDeviceToDevice, s));
    }
    // shots in chunks: V (chunk x chi), W (chunk x 2 chi), bits (chunk)
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(shots, (int64_t)1 << 18));
    DMem dV((size_t)chunk * chi * sizeof(c64)), dW((size_t)chunk * 2 * chi * sizeof(c64)),
        dB((size_t)chunk * sizeof(int64_t));
    for (int64_t t0 = 0; t0 < shots; t0 += chunk) {
        const int64_t n = std::min(chunk, shots - t0);
        hipLaunchKernelGGL(shot_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dV.as<c64>(),
                           (int64_t)chi, n, dB.as<int64_t>());
        YQ_HIP(hipGetLastError());
        for (int i = 0; i < L; ++i) {
            gemm<double, double, double, double>(desc((int)n, 2 * dr[i], dl[i], chi, 2 * dr[i], 2 * dr[i]),
                                                 dV.as<c64>(), dA.as<c64>() + off[i], dW.as<c64>(), s);
            hipLaunchKernelGGL(shot_step_kernel, dim3((unsigned)((n + SHOT_WAVES - 1) / SHOT_WAVES)),
                               dim3(SHOT_WAVES * 64), 0, s, dW.as<c64>(), dr[i], dV.as<c64>(), (int64_t)chi, i,
                               i == L - 1 ? 1 : 0, seed, (uint64_t)t0, n, dB.as<int64_t>());
            YQ_HIP(hipGetLastError());
        }
        YQ_HIP(hipMemcpyAsync(bits_out + t0, dB.as<int64_t>(), (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    }
    YQ_HIP(hipStreamSynchronize(s));
}

}  // namespace yq
