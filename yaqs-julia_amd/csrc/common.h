// Shared types and helpers of the MI355X TJM engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace yq {

// Complex scalar in interleaved (re, im) storage; T = float or double.
template <typename T>
struct cplx {
    T re, im;
};
using c32 = cplx<float>;
using c64 = cplx<double>;

template <typename T>
__host__ __device__ inline cplx<T> cmake(T r, T i) {
    cplx<T> z;
    z.re = r;
    z.im = i;
    return z;
}

// Fused multiply-add in the operands' own precision (no promotion of
// float arguments to double).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

struct HipError : std::runtime_error {
    explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define YQ_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::yq::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                                 std::to_string(__LINE__));                                            \
    } while (0)

enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

// Batched GEMM description: C = op(A) * op(B) (+ C if accumulate), with
// per-matrix pointers base + outer*t + inner*j.  Row-major, leading dims in
// elements; op(A) is M x K, op(B) is K x N.
struct GemmDesc {
    int M, N, K;
    int opA, opB;
    int64_t lda, ldb, ldc;
    int64_t sA_outer, sB_outer, sC_outer;
    int64_t sA_inner, sB_inner, sC_inner;
    int inner;       // inner batch count
    int outer;       // outer batch count (trajectories)
    int accumulate;  // 0: C = AB, 1: C += AB
    const int *active;  // optional per-outer mask
    int lower = 0;      // generic kernel: C tiles strictly above the diagonal are not computed
};

}  // namespace yq
