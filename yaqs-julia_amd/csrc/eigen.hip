// Batched Hermitian eigensolver for the FP64 Gram matrices of the bond
// splits (the SVD of split_mps_tensor_svd, src/Algorithms.jl:409-445).
//
// One 256-thread workgroup per matrix:
//   1. load the lower triangle into LDS (packed, FP64 complex; n <= 128 uses
//      132 KB of the CU's 160 KB LDS, larger n fall back to global scratch);
//   2. Householder reduction to a real symmetric tridiagonal T (the
//      reflectors follow LAPACK's zhetd2/zlarfg conventions: Q = H_0 ... H_{n-2},
//      H_k = I - tau_k v_k v_k^H, subdiagonal made real);
//   3. implicit QL iterations with Wilkinson shifts on T, run by one thread,
//      which records each sweep's Givens rotations in LDS while every thread
//      applies them to its own row of Z (global scratch);
//   4. eigenvalues sorted descending, kept rank by the requested truncation
//      rule, and only the columns that can be kept back-transformed, u = Q z
//      (one wavefront per column, reflectors read from LDS).
// Output: U[:, r] for r < min(n, max_bond) in sorted order, evals (sorted),
// perm[r] = r, keep.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace yq {

namespace {

__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_mulc(double2 a, double2 b) {  // conj(a) b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }

__device__ __forceinline__ int pidx(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

__device__ double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double bsum(double v, double *sh) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    for (int i = 0; i < nw; ++i) r += sh[i];
    __syncthreads();
    return r;
}

constexpr int LDS_MAX_N = 128;

}  // namespace

// P: packed lower triangle (LDS or global), n*(n+1)/2 double2.
__global__ __launch_bounds__(256) void heev_kernel(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Zs, int64_t sZ,
                                                   double2 *Pg, int64_t sP, int n, int ns, double *evals, int *perm,
                                                   int64_t sE, int *keep, int rule, double thr, int max_bond,
                                                   const int *mask) {
    extern __shared__ double2 smem2[];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const bool in_lds = n <= LDS_MAX_N;
    const int np = n * (n + 1) / 2;
    const int64_t zoff = ((int64_t)n * n + 1) & ~(int64_t)1;  // keep tau 16-byte aligned
    double2 *P, *tau;
    char *tail;
    if (in_lds) {
        P = smem2;
        tau = smem2 + np;
        tail = (char *)(smem2 + np + n);
    } else {
        P = Pg + t * sP;
        tau = (double2 *)(Zs + t * sZ + zoff);
        tail = (char *)smem2;
    }
    // d[n], e[n], rotation c/s [n], rotation index / sort order [n]
    double *dd = (double *)tail;
    double *ee = dd + n;
    double *rc = ee + n;
    double *rs = rc + n;
    int *ri = (int *)(rs + n);
    __shared__ double sh[16];
    __shared__ int nrot_sh, fail_sh;
    const c64 *A = G + t * sG;
    double *Z = Zs + t * sZ;
    for (int e = threadIdx.x; e < np; e += blockDim.x) {
        // e -> (i, j) with j <= i
        int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        const int j = e - i * (i + 1) / 2;
        P[e] = make_double2(A[(int64_t)i * n + j].re, A[(int64_t)i * n + j].im);
    }
    __syncthreads();
    // ---------------------------------------------- tridiagonalisation
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1;  // length of x = A[k+1.., k]
        double xs = 0;
        for (int i = 1 + threadIdx.x; i < m; i += blockDim.x) {
            const double2 x = P[pidx(k + 1 + i, k)];
            xs += x.x * x.x + x.y * x.y;
        }
        xs = bsum(xs, sh);
        const double2 alpha = P[pidx(k + 1, k)];
        double2 tk = make_double2(0, 0);
        double beta = alpha.x;
        if (!(xs == 0.0 && alpha.y == 0.0)) {
            const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs);
            beta = alpha.x >= 0 ? -nr : nr;
            tk = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            // v = x / (alpha - beta), v_0 = 1
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double dn = den.x * den.x + den.y * den.y;
            const double2 inv = make_double2(den.x / dn, -den.y / dn);
            for (int i = 1 + threadIdx.x; i < m; i += blockDim.x) P[pidx(k + 1 + i, k)] = c_mul(P[pidx(k + 1 + i, k)], inv);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            tau[k] = tk;
            dd[k] = P[pidx(k, k)].x;
            ee[k] = beta;
            P[pidx(k + 1, k)] = make_double2(1.0, 0.0);  // v_0 (restored as the reflector's head)
        }
        __syncthreads();
        if (tk.x != 0.0 || tk.y != 0.0) {
            // p = tau A22 v ; A22 is the trailing m x m block (rows/cols k+1..)
            // p is stored temporarily in Z row 0 (as complex pairs)
            double2 *p = (double2 *)Z;
            for (int r = threadIdx.x; r < m; r += blockDim.x) {
                double2 acc = make_double2(0, 0);
                const int gr = k + 1 + r;
                for (int c = 0; c < m; ++c) {
                    const int gc = k + 1 + c;
                    const double2 a = gc <= gr ? P[pidx(gr, gc)] : c_conj(P[pidx(gc, gr)]);
                    const double2 z = c_mul(a, P[pidx(gc, k)]);
                    acc.x += z.x;
                    acc.y += z.y;
                }
                p[r] = c_mul(tk, acc);
            }
            __syncthreads();
            // alpha' = -1/2 tau (p^H v)
            double2 pv = make_double2(0, 0);
            for (int r = threadIdx.x; r < m; r += blockDim.x) {
                const double2 z = c_mulc(p[r], P[pidx(k + 1 + r, k)]);
                pv.x += z.x;
                pv.y += z.y;
            }
            pv.x = bsum(pv.x, sh);
            pv.y = bsum(pv.y, sh);
            const double2 al = c_mul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
            // w = p + al v  (in place in p)
            for (int r = threadIdx.x; r < m; r += blockDim.x) {
                const double2 z = c_mul(al, P[pidx(k + 1 + r, k)]);
                p[r] = make_double2(p[r].x + z.x, p[r].y + z.y);
            }
            __syncthreads();
            // A22 -= v w^H + w v^H (lower triangle)
            const int nl = m * (m + 1) / 2;
            for (int e = threadIdx.x; e < nl; e += blockDim.x) {
                int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
                while (i * (i + 1) / 2 > e) --i;
                while ((i + 1) * (i + 2) / 2 <= e) ++i;
                const int j = e - i * (i + 1) / 2;
                const double2 vi = P[pidx(k + 1 + i, k)], vj = P[pidx(k + 1 + j, k)];
                const double2 a = c_mul(vi, c_conj(p[j]));
                const double2 b = c_mul(p[i], c_conj(vj));
                double2 &x = P[pidx(k + 1 + i, k + 1 + j)];
                x = make_double2(x.x - a.x - b.x, x.y - a.y - b.y);
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        dd[n - 1] = P[pidx(n - 1, n - 1)].x;
        ee[n - 1] = 0.0;
        fail_sh = 0;
    }
    // Z = I
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) Z[e] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    // ---------------------------------------------- implicit QL
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            // thread 0 finds the split point and, if needed, computes one QL
            // sweep on (l..m), recording its rotations
            if (threadIdx.x == 0) {
                int m = l;
                for (; m < n - 1; ++m) {
                    const double sc = fabs(dd[m]) + fabs(dd[m + 1]);
                    if (fabs(ee[m]) <= eps * sc) break;
                }
                int nr = 0;
                if (m != l && iter < 80) {
                    // Wilkinson-type shift from the leading 2x2 of the unreduced block
                    double g = (dd[l + 1] - dd[l]) / (2.0 * ee[l]);
                    double r = hypot(g, 1.0);
                    g = dd[m] - dd[l] + ee[l] / (g + (g >= 0 ? r : -r));
                    double s = 1.0, c = 1.0, pp = 0.0;
                    int i = m - 1;
                    bool deflated = false;
                    for (; i >= l; --i) {
                        double f = s * ee[i], b = c * ee[i];
                        r = hypot(f, g);
                        ee[i + 1] = r;
                        if (r == 0.0) {  // early deflation
                            dd[i + 1] -= pp;
                            ee[m] = 0.0;
                            deflated = true;
                            break;
                        }
                        s = f / r;
                        c = g / r;
                        g = dd[i + 1] - pp;
                        r = (dd[i] - g) * s + 2.0 * c * b;
                        pp = s * r;
                        dd[i + 1] = g + pp;
                        g = c * r - b;
                        rc[nr] = c;
                        rs[nr] = s;
                        ri[nr] = i;
                        ++nr;
                    }
                    if (!deflated) {
                        dd[l] -= pp;
                        ee[l] = g;
                        ee[m] = 0.0;
                    }
                } else if (m != l) {
                    fail_sh = 1;
                }
                nrot_sh = (m == l || iter >= 80) ? -1 : nr;
            }
            __syncthreads();
            const int nr = nrot_sh;
            if (nr < 0) break;
            // apply the sweep's rotations to the rows of Z
            for (int row = threadIdx.x; row < n; row += blockDim.x) {
                double *z = Z + (int64_t)row * n;
                for (int q = 0; q < nr; ++q) {
                    const int i = ri[q];
                    const double c = rc[q], s = rs[q];
                    const double f = z[i + 1];
                    z[i + 1] = s * z[i] + c * f;
                    z[i] = c * z[i] - s * f;
                }
            }
            __syncthreads();
        }
    }
    // ---------------------------------------------- sort, rule, back-transform
    double *ev = evals + t * sE;
    int *pm = perm + t * sE;
    int *order = ri;  // reuse: order[r] = index of the r-th largest eigenvalue
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double di = dd[i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = dd[j];
            if (dj > di || (dj == di && j < i)) ++rank;
        }
        order[rank] = i;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        ev[r] = dd[order[r]];
        pm[r] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int kp = ns;
        if (rule == 0) {  // T-A (src/Algorithms.jl:422-438)
            double disc = 0;
            for (int k = ns; k >= 1; --k) {
                disc += fmax(ev[k - 1], 0.0);
                if (disc >= thr) { kp = k > 2 ? k : 2; break; }
            }
        } else if (rule == 1) {  // T-C (src/Dissipation.jl:123-136)
            double tot = 0;
            for (int k = 0; k < ns; ++k) tot += fmax(ev[k], 0.0);
            double cur = 0;
            for (int k = ns; k >= 1; --k) {
                cur += fmax(ev[k - 1], 0.0);
                if (cur > thr * tot) { kp = k; break; }
            }
        } else {  // T-D (src/DigitalTJM.jl:283-292)
            double cur = 0;
            for (int k = ns; k >= 1; --k) {
                const double s2 = fmax(ev[k - 1], 0.0);
                if (cur + s2 > thr) break;
                cur += s2;
                kp -= 1;
            }
        }
        if (kp > max_bond) kp = max_bond;
        if (kp < 1) kp = 1;
        if (kp > ns) kp = ns;
        keep[t] = kp;
    }
    __syncthreads();
    // back-transform the columns that can be kept: u = H_0 ... H_{n-2} z
    int kout = ns < max_bond ? ns : max_bond;
    if (kout > n) kout = n;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    c64 *Ut = U + t * sU;
    for (int r = wid; r < kout; r += nw) {
        const int zc = order[r];
        const int CH = (n + 63) / 64;  // rows per lane
        // the output column doubles as working storage
        for (int q = 0; q < CH; ++q) {
            const int row = lane + 64 * q;
            if (row < n) Ut[(int64_t)row * n + r] = cmake<double>(Z[(int64_t)row * n + zc], 0.0);
        }
        for (int k = n - 2; k >= 0; --k) {
            const double2 tk = tau[k];
            if (tk.x == 0.0 && tk.y == 0.0) continue;
            // s = v^H u[k+1..]
            double sx = 0, sy = 0;
            for (int q = 0; q < CH; ++q) {
                const int row = lane + 64 * q;
                if (row >= k + 1 && row < n) {
                    const double2 v = P[pidx(row, k)];
                    const c64 uu = Ut[(int64_t)row * n + r];
                    const double2 z = c_mulc(v, make_double2(uu.re, uu.im));
                    sx += z.x;
                    sy += z.y;
                }
            }
            sx = wsum(sx);
            sy = wsum(sy);
            const double2 f = c_mul(tk, make_double2(sx, sy));
            for (int q = 0; q < CH; ++q) {
                const int row = lane + 64 * q;
                if (row >= k + 1 && row < n) {
                    const double2 z = c_mul(P[pidx(row, k)], f);
                    c64 &uu = Ut[(int64_t)row * n + r];
                    uu.re -= z.x;
                    uu.im -= z.y;
                }
            }
        }
    }
}

void heev(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP, int n, int ns,
          double *evals, int *perm, int64_t sE, int *keep, int rule, double thr, int max_bond, const int *mask, int B,
          hipStream_t s) {
    if (n > 1024) throw HipError("heev: n > 1024");
    const int np = n * (n + 1) / 2;
    size_t shmem = (size_t)n * (4 * sizeof(double) + sizeof(int)) + 16;
    if (n <= LDS_MAX_N) shmem += (size_t)(np + n) * sizeof(double2);
    else if (!Pg) throw HipError("heev: no global scratch for n > 128");
    hipLaunchKernelGGL(heev_kernel, dim3(B), dim3(256), shmem, s, G, sG, U, sU, Z, sZ, Pg, sP, n, ns, evals, perm, sE,
                       keep, rule, thr, max_bond, mask);
    YQ_HIP(hipGetLastError());
}

}  // namespace yq

// Test hook (not part of the public header): eigen-decomposition of one
// Hermitian n x n matrix (row-major interleaved complex doubles) on the
// device; writes the sorted eigenvalues, U (n x n, the first min(n, max_bond)
// columns valid) and the kept rank.
extern "C" int yaqs_debug_heev(int n, const double *G, double *evals, double *U, int *keep_out, int rule, double thr,
                               int max_bond) {
    using namespace yq;
    try {
        const int64_t nn = (int64_t)n * n;
        c64 *dG, *dU;
        double *dZ, *dE;
        double2 *dP = nullptr;
        int *dperm, *dkeep;
        YQ_HIP(hipMalloc(&dG, nn * sizeof(c64)));
        YQ_HIP(hipMalloc(&dU, nn * sizeof(c64)));
        YQ_HIP(hipMalloc(&dZ, (nn + 2 * n + 4) * sizeof(double)));
        YQ_HIP(hipMalloc(&dE, (n + 2) * sizeof(double)));
        YQ_HIP(hipMalloc(&dperm, (n + 2) * sizeof(int)));
        YQ_HIP(hipMalloc(&dkeep, sizeof(int)));
        if (n > 128) YQ_HIP(hipMalloc(&dP, (int64_t)n * (n + 1) / 2 * sizeof(double2)));
        YQ_HIP(hipMemcpy(dG, G, nn * sizeof(c64), hipMemcpyHostToDevice));
        YQ_HIP(hipMemset(dU, 0, nn * sizeof(c64)));
        heev(dG, nn, dU, nn, dZ, nn + 2 * n + 4, dP, (int64_t)n * (n + 1) / 2, n, n, dE, dperm, n + 2, dkeep, rule, thr,
             max_bond, nullptr, 1, 0);
        YQ_HIP(hipDeviceSynchronize());
        YQ_HIP(hipMemcpy(evals, dE, n * sizeof(double), hipMemcpyDeviceToHost));
        YQ_HIP(hipMemcpy(U, dU, nn * sizeof(c64), hipMemcpyDeviceToHost));
        YQ_HIP(hipMemcpy(keep_out, dkeep, sizeof(int), hipMemcpyDeviceToHost));
        (void)hipFree(dG);
        (void)hipFree(dU);
        (void)hipFree(dZ);
        (void)hipFree(dE);
        (void)hipFree(dperm);
        (void)hipFree(dkeep);
        if (dP) (void)hipFree(dP);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}
