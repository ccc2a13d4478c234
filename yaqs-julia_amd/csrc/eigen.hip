// Batched Hermitian eigensolver for the FP64 Gram matrices of the bond
// splits (the SVD of split_mps_tensor_svd, src/Algorithms.jl:409-445).
//
// Fast path (n <= 128, i.e. bond dimensions up to 64), four kernels:
//   1. Householder reduction to a real symmetric tridiagonal T with LAPACK
//      zhetd2/zlarfg conventions (Q = H_0 ... H_{n-2},
//      H_k = I - tau_k v_k v_k^H, subdiagonal real), one workgroup per
//      matrix: hetrd_reg_kernel (n > 32) keeps the whole matrix in the
//      registers of 1024 threads, two barriers per column; hetrd_kernel
//      (n <= 32) keeps the packed lower triangle in LDS, 512 threads, three
//      barriers per column (Hermitian product, p^H v, rank-2 update with the
//      next column's extraction).
//   2. stebz_kernel: every eigenvalue of T by bisection on Sturm counts, one
//      thread per eigenvalue (no serial chain across eigenvalues), to an
//      absolute width of 2 eps ||T||; sorted output and the kept rank.
//   3. stein_kernel: the kept eigenvectors of T by inverse iteration, one
//      lane per eigenvector (clusters closer than 1e-9 ||T|| handled in order
//      by one lane with re-orthogonalisation), then modified Gram-Schmidt of
//      the kept columns in LDS.
//   4. ormtr_kernel: u = H_0 ... H_{n-2} x for the kept columns, with the
//      n x 64 complex block held in registers (thread = (row residue, column))
//      and one barrier per reflector.
// Wide path (128 < n <= 512, at most 256 kept columns): the same four steps
// with the matrix in global memory (hetrd_wide_kernel, full storage, one
// 1024-thread workgroup per matrix), eigenvalues by the same bisection,
// inverse iteration with the vectors in global memory (stein_wide_kernel),
// the Gram-Schmidt pass as its own kernel (cgs_wide_kernel) and the
// back-transform with 16 row residues per 64 columns (n <= 256) or 32 per
// 32 columns.
// Matrices whose inverse iteration breaks down (zero or non-finite vector),
// and every matrix outside both paths, go through heev_lds_kernel: one
// workgroup per matrix doing Householder + implicit QL + back-transform
// (slower, same conventions).
// Output: U[:, r] for r < keep (eigenvalues sorted descending, row stride
// n; the single-workgroup solver also fills the columns up to
// min(n, max_bond)), evals (sorted descending), perm[r] = r, keep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "kernels.h"
#include "reduce.h"

namespace yq {

namespace {

__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_mulc(double2 a, double2 b) {  // conj(a) b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }

__host__ __device__ __forceinline__ int pidx(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

__device__ __forceinline__ double wsum(double v) { return wave_sum(v); }

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
constexpr double EPS = 2.220446049250313e-16;

// Kept rank from eigenvalues sorted descending (ev[0..ns)).
__device__ int truncation_rank(const double *ev, int ns, int rule, double thr, int max_bond) {
    int kp = ns;
    if (rule == 0) {  // T-A (src/Algorithms.jl:422-438)
        double disc = 0;
        for (int k = ns; k >= 1; --k) {
            disc += fmax(ev[k - 1], 0.0);
            if (disc >= thr) {
                kp = k > 2 ? k : 2;
                break;
            }
        }
    } else if (rule == 1) {  // T-C (src/Dissipation.jl:123-136)
        double tot = 0;
        for (int k = 0; k < ns; ++k) tot += fmax(ev[k], 0.0);
        double cur = 0;
        for (int k = ns; k >= 1; --k) {
            cur += fmax(ev[k - 1], 0.0);
            if (cur > thr * tot) {
                kp = k;
                break;
            }
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
    return kp;
}

// Per-matrix work area of the fast path (byte offsets).
struct HeevLay {
    int64_t V, tau, d, e, lu, X, total;
};
__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }
__host__ __device__ inline HeevLay heev_lay(int n) {
    HeevLay l;
    int64_t o = 0;
    l.V = o;
    o = align256(o + (int64_t)n * n * 16);
    l.tau = o;
    o = align256(o + (int64_t)n * 16);
    l.d = o;
    o = align256(o + (int64_t)n * 8);
    l.e = o;
    o = align256(o + (int64_t)n * 8);
    l.lu = o;
    o = align256(o + (int64_t)4 * n * 64 * 8);
    l.X = o;
    o = align256(o + (int64_t)n * 64 * 8);
    l.total = o;
    return l;
}

// Per-matrix work area of the wide path (128 < n <= 512, kw = kept columns
// rounded up to 64, <= 256): A is the full-storage working copy of the
// matrix (row k right of the diagonal receives v_k), LU / X / X2 are laid
// out [row][kw].
constexpr int WIDE_MAX_N = 512, WIDE_MAX_K = 256;
struct HeevLayW {
    int64_t A, tau, d, e, lu, X, X2, total;
};
__host__ __device__ inline HeevLayW heev_lay_wide(int n, int kw) {
    HeevLayW l;
    int64_t o = 0;
    l.A = o;
    o = align256(o + (int64_t)n * n * 16);
    l.tau = o;
    o = align256(o + (int64_t)n * 16);
    l.d = o;
    o = align256(o + (int64_t)n * 8);
    l.e = o;
    o = align256(o + (int64_t)n * 8);
    l.lu = o;
    o = align256(o + (int64_t)4 * n * kw * 8);
    l.X = o;
    o = align256(o + (int64_t)n * kw * 8);
    l.X2 = o;
    o = align256(o + (int64_t)n * kw * 8);
    l.total = o;
    return l;
}
inline int round64(int k) { return (k + 63) / 64 * 64; }

}  // namespace

// ------------------------------------------------------------------ fallback
// P: packed lower triangle (LDS or global), n*(n+1)/2 double2.
__global__ __launch_bounds__(256) void heev_lds_kernel(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Zs,
                                                       int64_t sZ, double2 *Pg, int64_t sP, int n, int ns,
                                                       double *evals, int *perm, int64_t sE, int *keep, int rule,
                                                       double thr, int max_bond, const int *mask, const int *only) {
    extern __shared__ double2 smem2[];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (only && !only[t]) return;
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
    __shared__ int nrot_sh;
    __shared__ double abstol_sh;
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
        double tn = 0;
        for (int i = 0; i < n; ++i) tn = fmax(tn, fabs(dd[i]) + fabs(ee[i]));
        abstol_sh = EPS * tn;
    }
    // Z = I
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) Z[e] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    const double abstol = abstol_sh;
    // ---------------------------------------------- implicit QL
    for (int l = 0; l < n; ++l) {
        for (int iter = 0;; ++iter) {
            // thread 0 finds the split point and, if needed, computes one QL
            // sweep on (l..m), recording its rotations
            if (threadIdx.x == 0) {
                int m = l;
                for (; m < n - 1; ++m) {
                    const double sc = fabs(dd[m]) + fabs(dd[m + 1]);
                    if (fabs(ee[m]) <= EPS * sc || fabs(ee[m]) <= abstol) break;
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
    if (threadIdx.x == 0) keep[t] = truncation_rank(ev, ns, rule, thr, max_bond);
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

// ------------------------------------------------------------------ fast path
// 1. Householder tridiagonalisation, one workgroup per matrix (n <= 128).
constexpr int TRD_WAVES = 8;  // 512 threads: two waves per SIMD
constexpr int TRD_PARTS = TRD_WAVES * 64 / 128;  // column parts of the Hermitian product
__global__ __launch_bounds__(TRD_WAVES * 64) void hetrd_kernel(const c64 *__restrict__ G, int64_t sG,
                                                               char *__restrict__ Wk, int64_t sW, int n,
                                                               const int *__restrict__ mask) {
    extern __shared__ double2 smem2[];
    constexpr int NT = TRD_WAVES * 64;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int np = n * (n + 1) / 2;
    double2 *P = smem2;
    double2 *xb = P + np;                // [2][n] current / next column below the diagonal
    double2 *pb = xb + 2 * n;            // [TRD_PARTS][n] partial Hermitian products
    double2 *pf = pb + TRD_PARTS * n;    // [n] p
    double2 *vf = pf + n;                // [n] v
    __shared__ double sxs[2][TRD_WAVES];
    __shared__ double2 spv[TRD_WAVES];
    const HeevLay lay = heev_lay(n);
    char *base = Wk + t * sW;
    double2 *Vg = (double2 *)(base + lay.V);
    double2 *taug = (double2 *)(base + lay.tau);
    double *dg = (double *)(base + lay.d);
    double *eg = (double *)(base + lay.e);
    const c64 *A = G + t * sG;
    // packed lower triangle, one wavefront per row
    for (int i = w; i < n; i += TRD_WAVES) {
        const int rb = pidx(i, 0);
        for (int j = lane; j <= i; j += 64) P[rb + j] = make_double2(A[(int64_t)i * n + j].re, A[(int64_t)i * n + j].im);
    }
    // first column below the diagonal and its tail norm
    double xsp = 0;
    for (int i = tid; i + 1 < n; i += NT) {
        const double2 x = make_double2(A[(int64_t)(i + 1) * n].re, A[(int64_t)(i + 1) * n].im);
        xb[i] = x;
        if (i >= 1) xsp += x.x * x.x + x.y * x.y;
    }
    xsp = wsum(xsp);
    if (lane == 0) sxs[0][w] = xsp;
    __syncthreads();
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1, cur = k & 1;
        const double2 *x = xb + cur * n;
        double xs = 0;
#pragma unroll
        for (int i = 0; i < TRD_WAVES; ++i) xs += sxs[cur][i];
        const double2 alpha = x[0];
        double2 tk = make_double2(0, 0), inv = make_double2(0, 0);
        double beta = alpha.x;
        if (!(xs == 0.0 && alpha.y == 0.0)) {
            const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs);
            beta = alpha.x >= 0 ? -nr : nr;
            tk = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double dn = den.x * den.x + den.y * den.y;
            inv = make_double2(den.x / dn, -den.y / dn);
        }
        // y = (alpha - beta, x_1, x_2, ...), so that v = inv * y
        const double2 y0 = make_double2(alpha.x - beta, alpha.y);
        if (tid == 0) {
            dg[k] = P[pidx(k, k)].x;
            eg[k] = beta;
            taug[k] = tk;
        }
        // Hermitian product q = A22 y: rows r = tid % 128, column parts h = tid / 128
        {
            const int h = tid >> 7;
            const int part = (m + TRD_PARTS - 1) / TRD_PARTS;
            const int c0 = min(m, h * part), c1 = min(m, c0 + part);
            for (int r = tid & 127; r < m; r += 128) {
                double2 acc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[u] = make_double2(0, 0);
                const int rowb = pidx(k + 1 + r, k + 1);  // A22(r, c) = P[rowb + c] for c <= r
                int colb = pidx(k + 1 + c0, k + 1 + r);    // A22(c, r) = P[colb] for c > r
                int c = c0;
                for (; c + 3 < c1; c += 4) {
                    double2 a[4], yv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const bool lo = c + u <= r;
                        a[u] = P[lo ? rowb + c + u : colb];
                        if (!lo) a[u].y = -a[u].y;
                        yv[u] = c + u == 0 ? y0 : x[c + u];
                        colb += k + 2 + c + u;  // next row of the packed triangle
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        acc[u].x += a[u].x * yv[u].x - a[u].y * yv[u].y;
                        acc[u].y += a[u].x * yv[u].y + a[u].y * yv[u].x;
                    }
                }
                for (; c < c1; ++c) {
                    const bool lo = c <= r;
                    double2 a = P[lo ? rowb + c : colb];
                    if (!lo) a.y = -a.y;
                    const double2 yc = c == 0 ? y0 : x[c];
                    acc[0].x += a.x * yc.x - a.y * yc.y;
                    acc[0].y += a.x * yc.y + a.y * yc.x;
                    colb += k + 2 + c;
                }
                pb[h * n + r] = make_double2(acc[0].x + acc[1].x + acc[2].x + acc[3].x,
                                             acc[0].y + acc[1].y + acc[2].y + acc[3].y);
            }
        }
        __syncthreads();
        // p = tau inv q, v = inv y; partial p^H v
        const double2 ti = c_mul(tk, inv);
        double2 pvp = make_double2(0, 0);
        for (int r = tid; r < m; r += NT) {
            double2 q = pb[r];
#pragma unroll
            for (int h = 1; h < TRD_PARTS; ++h) {
                q.x += pb[h * n + r].x;
                q.y += pb[h * n + r].y;
            }
            const double2 p = c_mul(ti, q);
            const double2 v = r == 0 ? make_double2(1.0, 0.0) : c_mul(inv, x[r]);
            pf[r] = p;
            vf[r] = v;
            Vg[(int64_t)k * n + k + 1 + r] = v;
            const double2 z = c_mulc(p, v);
            pvp.x += z.x;
            pvp.y += z.y;
        }
        pvp.x = wsum(pvp.x);
        pvp.y = wsum(pvp.y);
        if (lane == 0) spv[w] = pvp;
        __syncthreads();
        // A22 -= v w^H + w v^H with w = p + al v, al = -tau (p^H v) / 2
        double2 pv = make_double2(0, 0);
#pragma unroll
        for (int i = 0; i < TRD_WAVES; ++i) {
            pv.x += spv[i].x;
            pv.y += spv[i].y;
        }
        const double2 al = c_mul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
        double2 *xn = xb + (cur ^ 1) * n;
        double xsn = 0;
        // lane columns c0 = lane, c1 = lane + 64 (their v and w in registers);
        // four rows per wave and pass, all loads issued before the stores
        const int c0 = lane, c1 = lane + 64;
        double2 vc0 = make_double2(0, 0), wc0 = make_double2(0, 0), vc1 = make_double2(0, 0), wc1 = make_double2(0, 0);
        if (c0 < m) {
            vc0 = vf[c0];
            const double2 pc = pf[c0];
            wc0 = make_double2(pc.x + al.x * vc0.x - al.y * vc0.y, pc.y + al.x * vc0.y + al.y * vc0.x);
        }
        if (c1 < m) {
            vc1 = vf[c1];
            const double2 pc = pf[c1];
            wc1 = make_double2(pc.x + al.x * vc1.x - al.y * vc1.y, pc.y + al.x * vc1.y + al.y * vc1.x);
        }
        for (int r0 = w; r0 < m; r0 += 4 * TRD_WAVES) {
            double2 a[4][2], vr[4], wr[4];
            int rb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = r0 + u * TRD_WAVES;
                const bool ok = r < m;
                vr[u] = ok ? vf[r] : make_double2(0, 0);
                const double2 pr = ok ? pf[r] : make_double2(0, 0);
                wr[u] = make_double2(pr.x + al.x * vr[u].x - al.y * vr[u].y, pr.y + al.x * vr[u].y + al.y * vr[u].x);
                rb[u] = pidx(k + 1 + r, k + 1);
                a[u][0] = ok && c0 <= r ? P[rb[u] + c0] : make_double2(0, 0);
                a[u][1] = ok && c1 <= r ? P[rb[u] + c1] : make_double2(0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = r0 + u * TRD_WAVES;
                if (r >= m) continue;
                // a -= vr conj(wc) + wr conj(vc)
                if (c0 <= r) {
                    double2 x = a[u][0];
                    x.x -= vr[u].x * wc0.x + vr[u].y * wc0.y + wr[u].x * vc0.x + wr[u].y * vc0.y;
                    x.y -= vr[u].y * wc0.x - vr[u].x * wc0.y + wr[u].y * vc0.x - wr[u].x * vc0.y;
                    P[rb[u] + c0] = x;
                    if (c0 == 0 && r >= 1) {
                        xn[r - 1] = x;
                        if (r >= 2) xsn += x.x * x.x + x.y * x.y;
                    }
                }
                if (c1 <= r) {
                    double2 x = a[u][1];
                    x.x -= vr[u].x * wc1.x + vr[u].y * wc1.y + wr[u].x * vc1.x + wr[u].y * vc1.y;
                    x.y -= vr[u].y * wc1.x - vr[u].x * wc1.y + wr[u].y * vc1.x - wr[u].x * vc1.y;
                    P[rb[u] + c1] = x;
                }
            }
        }
        if (lane == 0) sxs[cur ^ 1][w] = xsn;  // only lane 0 (column 0) contributes
        __syncthreads();
    }
    if (tid == 0) {
        dg[n - 1] = P[pidx(n - 1, n - 1)].x;
        eg[n - 1] = 0.0;
    }
}

// 1b. The same reduction with the whole matrix in registers (n > 32): 1024
// threads, thread (r = tid / 8, cc = tid % 8) holds row r at columns
// cc + 8b, b < 16 (complex FP64, full Hermitian storage), so a column step
// reads no matrix element from memory.  Column k (k1 = k + 1),
// two barriers:
//   A: q = A22 y (y = x except y_k1 = alpha - beta) as A22 x minus
//      beta A[:, k1]: each thread its columns, the 8 lanes of a row summed
//      by DPP; p = tau inv q, v = inv y, per-wave partial of p^H v;
//   B: A22 -= v w^H + w v^H (w = p + al v) folded per row into
//      a_rc -= s1 conj(p_c) + s2 conj(y_c), s1 = v_r,
//      s2 = v_r conj(al inv) + w_r conj(inv).  Then row k1 writes the next
//      column (x_c = A[c][k1] = conj(A[k1][c]), c > k1), its tail norm and
//      d[k1], and row k1 + 1 writes A[c][k1 + 1] for the next correction.
// Every LDS vector is zero outside the trailing block, so partially covered
// column groups need no masks; waves whose rows are all reduced (or past n)
// and groups of 32 columns (4 blocks, their LDS reads issued together)
// outside the trailing block are skipped by uniform branches, so the
// per-column work shrinks with the trailing block.  The kernel is bound by
// the latency of its per-column chain (LDS round trips, DPP reductions, the
// scalar Householder arithmetic, two barriers); a variant with 4 lanes per
// row (512 threads, 32 values each) halves the per-wave scalar work but
// hides less latency and measured slower.  Conventions and outputs as
// hetrd_kernel.
constexpr int TRR_THREADS = 1024;
__device__ __forceinline__ double sum_row(double v) {  // over the 8 lanes of a row
    v += dpp_move<0xb1>(v);                             // xor 1
    v += dpp_move<0x4e>(v);                             // xor 2
    v += dpp_move<0x141>(v);                            // half-row mirror: the other quad
    return v;
}
// sum over lanes 0..15 (the other lanes' values are ignored), uniform
__device__ __forceinline__ double sum16(double v) {
    v += dpp_move<0xb1>(v);   // xor 1
    v += dpp_move<0x4e>(v);   // xor 2
    v += dpp_move<0x141>(v);  // half-row mirror
    v += dpp_move<0x140>(v);  // row mirror
    return lane_read(v, 0);
}
__global__ __launch_bounds__(TRR_THREADS) void hetrd_reg_kernel(const c64 *__restrict__ G, int64_t sG,
                                                                char *__restrict__ Wk, int64_t sW, int n,
                                                                const int *__restrict__ mask) {
    constexpr int LPR = 8, NB = 128 / LPR, RPW = 64 / LPR, NWV = TRR_THREADS / 64, GC = 4 * LPR;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    __shared__ double2 xb[2][128];  // x of column k by parity: A[c][k] for c > k, else 0
    __shared__ double2 cb[2][128];  // A[c][k1] for c >= k1, else 0
    __shared__ double2 xy[128];     // y for the update
    __shared__ double2 pb[128];     // p, 0 outside the trailing block
    __shared__ double sxs[2];       // tail norm of x
    __shared__ double2 rpv[NWV];
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = tid / LPR, cc = tid % LPR;
    const HeevLay lay = heev_lay(n);
    char *base = Wk + t * sW;
    double2 *Vg = (double2 *)(base + lay.V);
    double2 *taug = (double2 *)(base + lay.tau);
    double *dg = (double *)(base + lay.d);
    double *eg = (double *)(base + lay.e);
    const c64 *A = G + t * sG;
    const double2 zero = make_double2(0.0, 0.0);
    double2 a[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int c = cc + LPR * b;
        a[b] = zero;
        if (r < n && c < n) a[b] = make_double2(A[(int64_t)r * n + c].re, A[(int64_t)r * n + c].im);
    }
    if (tid < 128) {
        xb[1][tid] = zero;
        cb[1][tid] = zero;
        xy[tid] = zero;
        pb[tid] = zero;
    }
    // column 0 from row 0, the first correction from row 1
    if (r == 0) {
        double xsp = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int c = cc + LPR * b;
            xb[0][c] = c > 0 && c < n ? c_conj(a[b]) : zero;
            if (c >= 2) xsp += a[b].x * a[b].x + a[b].y * a[b].y;
        }
        if (cc == 0) dg[0] = a[0].x;
        xsp = sum_row(xsp);
        if (cc == 0) sxs[0] = xsp;
    } else if (r == 1) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int c = cc + LPR * b;
            cb[0][c] = c >= 1 && c < n ? c_conj(a[b]) : zero;
        }
    }
    __syncthreads();
    for (int k = 0; k + 1 < n; ++k) {
        const int par = k & 1, k1 = k + 1;
        const bool wact = RPW * w + RPW - 1 >= k1 && RPW * w < n;
        double2 vr = zero, pr = zero, tk = zero, inv = zero;
        double beta = 0.0;
        if (wact) {
            const double2 alpha = xb[par][k1];
            const double xs = sxs[par];
            const double2 ck = cb[par][r];
            const double2 xr = xb[par][r];
            beta = alpha.x;
            if (!(xs == 0.0 && alpha.y == 0.0)) {
                const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs);
                beta = alpha.x >= 0 ? -nr : nr;
                const double rb = 1.0 / beta;
                tk = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
                const double2 den = make_double2(alpha.x - beta, alpha.y);
                const double rdn = 1.0 / (den.x * den.x + den.y * den.y);
                inv = make_double2(den.x * rdn, -den.y * rdn);
            }
            const double2 ti = c_mul(tk, inv);
            double2 acc = zero;
#pragma unroll
            for (int g = 0; g < NB / 4; ++g) {
                if (GC * g + GC - 1 >= k1 && GC * g < n) {
                    double2 xc[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) xc[u] = xb[par][cc + LPR * (4 * g + u)];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double2 av = a[4 * g + u];
                        acc.x += av.x * xc[u].x - av.y * xc[u].y;
                        acc.y += av.x * xc[u].y + av.y * xc[u].x;
                    }
                }
            }
            acc.x = sum_row(acc.x);
            acc.y = sum_row(acc.y);
            acc.x -= beta * ck.x;
            acc.y -= beta * ck.y;
            const bool rin = r >= k1 && r < n;
            if (rin) {
                pr = c_mul(ti, acc);
                vr = r == k1 ? make_double2(1.0, 0.0) : c_mul(inv, xr);
            }
            if (cc == 0) {
                pb[r] = pr;
                if (rin) Vg[(int64_t)k * n + r] = vr;
                if (r == k1) {
                    eg[k] = beta;
                    taug[k] = tk;
                }
            } else if (cc == 1) {
                xy[r] = rin ? (r == k1 ? make_double2(xr.x - beta, xr.y) : xr) : zero;
            }
            // per-wave partial of p^H v: one lane per row carries the row's term
            const double2 z = cc == 0 ? c_mulc(pr, vr) : zero;
            const double zx = wsum(z.x), zy = wsum(z.y);
            if (lane == 0) rpv[w] = make_double2(zx, zy);
        }
        __syncthreads();
        if (wact) {
            // the active waves' partials, one per lane, summed across lanes
            double2 q = zero;
            if (lane < NWV && RPW * lane + RPW - 1 >= k1 && RPW * lane < n) q = rpv[lane];
            double2 pv;
            pv.x = sum16(q.x);
            pv.y = sum16(q.y);
            const double2 al = c_mul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
            const double2 wr = make_double2(pr.x + al.x * vr.x - al.y * vr.y, pr.y + al.x * vr.y + al.y * vr.x);
            const double2 s1 = vr;
            const double2 ai = c_mul(al, inv);
            const double2 s2 = make_double2(vr.x * ai.x + vr.y * ai.y + wr.x * inv.x + wr.y * inv.y,
                                            vr.y * ai.x - vr.x * ai.y + wr.y * inv.x - wr.x * inv.y);
#pragma unroll
            for (int g = 0; g < NB / 4; ++g) {
                if (GC * g + GC - 1 >= k1 && GC * g < n) {
                    double2 yc[4], pc[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        yc[u] = xy[cc + LPR * (4 * g + u)];
                        pc[u] = pb[cc + LPR * (4 * g + u)];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        double2 &av = a[4 * g + u];
                        av.x -= s1.x * pc[u].x + s1.y * pc[u].y + s2.x * yc[u].x + s2.y * yc[u].y;
                        av.y -= s1.y * pc[u].x - s1.x * pc[u].y + s2.y * yc[u].x - s2.x * yc[u].y;
                    }
                }
            }
            if (r == k1) {
                double xsp = 0, dk = 0;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int c = cc + LPR * b;
                    xb[par ^ 1][c] = c > k1 && c < n ? c_conj(a[b]) : zero;
                    if (c > k1 + 1) xsp += a[b].x * a[b].x + a[b].y * a[b].y;
                    dk = c == k1 ? a[b].x : dk;
                }
                if (cc == k1 % LPR) dg[k1] = dk;
                xsp = sum_row(xsp);
                if (cc == 0) sxs[par ^ 1] = xsp;
            } else if (r == k1 + 1) {
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int c = cc + LPR * b;
                    cb[par ^ 1][c] = c >= k1 + 1 && c < n ? c_conj(a[b]) : zero;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) eg[n - 1] = 0.0;
}

// 2. Eigenvalues of T by bisection on Sturm counts, one thread per
// eigenvalue (ascending index k = thread), to an absolute width of
// 2 eps ||T||; then sorted output, identity perm and the kept rank.
__device__ __forceinline__ double fast_rcp(double q) {
    double r = __builtin_amdgcn_rcp(q);  // hardware estimate, refined twice
    r = r * fma(-q, r, 2.0);
    return r * fma(-q, r, 2.0);
}

// NMAX: largest n of the launch (blockDim >= n, a multiple of 64); the
// tridiagonal (d, e) sits at byte offsets od / oe of each work area.
template <int NMAX>
__global__ __launch_bounds__(NMAX) void stebz_kernel(char *__restrict__ Wk, int64_t sW, int64_t od, int64_t oe, int n,
                                                     int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
                                                     double thr, int max_bond, const int *__restrict__ mask) {
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int k = threadIdx.x;
    __shared__ double dd[NMAX], e2[NMAX], red[3][NMAX / 64];
    const char *base = Wk + t * sW;
    const double *dg = (const double *)(base + od);
    const double *eg = (const double *)(base + oe);
    double lo = 0, hi = 0, ee = 0;
    if (k < n) {
        const double di = dg[k], ei = k + 1 < n ? fabs(eg[k]) : 0.0, ep = k > 0 ? fabs(eg[k - 1]) : 0.0;
        dd[k] = di;
        e2[k] = ei * ei;
        lo = di - ei - ep;
        hi = di + ei + ep;
        ee = ei * ei;
    } else {
        lo = 1e300;
        hi = -1e300;
    }
    // Gershgorin interval and max e^2 over the block (two wavefronts)
    lo = wave_min(lo);
    hi = wave_max(hi);
    ee = wave_max(ee);
    if ((k & 63) == 0) {
        red[0][k >> 6] = lo;
        red[1][k >> 6] = hi;
        red[2][k >> 6] = ee;
    }
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    lo = red[0][0];
    hi = red[1][0];
    ee = red[2][0];
    for (int i = 1; i < nw; ++i) {
        lo = fmin(lo, red[0][i]);
        hi = fmax(hi, red[1][i]);
        ee = fmax(ee, red[2][i]);
    }
    const double tn = fmax(fabs(lo), fabs(hi));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, ee);
    const double width = 2.0 * EPS * fmax(tn, 1e-300);
    lo -= width;
    hi += width;
    if (k < n) {
        // invariant: count(lo) <= k < count(hi); the width is reached in
        // about 53 halvings, the cap only guards against a stuck interval
        for (int it = 0; it < 128 && hi - lo > width; ++it) {
            const double x = 0.5 * (lo + hi);
            int cnt = 0;
            double q = dd[0] - x;
            if (fabs(q) <= pivmin) q = -pivmin;
            cnt += q < 0.0;
            for (int i = 1; i < n; ++i) {
                q = (dd[i] - x) - e2[i - 1] * fast_rcp(q);
                if (fabs(q) <= pivmin) q = -pivmin;
                cnt += q < 0.0;
            }
            if (cnt <= k) lo = x;
            else hi = x;
        }
        evals[t * sE + (n - 1 - k)] = 0.5 * (lo + hi);
        perm[t * sE + k] = k;
    }
    __syncthreads();
    if (k == 0) keep[t] = truncation_rank(evals + t * sE, ns, rule, thr, max_bond);
}

// 3. Kept eigenvectors of T by inverse iteration, one wavefront per matrix,
// lane c = the c-th largest eigenvalue, c < keep.  T - lambda I is factored
// with partial pivoting (U with two superdiagonals, stored as reciprocal
// diagonal and superdiagonals, plus the L multipliers; arrays in global
// memory laid out [i][lane]); pivots below eps ||T|| are replaced by
// +-eps ||T||; three solves from a fixed pseudo-random start.  Eigenvalues
// closer than 1e-9 ||T|| form a cluster whose vectors the first lane of the
// cluster computes in order, orthogonalising each against the previous ones
// after every solve.  Inverse iteration leaves vectors of eigenvalues g
// apart orthogonal to about eps ||T|| / g; one final classical Gram-Schmidt
// pass (largest eigenvalue first) makes them orthonormal.
__global__ __launch_bounds__(64) void stein_kernel(char *__restrict__ Wk, int64_t sW, int n, int kout,
                                                   const double *__restrict__ evals, int64_t sE,
                                                   const int *__restrict__ keep, const int *__restrict__ mask,
                                                   int *__restrict__ fallback) {
    extern __shared__ double lds_x[];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int c = threadIdx.x;
    const int kc = min(keep[t], kout);
    const HeevLay lay = heev_lay(n);
    char *base = Wk + t * sW;
    const double *dg = (const double *)(base + lay.d);
    const double *eg = (const double *)(base + lay.e);
    double *ui = (double *)(base + lay.lu);  // [i][64]: 1 / U diagonal
    double *ub = ui + n * 64;                // U first superdiagonal
    double *uc = ub + n * 64;                // U second superdiagonal
    double *ml = uc + n * 64;                // L multipliers
    double *X = lds_x;                       // X[row * 64 + col]
    double *dd = lds_x + n * 64, *es = dd + n;
    __shared__ int bad;
    for (int i = c; i < n; i += 64) {
        dd[i] = dg[i];
        es[i] = i + 1 < n ? eg[i] : 0.0;
    }
    if (c == 0) bad = 0;
    __syncthreads();
    const double *ev = evals + t * sE;
    double tn = 0;
    for (int i = 0; i < n; ++i) tn = fmax(tn, fabs(dd[i]) + fabs(es[i]) + (i ? fabs(es[i - 1]) : 0.0));
    const double tiny = EPS * fmax(tn, 1e-300);
    const double ctol = 1e-9 * tn;
    // cluster of lane c: [c, c1) when lane c leads it
    const bool lead = c < kc && (c == 0 || ev[c - 1] - ev[c] > ctol);
    int c1 = c + 1;
    if (lead)
        while (c1 < kc && ev[c1 - 1] - ev[c1] <= ctol) ++c1;
    if (lead) {
        for (int j = c; j < c1; ++j) {
            const double lam = ev[j];
            // factor T - lam I = P L U
            uint64_t piv[2] = {0, 0};
            double a_cur = dd[0] - lam, b_cur = n > 1 ? es[0] : 0.0;  // current row: diagonal, superdiagonal
            auto put_diag = [&](int i, double u) {
                if (fabs(u) < tiny) u = u < 0.0 ? -tiny : tiny;
                ui[i * 64 + c] = 1.0 / u;
            };
            for (int i = 0; i + 1 < n; ++i) {
                const double ci = es[i];  // subdiagonal below the current row
                const double a1 = dd[i + 1] - lam;
                const double b1 = i + 2 < n ? es[i + 1] : 0.0;
                if (fabs(a_cur) >= fabs(ci)) {
                    const double l = a_cur != 0.0 ? ci / a_cur : 0.0;
                    put_diag(i, a_cur);
                    ub[i * 64 + c] = b_cur;
                    uc[i * 64 + c] = 0.0;
                    ml[i * 64 + c] = l;
                    a_cur = a1 - l * b_cur;
                    b_cur = b1;
                } else {  // swap rows i and i+1
                    const double l = a_cur / ci;
                    put_diag(i, ci);
                    ub[i * 64 + c] = a1;
                    uc[i * 64 + c] = b1;
                    ml[i * 64 + c] = l;
                    piv[i >> 6] |= 1ull << (i & 63);
                    a_cur = b_cur - l * a1;
                    b_cur = -l * b1;
                }
            }
            put_diag(n - 1, a_cur);
            double *x = X + j;  // column j, stride 64
            for (int i = 0; i < n; ++i) {
                const uint32_t h = (uint32_t)(i * 2654435761u) ^ (uint32_t)((j + 1) * 2246822519u);
                x[i * 64] = 0.5 + (double)((h >> 8) & 0xffff) * (1.0 / 65536.0);
            }
            for (int it = 0; it < 3; ++it) {
                // forward: apply P and L (r0 carries row i)
                double r0 = x[0];
#pragma unroll 8
                for (int i = 0; i + 1 < n; ++i) {
                    const double l = ml[i * 64 + c];
                    double r1 = x[(i + 1) * 64];
                    if ((piv[i >> 6] >> (i & 63)) & 1ull) {
                        const double tmp = r0;
                        r0 = r1;
                        r1 = tmp;
                    }
                    x[i * 64] = r0;
                    r0 = r1 - l * r0;
                }
                x[(n - 1) * 64] = r0;
                // back substitution with U
                double y2 = 0.0, y1 = 0.0;
#pragma unroll 8
                for (int i = n - 1; i >= 0; --i) {
                    const double y = (x[i * 64] - ub[i * 64 + c] * y1 - uc[i * 64 + c] * y2) * ui[i * 64 + c];
                    x[i * 64] = y;
                    y2 = y1;
                    y1 = y;
                }
                // against the earlier vectors of the cluster
                for (int p = c; p < j; ++p) {
                    double s = 0;
                    for (int i = 0; i < n; ++i) s += X[i * 64 + p] * x[i * 64];
                    for (int i = 0; i < n; ++i) x[i * 64] -= s * X[i * 64 + p];
                }
                double nr = 0;
                for (int i = 0; i < n; ++i) nr += x[i * 64] * x[i * 64];
                if (!(nr > 0.0) || !isfinite(nr)) {
                    bad = 1;
                    break;
                }
                const double inv = 1.0 / sqrt(nr);
                for (int i = 0; i < n; ++i) x[i * 64] *= inv;
            }
        }
    }
    if (c >= kc)
        for (int i = 0; i < n; ++i) X[i * 64 + c] = 0.0;
    __syncthreads();
    // One classical Gram-Schmidt pass against all earlier (larger-eigenvalue)
    // columns, every projection computed from the inputs: with unit columns
    // whose pairwise products are at most ~eps ||T|| / gap, the result is
    // orthogonal to second order in those products.
    // (branch-free: all 64 products are formed, those with p >= c dropped)
    double sp[64];
#pragma unroll
    for (int p = 0; p < 64; ++p) sp[p] = 0.0;
    for (int i = 0; i < n; ++i) {
        const double xc = X[i * 64 + c];
#pragma unroll
        for (int p = 0; p < 64; ++p) sp[p] += X[i * 64 + p] * xc;
    }
#pragma unroll
    for (int p = 0; p < 64; ++p) sp[p] = p < c ? sp[p] : 0.0;
    // in place, row by row: the 64 lanes are one wavefront, so every lane
    // reads row i before any lane writes it
    double nr = 0;
    for (int i = 0; i < n; ++i) {
        double y0 = X[i * 64 + c], y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll
        for (int p = 0; p < 64; p += 4) {
            y0 -= sp[p] * X[i * 64 + p];
            y1 -= sp[p + 1] * X[i * 64 + p + 1];
            y2 -= sp[p + 2] * X[i * 64 + p + 2];
            y3 -= sp[p + 3] * X[i * 64 + p + 3];
        }
        const double y = (y0 + y1) + (y2 + y3);
        X[i * 64 + c] = y;
        nr += y * y;
    }
    const double inv = (c < kc && nr > 0.0) ? 1.0 / sqrt(nr) : 0.0;
    double *Xg = (double *)(base + lay.X);
    for (int i = 0; i < n; ++i) Xg[i * 64 + c] = X[i * 64 + c] * inv;
    if (c == 0) fallback[t] = bad;
}

// 4. u = H_0 ... H_{n-2} x for the kept columns: CW columns per workgroup
// (blockIdx.y selects them) and NR = NWV * 64 / CW row residues, each
// thread holding rows rres + NR j of its column in registers (n <= NMAX).
// v_k is row k of the matrix at byte offset ov (entries k+1 .. n-1), x is
// laid out [row][ldx] at offset ox.
template <int NMAX, int NWV, int CW>
__global__ __launch_bounds__(NWV * 64) void ormtr_kernel(char *__restrict__ Wk, int64_t sW, int64_t ov, int64_t otau,
                                                         int64_t ox, int ldx, int n, int kout, c64 *__restrict__ U,
                                                         int64_t sU, const int *__restrict__ mask,
                                                         const int *__restrict__ fallback) {
    constexpr int NR = NWV * 64 / CW, NJ = NMAX / NR;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (fallback[t]) return;
    const int tid = threadIdx.x, cl = tid % CW;
    const int rres = CW == 64 ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid / CW;  // wave-uniform for CW = 64
    const int c = blockIdx.y * CW + cl;
    __shared__ double2 vb[3][NMAX];
    __shared__ double2 part[2][NR][CW];
    __shared__ double2 taus[NMAX];
    const char *base = Wk + t * sW;
    const double2 *__restrict__ Vg = (const double2 *)(base + ov);
    const double2 *__restrict__ taug = (const double2 *)(base + otau);
    const double *__restrict__ Xg = (const double *)(base + ox);
    double2 u[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int r = rres + NR * j;
        u[j] = make_double2(r < n && c < ldx ? Xg[(int64_t)r * ldx + c] : 0.0, 0.0);
    }
    // Every global load of the reflector loop is issued one step before its
    // value is needed: the scalars tau in LDS up front, row k - 1 of V held
    // in a register while step k + 1 runs (a load consumed in the step that
    // issues it would stall every wave at the step's barrier).
    for (int i = tid; i + 1 < n; i += blockDim.x) taus[i] = taug[i];
    if (n >= 2 && tid < n) vb[(n - 2) % 3][tid] = tid >= n - 1 ? Vg[(int64_t)(n - 2) * n + tid] : make_double2(0, 0);
    double2 vnext = make_double2(0, 0);  // row k - 1 at the top of step k
    if (n >= 3 && tid < n && tid >= n - 2) vnext = Vg[(int64_t)(n - 3) * n + tid];
    __syncthreads();
    for (int k = n - 2; k >= 0; --k) {
        const int cb = k % 3;
        if (k >= 1 && tid < n) vb[(k - 1) % 3][tid] = vnext;
        if (k >= 2 && tid < n) vnext = tid >= k - 1 ? Vg[(int64_t)(k - 2) * n + tid] : make_double2(0, 0);
        const double2 tk = taus[k];
        const double2 *v = vb[cb];
        double2 acc = make_double2(0, 0);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int r = rres + NR * j;
            if (r > k && r < n) {
                const double2 z = c_mulc(v[r], u[j]);
                acc.x += z.x;
                acc.y += z.y;
            }
        }
        part[k & 1][rres][cl] = acc;
        __syncthreads();
        double2 s = make_double2(0, 0);
#pragma unroll
        for (int h = 0; h < NR; ++h) {
            s.x += part[k & 1][h][cl].x;
            s.y += part[k & 1][h][cl].y;
        }
        const double2 f = c_mul(tk, s);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int r = rres + NR * j;
            if (r > k && r < n) {
                const double2 z = c_mul(v[r], f);
                u[j].x -= z.x;
                u[j].y -= z.y;
            }
        }
    }
    if (c < kout) {
        c64 *Ut = U + t * sU;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int r = rres + NR * j;
            if (r < n) Ut[(int64_t)r * n + c] = cmake<double>(u[j].x, u[j].y);
        }
    }
}

// ------------------------------------------------------------- wide path
// 1w. Householder tridiagonalisation for 128 < n <= 512: one 1024-thread
// workgroup per matrix on a full-storage copy A in global memory (the whole
// Hermitian matrix is kept so that every access is a contiguous row: the
// column below the diagonal is read as the conjugate of row k, and one
// wavefront handles one row).  v_k goes to row k right of the diagonal,
// which the reduction no longer reads.  Same conventions as hetrd_kernel.
// The pass is HBM-bound, so the rank-2 update of step k and the Hermitian
// product of step k+1 share one sweep over the trailing matrix: row k+1 is
// updated first, v_{k+1} is formed from it, then every further row is
// updated and immediately multiplied by v_{k+1} (one read and one write of
// the trailing matrix per column).
constexpr int WT_WAVES = 16;
__global__ __launch_bounds__(WT_WAVES * 64) void hetrd_wide_kernel(const c64 *__restrict__ G, int64_t sG,
                                                                   char *__restrict__ Wk, int64_t sW, int n, int kw,
                                                                   const int *__restrict__ mask) {
    constexpr int NT = WT_WAVES * 64;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ double2 vbuf[2][WIDE_MAX_N], pbuf[2][WIDE_MAX_N];
    __shared__ double rs[WT_WAVES];
    __shared__ double2 rs2[WT_WAVES];
    const HeevLayW lay = heev_lay_wide(n, kw);
    char *base = Wk + t * sW;
    double2 *Am = (double2 *)(base + lay.A);
    double2 *taug = (double2 *)(base + lay.tau);
    double *dg = (double *)(base + lay.d);
    double *eg = (double *)(base + lay.e);
    const c64 *Gt = G + t * sG;
    for (int64_t e = tid; e < (int64_t)n * n; e += NT) Am[e] = make_double2(Gt[e].re, Gt[e].im);
    __syncthreads();
    // Householder vector of column k from row k of the current matrix (its
    // conjugate is the column below the diagonal): v into vs, v_k into row k,
    // d_k, e_k, tau_k; returns tau_k.  Two barriers.
    auto reflector = [&](int k, double2 *vs) -> double2 {
        const int m = n - k - 1;
        double2 *rowk = Am + (int64_t)k * n + k + 1;
        double xs = 0;
        for (int r = 1 + tid; r < m; r += NT) {
            const double2 a = rowk[r];
            xs += a.x * a.x + a.y * a.y;
        }
        xs = wsum(xs);
        if (lane == 0) rs[w] = xs;
        __syncthreads();
        xs = 0;
#pragma unroll
        for (int i = 0; i < WT_WAVES; ++i) xs += rs[i];
        const double2 alpha = c_conj(rowk[0]);
        double2 tk = make_double2(0, 0), inv = make_double2(0, 0);
        double beta = alpha.x;
        if (!(xs == 0.0 && alpha.y == 0.0)) {
            const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs);
            beta = alpha.x >= 0 ? -nr : nr;
            tk = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double dn = den.x * den.x + den.y * den.y;
            inv = make_double2(den.x / dn, -den.y / dn);
        }
        for (int r = tid; r < m; r += NT) vs[r] = r == 0 ? make_double2(1.0, 0.0) : c_mul(inv, c_conj(rowk[r]));
        if (tid == 0) {
            dg[k] = Am[(int64_t)k * n + k].x;
            eg[k] = beta;
            taug[k] = tk;
        }
        __syncthreads();
        for (int r = tid; r < m; r += NT) rowk[r] = vs[r];
        return tk;
    };
    // al = -tau (p^H v) / 2 over m entries.  One barrier.
    auto pv_alpha = [&](int m, const double2 *vs, const double2 *ps, double2 tk) -> double2 {
        double2 pvp = make_double2(0, 0);
        for (int r = tid; r < m; r += NT) {
            const double2 z = c_mulc(ps[r], vs[r]);
            pvp.x += z.x;
            pvp.y += z.y;
        }
        pvp.x = wsum(pvp.x);
        pvp.y = wsum(pvp.y);
        if (lane == 0) rs2[w] = pvp;
        __syncthreads();
        double2 pv = make_double2(0, 0);
#pragma unroll
        for (int i = 0; i < WT_WAVES; ++i) {
            pv.x += rs2[i].x;
            pv.y += rs2[i].y;
        }
        return c_mul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
    };
    // step 0: reflector and the Hermitian product p = tau A22 v from scratch
    int cur = 0;
    double2 tk = reflector(0, vbuf[0]);
    {
        const int m = n - 1;
        const double2 *vs = vbuf[0];
        for (int rr = w; rr < m; rr += WT_WAVES) {
            const double2 *arow = Am + (int64_t)(1 + rr) * n + 1;
            double2 acc = make_double2(0, 0);
            for (int c = lane; c < m; c += 64) {
                const double2 z = c_mul(arow[c], vs[c]);
                acc.x += z.x;
                acc.y += z.y;
            }
            acc.x = wsum(acc.x);
            acc.y = wsum(acc.y);
            if (lane == 0) pbuf[0][rr] = c_mul(tk, acc);
        }
        __syncthreads();
    }
    double2 al = pv_alpha(n - 1, vbuf[0], pbuf[0], tk);
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1;
        const double2 *vs = vbuf[cur], *ps = pbuf[cur];
        // A22 -= v w^H + w v^H with w = p + al v; row k+1 first
        auto update_row = [&](int rr, double2 *arow, double2 *vnext, double2 *acc) {
            const double2 vr = vs[rr], pr = ps[rr];
            const double2 wr = make_double2(pr.x + al.x * vr.x - al.y * vr.y, pr.y + al.x * vr.y + al.y * vr.x);
            for (int c = lane; c < m; c += 64) {
                const double2 vc = vs[c], pc = ps[c];
                const double2 wc = make_double2(pc.x + al.x * vc.x - al.y * vc.y, pc.y + al.x * vc.y + al.y * vc.x);
                double2 a = arow[c];
                a.x -= vr.x * wc.x + vr.y * wc.y + wr.x * vc.x + wr.y * vc.y;
                a.y -= vr.y * wc.x - vr.x * wc.y + wr.y * vc.x - wr.x * vc.y;
                arow[c] = a;
                if (vnext && c >= 1) {
                    const double2 z = c_mul(a, vnext[c - 1]);
                    acc->x += z.x;
                    acc->y += z.y;
                }
            }
        };
        if (w == 0) update_row(0, Am + (int64_t)(k + 1) * n + k + 1, nullptr, nullptr);
        __syncthreads();
        if (k + 2 >= n) break;  // the last column needs no reflector
        const int nxt = cur ^ 1;
        const double2 tk1 = reflector(k + 1, vbuf[nxt]);
        // remaining rows: update, then their product with v_{k+1}
        for (int rr = 1 + w; rr < m; rr += WT_WAVES) {
            double2 acc = make_double2(0, 0);
            update_row(rr, Am + (int64_t)(k + 1 + rr) * n + k + 1, vbuf[nxt], &acc);
            acc.x = wsum(acc.x);
            acc.y = wsum(acc.y);
            if (lane == 0) pbuf[nxt][rr - 1] = c_mul(tk1, acc);
        }
        __syncthreads();
        al = pv_alpha(m - 1, vbuf[nxt], pbuf[nxt], tk1);
        tk = tk1;
        cur = nxt;
    }
    if (tid == 0) {
        dg[n - 1] = Am[(int64_t)(n - 1) * n + n - 1].x;
        eg[n - 1] = 0.0;
    }
}

// 1w'. Blocked form of the same reduction (LAPACK zlatrd/zhetrd structure):
// the columns are taken in panels of NB; inside a panel the trailing matrix
// is left as it was at the panel's start and the pending rank-2 updates are
// applied through the panel's reflectors V and vectors W, held in LDS:
//   column k:  x = A[k:, k] - sum_q (v_q conj(w_q[k]) + w_q conj(v_q[k]))
//   reflector of x[k+1:] (same conventions as above), v_k into row k
//   p = tau (A22 v - sum_q v_q (w_q^H v) - sum_q w_q (v_q^H v))
//   w = p - tau/2 (p^H v) v
// and after the panel one rank-2 NB update A22 -= V W^H + W V^H of the
// trailing square.  Each column so reads the trailing matrix once (the
// product) and the trailing matrix is written once per panel, where the
// unblocked pass reads and writes it once per column.  Same outputs (rows
// holding v, d, e, tau) as hetrd_wide_kernel.
template <int NB, int NMX, bool LOW>
__global__ __launch_bounds__(WT_WAVES * 64) void hetrd_wide_blk_kernel(const c64 *__restrict__ G, int64_t sG,
                                                                       char *__restrict__ Wk, int64_t sW, int n,
                                                                       int kw, const int *__restrict__ mask) {
    constexpr int NT = WT_WAVES * 64;
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ double2 Vs[NB][NMX], Ws[NB][NMX];
    __shared__ double2 xb[NMX], pb[NMX];
    // LOW: per-wavefront partials of the product's upper half (column sums)
    __shared__ double2 tpart[LOW ? WT_WAVES : 1][LOW ? NMX : 1];
    __shared__ double2 cw[NB], cv[NB];  // w_q^H v, v_q^H v
    __shared__ double rs[WT_WAVES];
    __shared__ double2 rs2[WT_WAVES];
    __shared__ double2 tk_s, al_s;
    const HeevLayW lay = heev_lay_wide(n, kw);
    char *base = Wk + t * sW;
    double2 *Am = (double2 *)(base + lay.A);
    double2 *taug = (double2 *)(base + lay.tau);
    double *dg = (double *)(base + lay.d);
    double *eg = (double *)(base + lay.e);
    const c64 *Gt = G + t * sG;
    for (int64_t e = tid; e < (int64_t)n * n; e += NT) Am[e] = make_double2(Gt[e].re, Gt[e].im);
    __syncthreads();
    for (int k0 = 0; k0 + 1 < n; k0 += NB) {
        const int nbp = min(NB, n - 1 - k0);  // reflectors of this panel (columns k0 .. k0 + nbp - 1)
        for (int j = 0; j < nbp; ++j) {
            const int k = k0 + j;
            // corrected column k, rows k .. n-1 (row k of the full storage, conjugated)
            for (int r = k + tid; r < n; r += NT) {
                double2 x;
                if (LOW) x = Am[(int64_t)r * n + k];  // column k of the lower triangle
                else x = r == k ? Am[(int64_t)k * n + k] : c_conj(Am[(int64_t)k * n + r]);
                for (int q = 0; q < j; ++q) {
                    const double2 a = c_mul(Vs[q][r], c_conj(Ws[q][k])), b = c_mul(Ws[q][r], c_conj(Vs[q][k]));
                    x.x -= a.x + b.x;
                    x.y -= a.y + b.y;
                }
                xb[r] = x;
            }
            __syncthreads();
            // reflector of x[k+1 ..]: sum |x|^2 over rows k+2 ..
            double xs = 0;
            for (int r = k + 2 + tid; r < n; r += NT) xs += xb[r].x * xb[r].x + xb[r].y * xb[r].y;
            xs = wsum(xs);
            if (lane == 0) rs[w] = xs;
            __syncthreads();
            xs = 0;
#pragma unroll
            for (int i = 0; i < WT_WAVES; ++i) xs += rs[i];
            const double2 alpha = xb[k + 1];
            double2 tk = make_double2(0, 0), inv = make_double2(0, 0);
            double beta = alpha.x;
            if (!(xs == 0.0 && alpha.y == 0.0)) {
                const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xs);
                beta = alpha.x >= 0 ? -nr : nr;
                tk = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
                const double2 den = make_double2(alpha.x - beta, alpha.y);
                const double dn = den.x * den.x + den.y * den.y;
                inv = make_double2(den.x / dn, -den.y / dn);
            }
            for (int r = tid; r < n; r += NT) {
                double2 v = make_double2(0, 0);
                if (r == k + 1) v = make_double2(1.0, 0.0);
                else if (r > k + 1) v = c_mul(inv, xb[r]);
                Vs[j][r] = v;
                if (r > k) Am[(int64_t)k * n + r] = v;
            }
            if (tid == 0) {
                dg[k] = xb[k].x;
                eg[k] = beta;
                taug[k] = tk;
            }
            __syncthreads();
            // w_q^H v and v_q^H v of the panel's earlier columns (one wavefront each)
            for (int u = w; u < 2 * j; u += WT_WAVES) {
                const int q = u >> 1;
                const double2 *a = (u & 1) ? Vs[q] : Ws[q];
                double2 acc = make_double2(0, 0);
                for (int r = k + 1 + lane; r < n; r += 64) {
                    const double2 z = c_mulc(a[r], Vs[j][r]);
                    acc.x += z.x;
                    acc.y += z.y;
                }
                acc.x = wsum(acc.x);
                acc.y = wsum(acc.y);
                if (lane == 0) {
                    if (u & 1) cv[q] = acc;
                    else cw[q] = acc;
                }
            }
            // A22 v with the panel-start matrix, one wavefront per row
            const int m = n - k - 1;
            if (LOW) {
                // lower triangle only: row rr gives sum_{c <= rr} A[rr][c] v_c to
                // y_rr and conj(A[rr][c]) v_rr to y_c (c < rr), the latter kept per
                // lane (lane owns columns c = lane + 64 q) and summed over the
                // wavefronts below in a fixed order
                const double2 *vs = &Vs[j][k + 1];
                constexpr int NQ = NMX / 64;
                double2 tacc[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) tacc[q] = make_double2(0, 0);
                // two rows per wavefront and pass (rr and rr + WT_WAVES), both
                // rows' loads issued before either is used; the per-lane
                // column sums still take the rows in increasing order
                for (int rr = w; rr < m; rr += 2 * WT_WAVES) {
                    const int r2 = rr + WT_WAVES;
                    const bool two = r2 < m;
                    const double2 *arow = Am + (int64_t)(k + 1 + rr) * n + k + 1;
                    const double2 *arow2 = Am + (int64_t)(k + 1 + (two ? r2 : rr)) * n + k + 1;
                    double2 a1[NQ], a2[NQ];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int c = lane + 64 * q;
                        a1[q] = c <= rr ? arow[c] : make_double2(0, 0);
                        a2[q] = two && c <= r2 ? arow2[c] : make_double2(0, 0);
                    }
                    const double2 vr = vs[rr], vr2 = vs[two ? r2 : rr];
                    double2 acc = make_double2(0, 0), acc2 = make_double2(0, 0);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int c = lane + 64 * q;
                        if (c <= rr) {
                            const double2 a = a1[q];
                            const double2 z = c_mul(a, vs[c]);
                            acc.x += z.x;
                            acc.y += z.y;
                            if (c < rr) {
                                const double2 u = c_mulc(a, vr);
                                tacc[q].x += u.x;
                                tacc[q].y += u.y;
                            }
                        }
                    }
                    if (two) {
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const int c = lane + 64 * q;
                            if (c <= r2) {
                                const double2 a = a2[q];
                                const double2 z = c_mul(a, vs[c]);
                                acc2.x += z.x;
                                acc2.y += z.y;
                                if (c < r2) {
                                    const double2 u = c_mulc(a, vr2);
                                    tacc[q].x += u.x;
                                    tacc[q].y += u.y;
                                }
                            }
                        }
                    }
                    acc.x = wsum(acc.x);
                    acc.y = wsum(acc.y);
                    acc2.x = wsum(acc2.x);
                    acc2.y = wsum(acc2.y);
                    if (lane == 0) {
                        pb[k + 1 + rr] = acc;
                        if (two) pb[k + 1 + r2] = acc2;
                    }
                }
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (lane + 64 * q < m) tpart[w][lane + 64 * q] = tacc[q];
                __syncthreads();
                for (int c = tid; c < m; c += NT) {
                    double2 y = pb[k + 1 + c];
#pragma unroll
                    for (int i = 0; i < WT_WAVES; ++i) {
                        y.x += tpart[i][c].x;
                        y.y += tpart[i][c].y;
                    }
                    pb[k + 1 + c] = y;
                }
                __syncthreads();
            } else {
                for (int rr = w; rr < m; rr += WT_WAVES) {
                    const double2 *arow = Am + (int64_t)(k + 1 + rr) * n + k + 1;
                    const double2 *vs = &Vs[j][k + 1];
                    double2 acc = make_double2(0, 0);
                    for (int c = lane; c < m; c += 64) {
                        const double2 z = c_mul(arow[c], vs[c]);
                        acc.x += z.x;
                        acc.y += z.y;
                    }
                    acc.x = wsum(acc.x);
                    acc.y = wsum(acc.y);
                    if (lane == 0) pb[k + 1 + rr] = acc;
                }
                __syncthreads();
            }
            // p = tau (A22 v - V cw - W cv), and p^H v
            double2 pvp = make_double2(0, 0);
            for (int r = k + 1 + tid; r < n; r += NT) {
                double2 y = pb[r];
                for (int q = 0; q < j; ++q) {
                    const double2 a = c_mul(Vs[q][r], cw[q]), b = c_mul(Ws[q][r], cv[q]);
                    y.x -= a.x + b.x;
                    y.y -= a.y + b.y;
                }
                y = c_mul(tk, y);
                pb[r] = y;
                const double2 z = c_mulc(y, Vs[j][r]);
                pvp.x += z.x;
                pvp.y += z.y;
            }
            pvp.x = wsum(pvp.x);
            pvp.y = wsum(pvp.y);
            if (lane == 0) rs2[w] = pvp;
            __syncthreads();
            double2 pv = make_double2(0, 0);
#pragma unroll
            for (int i = 0; i < WT_WAVES; ++i) {
                pv.x += rs2[i].x;
                pv.y += rs2[i].y;
            }
            const double2 al = c_mul(make_double2(-0.5 * tk.x, -0.5 * tk.y), pv);
            for (int r = tid; r < n; r += NT) {
                double2 wv = make_double2(0, 0);
                if (r > k) {
                    const double2 z = c_mul(al, Vs[j][r]);
                    wv = make_double2(pb[r].x + z.x, pb[r].y + z.y);
                }
                Ws[j][r] = wv;
            }
            __syncthreads();
        }
        // the panel's rank-2 nbp update of the trailing square (both triangles,
        // or with LOW the lower one)
        const int k1 = k0 + nbp, m1 = n - k1;
        // (the row's v and w of eight reflectors at a time in registers)
        constexpr int QC = 8;
        for (int rr = w; rr < m1; rr += WT_WAVES) {
            const int r = k1 + rr;
            double2 *arow = Am + (int64_t)r * n;
            for (int q0 = 0; q0 < nbp; q0 += QC) {
                double2 vr[QC], wr[QC];
#pragma unroll
                for (int q = 0; q < QC; ++q) {
                    vr[q] = q0 + q < nbp ? Vs[q0 + q][r] : make_double2(0, 0);
                    wr[q] = q0 + q < nbp ? Ws[q0 + q][r] : make_double2(0, 0);
                }
                const int qn = min(QC, nbp - q0);
                const int cend = LOW ? r + 1 : n;  // LOW: the lower triangle only
                for (int c = k1 + lane; c < cend; c += 64) {
                    double2 a = arow[c];
#pragma unroll
                    for (int q = 0; q < QC; ++q) {
                        if (q >= qn) break;
                        const double2 x1 = c_mul(vr[q], c_conj(Ws[q0 + q][c]));
                        const double2 x2 = c_mul(wr[q], c_conj(Vs[q0 + q][c]));
                        a.x -= x1.x + x2.x;
                        a.y -= x1.y + x2.y;
                    }
                    arow[c] = a;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        dg[n - 1] = Am[(int64_t)(n - 1) * n + n - 1].x;
        eg[n - 1] = 0.0;
    }
}

// 3w. Kept eigenvectors of T by inverse iteration for the wide path: the
// algorithm of stein_kernel with kw lanes (kw / 64 wavefronts, lane c = the
// c-th largest eigenvalue) and X in global memory [row][kw]; the
// orthogonalisation pass is cgs_wide_kernel.
__global__ __launch_bounds__(WIDE_MAX_K) void stein_wide_kernel(char *__restrict__ Wk, int64_t sW, int n, int kw,
                                                                const double *__restrict__ evals, int64_t sE,
                                                                const int *__restrict__ keep,
                                                                const int *__restrict__ mask, int *__restrict__ fallback) {
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    const int c = threadIdx.x;
    const int kc = min(keep[t], kw);
    const HeevLayW lay = heev_lay_wide(n, kw);
    char *base = Wk + t * sW;
    const double *dg = (const double *)(base + lay.d);
    const double *eg = (const double *)(base + lay.e);
    double *ui = (double *)(base + lay.lu);  // [i][kw]: 1 / U diagonal
    double *ub = ui + (int64_t)n * kw;       // U first superdiagonal
    double *uc = ub + (int64_t)n * kw;       // U second superdiagonal
    double *ml = uc + (int64_t)n * kw;       // L multipliers
    double *X = (double *)(base + lay.X);    // X[row * kw + col]
    __shared__ double dd[WIDE_MAX_N], es[WIDE_MAX_N];
    __shared__ int bad;
    for (int i = c; i < n; i += blockDim.x) {
        dd[i] = dg[i];
        es[i] = i + 1 < n ? eg[i] : 0.0;
    }
    if (c == 0) bad = 0;
    __syncthreads();
    const double *ev = evals + t * sE;
    double tn = 0;
    for (int i = 0; i < n; ++i) tn = fmax(tn, fabs(dd[i]) + fabs(es[i]) + (i ? fabs(es[i - 1]) : 0.0));
    const double tiny = EPS * fmax(tn, 1e-300);
    const double ctol = 1e-9 * tn;
    const bool lead = c < kc && (c == 0 || ev[c - 1] - ev[c] > ctol);
    int c1 = c + 1;
    if (lead)
        while (c1 < kc && ev[c1 - 1] - ev[c1] <= ctol) ++c1;
    if (lead) {
        for (int j = c; j < c1; ++j) {
            const double lam = ev[j];
            uint64_t piv[WIDE_MAX_N / 64];
#pragma unroll
            for (int q = 0; q < WIDE_MAX_N / 64; ++q) piv[q] = 0;
            double a_cur = dd[0] - lam, b_cur = n > 1 ? es[0] : 0.0;
            auto put_diag = [&](int i, double u) {
                if (fabs(u) < tiny) u = u < 0.0 ? -tiny : tiny;
                ui[(int64_t)i * kw + c] = 1.0 / u;
            };
            for (int i = 0; i + 1 < n; ++i) {
                const double ci = es[i];
                const double a1 = dd[i + 1] - lam;
                const double b1 = i + 2 < n ? es[i + 1] : 0.0;
                const int64_t o = (int64_t)i * kw + c;
                if (fabs(a_cur) >= fabs(ci)) {
                    const double l = a_cur != 0.0 ? ci / a_cur : 0.0;
                    put_diag(i, a_cur);
                    ub[o] = b_cur;
                    uc[o] = 0.0;
                    ml[o] = l;
                    a_cur = a1 - l * b_cur;
                    b_cur = b1;
                } else {
                    const double l = a_cur / ci;
                    put_diag(i, ci);
                    ub[o] = a1;
                    uc[o] = b1;
                    ml[o] = l;
                    piv[i >> 6] |= 1ull << (i & 63);
                    a_cur = b_cur - l * a1;
                    b_cur = -l * b1;
                }
            }
            put_diag(n - 1, a_cur);
            double *x = X + j;  // column j, stride kw
            for (int i = 0; i < n; ++i) {
                const uint32_t h = (uint32_t)(i * 2654435761u) ^ (uint32_t)((j + 1) * 2246822519u);
                x[(int64_t)i * kw] = 0.5 + (double)((h >> 8) & 0xffff) * (1.0 / 65536.0);
            }
            // X sits in global memory beside the factors, so the compiler
            // may not move a row's loads above the previous row's store:
            // the solves load the operands of SCH rows before using them
            // (same arithmetic, same order)
            constexpr int SCH = 8;
            for (int it = 0; it < 3; ++it) {
                double r0 = x[0];
                for (int i0 = 0; i0 + 1 < n; i0 += SCH) {
                    double lv[SCH], xv[SCH];
#pragma unroll
                    for (int u = 0; u < SCH; ++u) {
                        const int i = i0 + u;
                        lv[u] = xv[u] = 0.0;
                        if (i + 1 < n) {
                            lv[u] = ml[(int64_t)i * kw + c];
                            xv[u] = x[(int64_t)(i + 1) * kw];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < SCH; ++u) {
                        const int i = i0 + u;
                        if (i + 1 < n) {
                            double r1 = xv[u];
                            if ((piv[i >> 6] >> (i & 63)) & 1ull) {
                                const double tmp = r0;
                                r0 = r1;
                                r1 = tmp;
                            }
                            x[(int64_t)i * kw] = r0;
                            r0 = r1 - lv[u] * r0;
                        }
                    }
                }
                x[(int64_t)(n - 1) * kw] = r0;
                double y2 = 0.0, y1 = 0.0;
                for (int i0 = n - 1; i0 >= 0; i0 -= SCH) {
                    double xv[SCH], bv[SCH], cv[SCH], dv[SCH];
#pragma unroll
                    for (int u = 0; u < SCH; ++u) {
                        const int i = i0 - u;
                        xv[u] = bv[u] = cv[u] = dv[u] = 0.0;
                        if (i >= 0) {
                            const int64_t o = (int64_t)i * kw + c;
                            xv[u] = x[(int64_t)i * kw];
                            bv[u] = ub[o];
                            cv[u] = uc[o];
                            dv[u] = ui[o];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < SCH; ++u) {
                        if (i0 - u >= 0) {
                            const double y = (xv[u] - bv[u] * y1 - cv[u] * y2) * dv[u];
                            x[(int64_t)(i0 - u) * kw] = y;
                            y2 = y1;
                            y1 = y;
                        }
                    }
                }
                for (int p = c; p < j; ++p) {
                    double sp = 0;
                    for (int i = 0; i < n; ++i) sp += X[(int64_t)i * kw + p] * x[(int64_t)i * kw];
                    for (int i = 0; i < n; ++i) x[(int64_t)i * kw] -= sp * X[(int64_t)i * kw + p];
                }
                double nr = 0;
                for (int i = 0; i < n; ++i) nr += x[(int64_t)i * kw] * x[(int64_t)i * kw];
                if (!(nr > 0.0) || !isfinite(nr)) {
                    bad = 1;
                    break;
                }
                const double inv = 1.0 / sqrt(nr);
                for (int i = 0; i < n; ++i) x[(int64_t)i * kw] *= inv;
            }
        }
    }
    __syncthreads();
    if (c >= kc)
        for (int i = 0; i < n; ++i) X[(int64_t)i * kw + c] = 0.0;
    if (c == 0) fallback[t] = bad;
}

// 3w'. One classical Gram-Schmidt pass over the kept columns (as at the end
// of stein_kernel): S[p][c] = x_p . x_c for p < c (into the LU area), then
// X2[:, c] = x_c - sum_{p<c} S[p][c] x_p from the unmodified inputs, then
// unit columns written back to X.  Thread c owns column c.
__global__ __launch_bounds__(WIDE_MAX_K) void cgs_wide_kernel(char *__restrict__ Wk, int64_t sW, int n, int kw,
                                                              const int *__restrict__ keep, const int *__restrict__ mask,
                                                              const int *__restrict__ fallback) {
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (fallback[t]) return;
    const int c = threadIdx.x;
    const int kc = min(keep[t], kw);
    const HeevLayW lay = heev_lay_wide(n, kw);
    char *base = Wk + t * sW;
    double *S = (double *)(base + lay.lu);  // [p][kw]
    double *X = (double *)(base + lay.X);
    double *X2 = (double *)(base + lay.X2);
    constexpr int CH = 16;
    // wave-uniform bound on the columns any lane of this wave needs
    const int cmax = min(kw, (c | 63) + 1);
    for (int p0 = 0; p0 < cmax; p0 += CH) {
        double acc[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) acc[q] = 0.0;
        for (int i = 0; i < n; ++i) {
            const double xc = X[(int64_t)i * kw + c];
#pragma unroll
            for (int q = 0; q < CH; ++q) acc[q] += X[(int64_t)i * kw + p0 + q] * xc;
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) S[(int64_t)(p0 + q) * kw + c] = p0 + q < c ? acc[q] : 0.0;
    }
    for (int i = 0; i < n; ++i) X2[(int64_t)i * kw + c] = X[(int64_t)i * kw + c];
    for (int p0 = 0; p0 < cmax; p0 += CH) {
        double sp[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) sp[q] = S[(int64_t)(p0 + q) * kw + c];
        for (int i = 0; i < n; ++i) {
            double y = X2[(int64_t)i * kw + c];
#pragma unroll
            for (int q = 0; q < CH; ++q) y -= sp[q] * X[(int64_t)i * kw + p0 + q];
            X2[(int64_t)i * kw + c] = y;
        }
    }
    double nr = 0;
    for (int i = 0; i < n; ++i) {
        const double y = X2[(int64_t)i * kw + c];
        nr += y * y;
    }
    const double inv = (c < kc && nr > 0.0) ? 1.0 / sqrt(nr) : 0.0;
    __syncthreads();  // every lane has read the original X
    for (int i = 0; i < n; ++i) X[(int64_t)i * kw + c] = X2[(int64_t)i * kw + c] * inv;
}

// 3w''. The pass of cgs_wide_kernel for n <= 256, kw <= 128 as two
// register-blocked products staged through LDS by 1024 threads (the
// per-column loop above runs one dependent global load per row):
//   S = X^T X: thread (ta, tb) = (tid / 32, tid % 32) forms the 4 x 4 block
//   S[4 ta + u][4 tb + v] over chunks of 32 rows of X in LDS; the entries
//   with p >= c are dropped (stored as 0) and S goes to the LU area;
//   Y = X - X S: thread (ti, tc) keeps Y[8 ti + r][4 tc + v] in registers
//   while p runs over chunks of 16 (X[:, p0 .. p0 + 16) and
//   S[p0 .. p0 + 16][:] in LDS);
//   column norms: per-thread partials over its 8 rows, summed over ti in a
//   fixed order; unit columns (zero beyond the kept rank) written to X.
constexpr int CGW_THREADS = 1024;
__global__ __launch_bounds__(CGW_THREADS) void cgs_wide2_kernel(char *__restrict__ Wk, int64_t sW, int n, int kw,
                                                                const int *__restrict__ keep,
                                                                const int *__restrict__ mask,
                                                                const int *__restrict__ fallback) {
    constexpr int R = 32, PC = 16, KM = 128, NM = 256;
    // phase 1: rows of X [R][KM]; phase 2: X[:, p0 ..) [NM][PC] and
    // S[p0 ..)[:] [PC][KM]; then the norm partials [32][KM]
    __shared__ double lds[NM * PC + PC * KM + 32 * KM];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (fallback[t]) return;
    const int tid = threadIdx.x;
    const int kc = min(keep[t], kw);
    const HeevLayW lay = heev_lay_wide(n, kw);
    char *base = Wk + t * sW;
    double *S = (double *)(base + lay.lu);  // [p][kw]
    double *X = (double *)(base + lay.X);   // [i][kw]
    const int ta = tid >> 5, tb = tid & 31;
    {
        double acc[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
        const bool act = 4 * ta < kw && 4 * tb < kw;
        for (int i0 = 0; i0 < n; i0 += R) {
            __syncthreads();
            for (int e = tid; e < R * kw; e += CGW_THREADS) {
                const int r = e / kw, c = e - r * kw;
                lds[r * KM + c] = i0 + r < n ? X[(int64_t)(i0 + r) * kw + c] : 0.0;
            }
            __syncthreads();
            if (act)
#pragma unroll 2
                for (int r = 0; r < R; ++r) {
                    const double *row = lds + r * KM;
                    double a[4], b[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        a[u] = row[4 * ta + u];
                        b[u] = row[4 * tb + u];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
                }
        }
        if (act)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int p = 4 * ta + u, c = 4 * tb + v;
                    S[(int64_t)p * kw + c] = p < c ? acc[u][v] : 0.0;
                }
    }
    const int ti = ta, tc = tb;
    const bool act = 4 * tc < kw;
    double y[8][4];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 8 * ti + r;
            y[r][v] = act && i < n ? X[(int64_t)i * kw + 4 * tc + v] : 0.0;
        }
    double *xp = lds, *sp = lds + NM * PC, *nq = lds + NM * PC + PC * KM;
    for (int p0 = 0; p0 < kw; p0 += PC) {
        __syncthreads();  // S complete; the previous chunk consumed
        for (int e = tid; e < NM * PC; e += CGW_THREADS) {
            const int i = e / PC, q = e % PC;
            xp[e] = i < n ? X[(int64_t)i * kw + p0 + q] : 0.0;
        }
        for (int e = tid; e < PC * kw; e += CGW_THREADS) {
            const int q = e / kw, c = e - q * kw;
            sp[q * KM + c] = S[(int64_t)(p0 + q) * kw + c];
        }
        __syncthreads();
        if (act && 8 * ti < n)
#pragma unroll 1
            for (int q = 0; q < PC; ++q) {
                double s4[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) s4[v] = sp[q * KM + 4 * tc + v];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const double xv = xp[(8 * ti + r) * PC + q];
#pragma unroll
                    for (int v = 0; v < 4; ++v) y[r][v] = fma(-s4[v], xv, y[r][v]);
                }
            }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double sq = 0.0;
#pragma unroll
        for (int r = 0; r < 8; ++r) sq += y[r][v] * y[r][v];
        nq[ti * KM + 4 * tc + v] = sq;
    }
    __syncthreads();  // every thread has staged its last chunk of X
    if (!act) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = 4 * tc + v;
        double sq = 0.0;
        for (int w = 0; w < 32; ++w) sq += nq[w * KM + c];
        const double inv = (c < kc && sq > 0.0) ? 1.0 / sqrt(sq) : 0.0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = 8 * ti + r;
            if (i < n) X[(int64_t)i * kw + c] = y[r][v] * inv;
        }
    }
}

int64_t heev_work_bytes(int n, int kmax) {
    // calls with any n' <= n use the same slots
    const int64_t narrow = heev_lay(std::min(n, LDS_MAX_N)).total;
    if (n > WIDE_MAX_N) return narrow;
    const int kw = round64(std::min(std::min(n, kmax), WIDE_MAX_K));
    if (n <= LDS_MAX_N && kw <= 64) return narrow;
    return std::max(narrow, heev_lay_wide(n, kw).total);
}

static size_t heev_lds_shmem(int n) {
    size_t shmem = (size_t)n * (4 * sizeof(double) + sizeof(int)) + 16;
    if (n <= LDS_MAX_N) shmem += (size_t)(n * (n + 1) / 2 + n) * sizeof(double2);
    return shmem;
}

void heev_fixup(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP,
                const int *fallback, int n, int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
                double thr, int max_bond, const int *mask, int B, hipStream_t s) {
    if (n > LDS_MAX_N && !Pg) throw HipError("heev: no global scratch for n > 128");
    hipLaunchKernelGGL(heev_lds_kernel, dim3(B), dim3(256), heev_lds_shmem(n), s, G, sG, U, sU, Z, sZ, Pg, sP, n, ns,
                       evals, perm, sE, keep, rule, thr, max_bond, mask, fallback);
    YQ_HIP(hipGetLastError());
}

// Wide-path tridiagonalisation: 1 = blocked, lower triangle only up to
// n = 256 (the default), 2 = blocked with full rows, 0 = unblocked
// (YAQS_HETRD_BLOCKED overrides).
static int hetrd_wide_mode() {
    static const int m = [] {
        const char *e = std::getenv("YAQS_HETRD_BLOCKED");
        return e ? std::atoi(e) : 1;
    }();
    return m;
}

bool heev_gram_lower_only(int n) { return n > LDS_MAX_N && n <= 256 && hetrd_wide_mode() == 1; }

bool heev(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP, void *work,
          int64_t sW, int *fallback, int n, int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
          double thr, int max_bond, const int *mask, int B, hipStream_t s, bool fixup) {
    if (n > 1024) throw HipError("heev: n > 1024");
    int kout = ns < max_bond ? ns : max_bond;
    if (kout > n) kout = n;
    const bool fast = n <= LDS_MAX_N && kout <= 64 && work && fallback && heev_lay(n).total <= sW;
    const int kw = round64(kout);
    // the global-memory kernels also take n <= 128 when more than 64
    // columns are kept (the LDS kernels hold 64)
    const bool wide = !fast && n <= WIDE_MAX_N && kw <= WIDE_MAX_K && work && fallback &&
                      heev_lay_wide(n, kw).total <= sW;
    const int np = n * (n + 1) / 2;
    if (n > LDS_MAX_N && !Pg) throw HipError("heev: no global scratch for n > 128");
    char *W = (char *)work;
    if (fast) {
        const HeevLay lay = heev_lay(n);
        const size_t sh1 = (size_t)(np + (4 + TRD_PARTS) * n) * sizeof(double2);
        if (n > 32)
            hipLaunchKernelGGL(hetrd_reg_kernel, dim3(B), dim3(TRR_THREADS), 0, s, G, sG, W, sW, n, mask);
        else
            hipLaunchKernelGGL(hetrd_kernel, dim3(B), dim3(TRD_WAVES * 64), sh1, s, G, sG, W, sW, n, mask);
        YQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(stebz_kernel<LDS_MAX_N>, dim3(B), dim3(LDS_MAX_N), 0, s, W, sW, lay.d, lay.e, n, ns, evals,
                           perm, sE, keep, rule, thr, max_bond, mask);
        YQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(stein_kernel, dim3(B), dim3(64), (size_t)(n * 64 + 2 * n) * sizeof(double), s, W, sW, n,
                           kout, (const double *)evals, sE, (const int *)keep, mask, fallback);
        YQ_HIP(hipGetLastError());
        hipLaunchKernelGGL((ormtr_kernel<LDS_MAX_N, 8, 64>), dim3(B, 1), dim3(512), 0, s, W, sW, lay.V, lay.tau, lay.X, 64,
                           n, kout, U, sU, mask, (const int *)fallback);
        YQ_HIP(hipGetLastError());
    } else if (wide) {
        const HeevLayW lay = heev_lay_wide(n, kw);
        const int blk = hetrd_wide_mode();
        if (blk == 2 && n <= 256)
            hipLaunchKernelGGL((hetrd_wide_blk_kernel<16, 256, false>), dim3(B), dim3(WT_WAVES * 64), 0, s, G, sG, W,
                               sW, n, kw, mask);
        else if (blk && n <= 256)
            hipLaunchKernelGGL((hetrd_wide_blk_kernel<8, 256, true>), dim3(B), dim3(WT_WAVES * 64), 0, s, G, sG, W,
                               sW, n, kw, mask);
        else if (blk)
            hipLaunchKernelGGL((hetrd_wide_blk_kernel<8, WIDE_MAX_N, false>), dim3(B), dim3(WT_WAVES * 64), 0, s, G,
                               sG, W, sW, n, kw, mask);
        else
            hipLaunchKernelGGL(hetrd_wide_kernel, dim3(B), dim3(WT_WAVES * 64), 0, s, G, sG, W, sW, n, kw, mask);
        YQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(stebz_kernel<WIDE_MAX_N>, dim3(B), dim3(round64(n)), 0, s, W, sW, lay.d, lay.e, n, ns,
                           evals, perm, sE, keep, rule, thr, max_bond, mask);
        YQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(stein_wide_kernel, dim3(B), dim3(kw), 0, s, W, sW, n, kw, (const double *)evals, sE,
                           (const int *)keep, mask, fallback);
        YQ_HIP(hipGetLastError());
        if (n <= 256 && kw <= 128)
            hipLaunchKernelGGL(cgs_wide2_kernel, dim3(B), dim3(CGW_THREADS), 0, s, W, sW, n, kw, (const int *)keep, mask,
                               (const int *)fallback);
        else
            hipLaunchKernelGGL(cgs_wide_kernel, dim3(B), dim3(kw), 0, s, W, sW, n, kw, (const int *)keep, mask,
                               (const int *)fallback);
        YQ_HIP(hipGetLastError());
        if (n <= 256)  // 16 row residues per 64 columns
            hipLaunchKernelGGL((ormtr_kernel<256, 16, 64>), dim3(B, kw / 64), dim3(1024), 0, s, W, sW, lay.A,
                               lay.tau, lay.X, kw, n, kout, U, sU, mask, (const int *)fallback);
        else
            hipLaunchKernelGGL((ormtr_kernel<WIDE_MAX_N, 16, 32>), dim3(B, kw / 32), dim3(1024), 0, s, W, sW, lay.A,
                               lay.tau, lay.X, kw, n, kout, U, sU, mask, (const int *)fallback);
        YQ_HIP(hipGetLastError());
    } else {
        heev_fixup(G, sG, U, sU, Z, sZ, Pg, sP, nullptr, n, ns, evals, perm, sE, keep, rule, thr, max_bond, mask, B, s);
        return false;
    }
    // matrices whose inverse iteration broke down: full single-workgroup solver
    if (fixup)
        heev_fixup(G, sG, U, sU, Z, sZ, Pg, sP, fallback, n, ns, evals, perm, sE, keep, rule, thr, max_bond, mask, B,
                   s);
    return true;
}

}  // namespace yq

// Test hook (not part of the public header): eigen-decomposition of `count`
// Hermitian n x n matrices (row-major interleaved complex doubles, stored
// back to back) on the device; writes the sorted eigenvalues, U (n x n, the
// first min(n, max_bond) columns valid) and the kept rank of each.
// path: 0 = automatic, 1 = single-workgroup solver only.
extern "C" int yaqs_debug_heev_batch(int n, int count, const double *G, double *evals, double *U, int *keep_out,
                                     int *fallback_out, int rule, double thr, int max_bond, int path) {
    using namespace yq;
    try {
        const int64_t nn = (int64_t)n * n, sZ = nn + 2 * n + 4;
        c64 *dG, *dU;
        double *dZ, *dE;
        double2 *dP = nullptr;
        int *dperm, *dkeep, *dfb;
        char *dW = nullptr;
        const int64_t sW = path == 1 ? 0 : heev_work_bytes(n, max_bond);
        YQ_HIP(hipMalloc(&dG, nn * count * sizeof(c64)));
        YQ_HIP(hipMalloc(&dU, nn * count * sizeof(c64)));
        YQ_HIP(hipMalloc(&dZ, sZ * count * sizeof(double)));
        YQ_HIP(hipMalloc(&dE, (n + 2) * count * sizeof(double)));
        YQ_HIP(hipMalloc(&dperm, (n + 2) * count * sizeof(int)));
        YQ_HIP(hipMalloc(&dkeep, count * sizeof(int)));
        YQ_HIP(hipMalloc(&dfb, count * sizeof(int)));
        YQ_HIP(hipMemset(dfb, 0, count * sizeof(int)));
        if (sW) YQ_HIP(hipMalloc(&dW, sW * count));
        if (n > 128) YQ_HIP(hipMalloc(&dP, (int64_t)n * (n + 1) / 2 * count * sizeof(double2)));
        YQ_HIP(hipMemcpy(dG, G, nn * count * sizeof(c64), hipMemcpyHostToDevice));
        YQ_HIP(hipMemset(dU, 0, nn * count * sizeof(c64)));
        heev(dG, nn, dU, nn, dZ, sZ, dP, (int64_t)n * (n + 1) / 2, dW, sW, sW ? dfb : nullptr, n, n, dE, dperm, n + 2,
             dkeep, rule, thr, max_bond, nullptr, count, 0);
        YQ_HIP(hipDeviceSynchronize());
        for (int b = 0; b < count; ++b)
            YQ_HIP(hipMemcpy(evals + (int64_t)b * n, dE + (int64_t)b * (n + 2), n * sizeof(double),
                             hipMemcpyDeviceToHost));
        YQ_HIP(hipMemcpy(U, dU, nn * count * sizeof(c64), hipMemcpyDeviceToHost));
        YQ_HIP(hipMemcpy(keep_out, dkeep, count * sizeof(int), hipMemcpyDeviceToHost));
        if (fallback_out) YQ_HIP(hipMemcpy(fallback_out, dfb, count * sizeof(int), hipMemcpyDeviceToHost));
        (void)hipFree(dG);
        (void)hipFree(dU);
        (void)hipFree(dZ);
        (void)hipFree(dE);
        (void)hipFree(dperm);
        (void)hipFree(dkeep);
        (void)hipFree(dfb);
        if (dW) (void)hipFree(dW);
        if (dP) (void)hipFree(dP);
        return 0;
    } catch (const std::exception &) {
        return 3;
    }
}

extern "C" int yaqs_debug_heev(int n, const double *G, double *evals, double *U, int *keep_out, int rule, double thr,
                               int max_bond) {
    return yaqs_debug_heev_batch(n, 1, G, evals, U, keep_out, nullptr, rule, thr, max_bond, 0);
}
