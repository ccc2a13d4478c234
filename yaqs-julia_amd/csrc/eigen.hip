// Batched Hermitian eigensolver for the FP64 Gram matrices of the bond
// splits (the SVD of split_mps_tensor_svd, src/Algorithms.jl:409-445).
//
// Fast path (n <= 128, i.e. bond dimensions up to 64), four kernels:
//   1. hetrd_kernel: one 256-thread workgroup per matrix reduces the packed
//      lower triangle (LDS, <= 144 KB) to a real symmetric tridiagonal T with
//      LAPACK zhetd2/zlarfg conventions (Q = H_0 ... H_{n-2},
//      H_k = I - tau_k v_k v_k^H, subdiagonal real).  Three barriers per
//      column: the Hermitian product p = tau A22 v (two column halves per
//      row), the reduction of p^H v, and the rank-2 update, which also
//      extracts the next column and its norm.
//   2. tql_kernel: one LANE per matrix runs implicit QL with Wilkinson shifts
//      on (d, e) in LDS and records every Givens rotation (c, s, i) to global
//      memory; the serial rotation chain of 64 matrices shares one wavefront,
//      so the latency of 1000 chains overlaps.  It then sorts the
//      eigenvalues, applies the truncation rule and writes the kept rank.
//   3. tqlvec_kernel: one wavefront per matrix rebuilds only the eigenvector
//      columns that can be kept, lane = column: x = G_1 ... G_K e_order[c],
//      applying the recorded rotations in reverse order to an LDS-resident
//      n x 64 block.
//   4. ormtr_kernel: u = H_0 ... H_{n-2} x for the kept columns, with the
//      n x 64 complex block held in registers (thread = (row residue, column))
//      and one barrier per reflector.
// Matrices whose QL does not converge within the rotation budget, and every
// matrix with n > 128, go through heev_lds_kernel: one workgroup per matrix
// doing all stages (slower, same conventions).
// Output: U[:, r] for r < min(n, max_bond) in sorted order (row stride n),
// evals (sorted descending), perm[r] = r, keep.
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

__host__ __device__ __forceinline__ int pidx(int i, int j) { return i * (i + 1) / 2 + j; }  // j <= i

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
constexpr double EPS = 2.220446049250313e-16;
constexpr int QL_MAX_ITER = 60;

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
    int64_t V, tau, d, e, rot, ri, order, nrot, X, total;
    int cap;
};
__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }
__host__ __device__ inline HeevLay heev_lay(int n) {
    HeevLay l;
    l.cap = 3 * n * n + 64;
    int64_t o = 0;
    l.V = o;
    o = align256(o + (int64_t)n * n * 16);
    l.tau = o;
    o = align256(o + (int64_t)n * 16);
    l.d = o;
    o = align256(o + (int64_t)n * 8);
    l.e = o;
    o = align256(o + (int64_t)n * 8);
    l.rot = o;
    o = align256(o + (int64_t)l.cap * 16);
    l.ri = o;
    o = align256(o + (int64_t)l.cap * 4);
    l.order = o;
    o = align256(o + (int64_t)n * 4);
    l.nrot = o;
    o = align256(o + 16);
    l.X = o;
    o = align256(o + (int64_t)n * 64 * 8);
    l.total = o;
    return l;
}

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
constexpr int TRD_WAVES = 8;                  // 512 threads: two waves per SIMD hide LDS latency
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
                double2 acc0 = make_double2(0, 0), acc1 = make_double2(0, 0);
                const int rowb = pidx(k + 1 + r, k + 1);  // A22(r, c) = P[rowb + c] for c <= r
                int colb = pidx(k + 1 + c0, k + 1 + r);    // A22(c, r) = P[colb] for c > r
                int c = c0;
                for (; c + 1 < c1; c += 2) {
                    const bool lo0 = c <= r, lo1 = c + 1 <= r;
                    const int colb1 = colb + k + 2 + c;
                    double2 a0 = P[lo0 ? rowb + c : colb];
                    double2 a1 = P[lo1 ? rowb + c + 1 : colb1];
                    if (!lo0) a0.y = -a0.y;
                    if (!lo1) a1.y = -a1.y;
                    const double2 y0c = c == 0 ? y0 : x[c];
                    const double2 y1c = x[c + 1];
                    acc0.x += a0.x * y0c.x - a0.y * y0c.y;
                    acc0.y += a0.x * y0c.y + a0.y * y0c.x;
                    acc1.x += a1.x * y1c.x - a1.y * y1c.y;
                    acc1.y += a1.x * y1c.y + a1.y * y1c.x;
                    colb = colb1 + k + 3 + c;  // two rows further down the packed triangle
                }
                if (c < c1) {
                    const bool lo = c <= r;
                    double2 a = P[lo ? rowb + c : colb];
                    if (!lo) a.y = -a.y;
                    const double2 yc = c == 0 ? y0 : x[c];
                    acc0.x += a.x * yc.x - a.y * yc.y;
                    acc0.y += a.x * yc.y + a.y * yc.x;
                }
                pb[h * n + r] = make_double2(acc0.x + acc1.x, acc0.y + acc1.y);
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
        for (int r = w; r < m; r += TRD_WAVES) {
            const double2 vr = vf[r];
            const double2 pr = pf[r];
            const double2 wr = make_double2(pr.x + al.x * vr.x - al.y * vr.y, pr.y + al.x * vr.y + al.y * vr.x);
            const int rowb = pidx(k + 1 + r, k + 1);
            for (int c = lane; c <= r; c += 64) {
                const double2 vc = vf[c];
                const double2 pc = pf[c];
                const double2 wc = make_double2(pc.x + al.x * vc.x - al.y * vc.y, pc.y + al.x * vc.y + al.y * vc.x);
                double2 a = P[rowb + c];
                // a -= vr conj(wc) + wr conj(vc)
                a.x -= vr.x * wc.x + vr.y * wc.y + wr.x * vc.x + wr.y * vc.y;
                a.y -= vr.y * wc.x - vr.x * wc.y + wr.y * vc.x - wr.x * vc.y;
                P[rowb + c] = a;
                if (c == 0 && r >= 1) {
                    xn[r - 1] = a;
                    if (r >= 2) xsn += a.x * a.x + a.y * a.y;
                }
            }
        }
        xsn = wsum(xsn);
        if (lane == 0) sxs[cur ^ 1][w] = xsn;
        __syncthreads();
    }
    if (tid == 0) {
        dg[n - 1] = P[pidx(n - 1, n - 1)].x;
        eg[n - 1] = 0.0;
    }
}

// 2. Implicit QL on (d, e), one lane per matrix; rotations recorded.
__global__ __launch_bounds__(64) void tql_kernel(char *__restrict__ Wk, int64_t sW, int n, int ns, double *evals,
                                                 int *perm, int64_t sE, int *keep, int rule, double thr, int max_bond,
                                                 const int *__restrict__ mask, int *__restrict__ fallback, int B) {
    extern __shared__ double lds_de[];
    const int lane = threadIdx.x;
    const int t = blockIdx.x * 64 + lane;
    if (t >= B) return;
    if (mask && !mask[t]) return;
    const HeevLay lay = heev_lay(n);
    char *base = Wk + t * sW;
    const double *dg = (const double *)(base + lay.d);
    const double *eg = (const double *)(base + lay.e);
    double2 *rot = (double2 *)(base + lay.rot);
    int *ri = (int *)(base + lay.ri);
    double *dd = lds_de + lane;          // dd[i * 64]
    double *ee = lds_de + n * 64 + lane;  // ee[i * 64]
#define D(i) dd[(i) * 64]
#define E(i) ee[(i) * 64]
    double tn = 0;
    for (int i = 0; i < n; ++i) {
        const double di = dg[i], ei = eg[i];
        D(i) = di;
        E(i) = ei;
        tn = fmax(tn, fabs(di) + fabs(ei));
    }
    const double abstol = EPS * tn;
    // e_j negligible: relative to its diagonal neighbours or absolutely (eps ||T||)
    auto negl = [&](int j) {
        const double ej = fabs(E(j));
        return ej <= EPS * (fabs(D(j)) + fabs(D(j + 1))) || ej <= abstol;
    };
    auto scan = [&](int from) {
        int m = from;
        while (m < n - 1 && !negl(m)) ++m;
        return m;
    };
    int q = 0, l = 0, iter = 0;
    bool fail = false;
    int m = scan(0);
    for (;;) {
        // d_l has converged while the unreduced block starting at l is 1 x 1
        bool done = false;
        while (m == l) {
            if (++l >= n) {
                done = true;
                break;
            }
            iter = 0;
            if (m < l) m = scan(l);
        }
        if (done) break;
        if (iter >= QL_MAX_ITER || q + (m - l) > lay.cap) {
            fail = true;
            break;
        }
        ++iter;
        // Wilkinson-type shift from the leading 2x2 of the block [l, m]
        double g = (D(l + 1) - D(l)) / (2.0 * E(l));
        const double rr = hypot(g, 1.0);
        g = D(m) - D(l) + E(l) / (g + (g >= 0 ? rr : -rr));
        double s = 1.0, c = 1.0, pp = 0.0;
        int mn = m;  // smallest negligible index in (l, m] after the sweep
        bool deflated = false;
        // registers: ei = e_i, di = d_i (prefetched), di1 = d_{i+1} (original),
        // dn1 = d_{i+2} after this sweep
        double ei = E(m - 1), di = D(m - 1), di1 = D(m), dn1 = 0.0;
        for (int i = m - 1; i >= l; --i) {
            const double f = s * ei, b = c * ei;
            const double x = f * f + g * g;
            if (x == 0.0) {  // early deflation: the block splits at i + 1
                E(i + 1) = 0.0;
                D(i + 1) = di1 - pp;
                E(m) = 0.0;
                mn = i + 1;
                deflated = true;
                break;
            }
            // 1/sqrt(x): hardware estimate plus two Newton steps
            double ir = __builtin_amdgcn_rsq(x);
            const double hx = 0.5 * x;
            ir = ir * fma(-hx, ir * ir, 1.5);
            ir = ir * fma(-hx, ir * ir, 1.5);
            const double r = x * ir;
            E(i + 1) = r;
            s = f * ir;
            c = g * ir;
            const double ein = i > l ? E(i - 1) : 0.0, din = i > l ? D(i - 1) : 0.0;
            g = di1 - pp;
            const double r2 = (di - g) * s + 2.0 * c * b;
            pp = s * r2;
            const double dnew = g + pp;
            D(i + 1) = dnew;
            g = c * r2 - b;
            if (i + 1 < m && (r <= EPS * (fabs(dnew) + fabs(dn1)) || r <= abstol)) mn = i + 1;
            rot[q] = make_double2(c, s);
            ri[q] = i;
            ++q;
            dn1 = dnew;
            di1 = di;
            ei = ein;
            di = din;
        }
        if (!deflated) {
            D(l) = di1 - pp;
            E(l) = g;
            E(m) = 0.0;
        }
        m = negl(l) ? l : mn;
        if (m == l) {
            // converged at l: the next block starts at l + 1 and ends at mn
            if (++l >= n) break;
            iter = 0;
            m = mn;
        }
    }
    *(int *)(base + lay.nrot) = q;
    fallback[t] = fail ? 1 : 0;
    if (fail) return;
    // sort descending (stable by index), rule
    double *ev = evals + t * sE;
    int *pm = perm + t * sE;
    int *order = (int *)(base + lay.order);
    for (int i = 0; i < n; ++i) {
        const double di = D(i);
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = D(j);
            rank += (dj > di || (dj == di && j < i)) ? 1 : 0;
        }
        order[rank] = i;
        ev[rank] = di;
        pm[i] = i;
    }
    keep[t] = truncation_rank(ev, ns, rule, thr, max_bond);
#undef D
#undef E
}

// 3. Kept eigenvector columns of T, one wavefront per matrix, lane = column.
__global__ __launch_bounds__(64) void tqlvec_kernel(char *__restrict__ Wk, int64_t sW, int n, int kout,
                                                    const int *__restrict__ mask, const int *__restrict__ fallback) {
    extern __shared__ double lds_x[];
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (fallback[t]) return;
    const int c = threadIdx.x;
    const HeevLay lay = heev_lay(n);
    char *base = Wk + t * sW;
    const double2 *__restrict__ rot = (const double2 *)(base + lay.rot);
    const int *__restrict__ ri = (const int *)(base + lay.ri);
    const int K = *(const int *)(base + lay.nrot);
    const int oc = c < kout ? ((const int *)(base + lay.order))[c] : -1;
    double *X = lds_x + c;  // X[row * 64]
    for (int row = 0; row < n; ++row) X[row * 64] = row == oc ? 1.0 : 0.0;
    // Reverse order, G(i): x_i' = c x_i + s x_{i+1}, x_{i+1}' = -s x_i + c x_{i+1}.
    // Blocks of 64 rotations are fetched one block ahead, lane j holding
    // rotation q0 - j, and broadcast with readlane.  Within a QL sweep the
    // indices ascend by one, so x_{i+1} is carried in a register and
    // x_{i+2} is read one rotation ahead.
    int ci = -2;
    double carry = 0.0, pre = 0.0;
    auto fetch = [&](int q0, double &cx, double &sx, int &ix) {
        const int qj = q0 - c;
        if (qj >= 0) {
            const double2 v = rot[qj];
            cx = v.x;
            sx = v.y;
            ix = ri[qj];
        } else {
            cx = 1.0;
            sx = 0.0;
            ix = 0;
        }
    };
    double cxn, sxn;
    int ixn;
    fetch(K - 1, cxn, sxn, ixn);
    for (int q0 = K - 1; q0 >= 0; q0 -= 64) {
        const double cxb = cxn, sxb = sxn;
        const int ixb = ixn;
        if (q0 - 64 >= 0) fetch(q0 - 64, cxn, sxn, ixn);
        const int nb = q0 + 1 < 64 ? q0 + 1 : 64;
        for (int j = 0; j < nb; ++j) {
            const int i = __builtin_amdgcn_readlane(ixb, j);
            const double cc = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(cxb), j),
                                               __builtin_amdgcn_readlane(__double2loint(cxb), j));
            const double ss = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(sxb), j),
                                               __builtin_amdgcn_readlane(__double2loint(sxb), j));
            double xa, xb;
            if (i == ci) {
                xa = carry;
                xb = pre;
            } else {
                if (ci >= 0) X[ci * 64] = carry;
                xa = X[i * 64];
                xb = X[(i + 1) * 64];
            }
            pre = i + 2 < n ? X[(i + 2) * 64] : 0.0;
            X[i * 64] = cc * xa + ss * xb;
            carry = cc * xb - ss * xa;
            ci = i + 1;
        }
    }
    if (ci >= 0) X[ci * 64] = carry;
    double *Xg = (double *)(base + lay.X);
    for (int row = 0; row < n; ++row) Xg[row * 64 + c] = X[row * 64];
}

// 4. u = H_0 ... H_{n-2} x for the kept columns (n <= 128, kout <= 64).
__global__ __launch_bounds__(256) void ormtr_kernel(char *__restrict__ Wk, int64_t sW, int n, int kout,
                                                    c64 *__restrict__ U, int64_t sU, const int *__restrict__ mask,
                                                    const int *__restrict__ fallback) {
    const int t = blockIdx.x;
    if (mask && !mask[t]) return;
    if (fallback[t]) return;
    const int tid = threadIdx.x, c = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ double2 vb[3][LDS_MAX_N];
    __shared__ double2 part[2][4][64];
    const HeevLay lay = heev_lay(n);
    const char *base = Wk + t * sW;
    const double2 *__restrict__ Vg = (const double2 *)(base + lay.V);
    const double2 *__restrict__ taug = (const double2 *)(base + lay.tau);
    const double *__restrict__ Xg = (const double *)(base + lay.X);
    constexpr int NJ = LDS_MAX_N / 4;
    double2 u[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int r = w + 4 * j;
        u[j] = make_double2(r < n ? Xg[r * 64 + c] : 0.0, 0.0);
    }
    if (n >= 2 && tid < n) vb[(n - 2) % 3][tid] = tid >= n - 1 ? Vg[(int64_t)(n - 2) * n + tid] : make_double2(0, 0);
    __syncthreads();
    for (int k = n - 2; k >= 0; --k) {
        const int cb = k % 3;
        if (k >= 1 && tid < n) vb[(k - 1) % 3][tid] = tid >= k ? Vg[(int64_t)(k - 1) * n + tid] : make_double2(0, 0);
        const double2 tk = taug[k];
        const double2 *v = vb[cb];
        double2 acc = make_double2(0, 0);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int r = w + 4 * j;
            if (r > k && r < n) {
                const double2 z = c_mulc(v[r], u[j]);
                acc.x += z.x;
                acc.y += z.y;
            }
        }
        part[k & 1][w][c] = acc;
        __syncthreads();
        const double2 s = make_double2(part[k & 1][0][c].x + part[k & 1][1][c].x + part[k & 1][2][c].x + part[k & 1][3][c].x,
                                       part[k & 1][0][c].y + part[k & 1][1][c].y + part[k & 1][2][c].y + part[k & 1][3][c].y);
        const double2 f = c_mul(tk, s);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int r = w + 4 * j;
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
            const int r = w + 4 * j;
            if (r < n) Ut[(int64_t)r * n + c] = cmake<double>(u[j].x, u[j].y);
        }
    }
}

int64_t heev_work_bytes(int n) { return n <= LDS_MAX_N ? heev_lay(n).total : 0; }

void heev(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP, void *work,
          int64_t sW, int *fallback, int n, int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
          double thr, int max_bond, const int *mask, int B, hipStream_t s) {
    if (n > 1024) throw HipError("heev: n > 1024");
    int kout = ns < max_bond ? ns : max_bond;
    if (kout > n) kout = n;
    const bool fast = n <= LDS_MAX_N && kout <= 64 && work && fallback;
    const int np = n * (n + 1) / 2;
    size_t shmem = (size_t)n * (4 * sizeof(double) + sizeof(int)) + 16;
    if (n <= LDS_MAX_N) shmem += (size_t)(np + n) * sizeof(double2);
    else if (!Pg) throw HipError("heev: no global scratch for n > 128");
    if (!fast) {
        hipLaunchKernelGGL(heev_lds_kernel, dim3(B), dim3(256), shmem, s, G, sG, U, sU, Z, sZ, Pg, sP, n, ns, evals,
                           perm, sE, keep, rule, thr, max_bond, mask, (const int *)nullptr);
        YQ_HIP(hipGetLastError());
        return;
    }
    char *W = (char *)work;
    const size_t sh1 = (size_t)(np + (4 + TRD_PARTS) * n) * sizeof(double2);
    hipLaunchKernelGGL(hetrd_kernel, dim3(B), dim3(TRD_WAVES * 64), sh1, s, G, sG, W, sW, n, mask);
    YQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(tql_kernel, dim3((B + 63) / 64), dim3(64), (size_t)2 * n * 64 * sizeof(double), s, W, sW, n,
                       ns, evals, perm, sE, keep, rule, thr, max_bond, mask, fallback, B);
    YQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(tqlvec_kernel, dim3(B), dim3(64), (size_t)n * 64 * sizeof(double), s, W, sW, n, kout, mask,
                       (const int *)fallback);
    YQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(ormtr_kernel, dim3(B), dim3(256), 0, s, W, sW, n, kout, U, sU, mask, (const int *)fallback);
    YQ_HIP(hipGetLastError());
    // matrices whose QL exceeded the rotation budget: full single-workgroup solver
    hipLaunchKernelGGL(heev_lds_kernel, dim3(B), dim3(256), shmem, s, G, sG, U, sU, Z, sZ, Pg, sP, n, ns, evals, perm,
                       sE, keep, rule, thr, max_bond, mask, (const int *)fallback);
    YQ_HIP(hipGetLastError());
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
        const int64_t sW = path == 1 ? 0 : heev_work_bytes(n);
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
