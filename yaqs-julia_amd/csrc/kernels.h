// Host-side launchers of the TJM engine kernels (kernels.hip, gemm_mfma.hip).
#pragma once

#include "common.h"

namespace yq {

template <typename TA, typename TB, typename TC, typename Tacc>
void gemm(const GemmDesc &d, const cplx<TA> *A, const cplx<TB> *B, cplx<TC> *C, hipStream_t s);

// MFMA GEMM for op(A) in {N, C}, op(B) = N, same precision in and out
// (float: v_mfma_f32_16x16x4f32, double: v_mfma_f64_16x16x4f64).  Returns
// false when the descriptor is outside what the kernel covers.  tile picks
// the kernel variant (all give bit-identical results): 0-3 the C tile per
// workgroup (GEMM_TILE_BM/BN) with register-staged k tiles, for op(A) in
// {N, T, C} with op(B) = N (other ops take tile 0); 4 and 5 the 64 x 64
// tile with k tiles loaded straight into LDS, three stages of 16 k or four
// of 8 k (complex-FP32, op N operands, whole tiles, 16-byte aligned
// operands; tile 0 otherwise).
constexpr int GEMM_TILES = 6;
constexpr int GEMM_TILE_BM[GEMM_TILES] = {64, 32, 64, 32, 64, 64};
constexpr int GEMM_TILE_BN[GEMM_TILES] = {64, 64, 32, 32, 64, 64};
template <typename T>
bool gemm_mfma(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, cplx<T> *C, hipStream_t s, int tile = 0);
// FP64 Gram matrix of complex-FP32 operands on the FP64 matrix cores
// (op pairs (N, C) and (C, N); honours d.lower); false if not launched.
bool gemm_mfma_gram(const GemmDesc &d, const c32 *A, const c32 *B, c64 *C, hipStream_t s);
// True when tile 4 (k tiles straight into LDS) applies to the descriptor.
template <typename T>
bool gemm_dl_fits(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B);

// T1 = A B (both op N) with rows (p, m < Dl) and columns (s < dph, q < Q),
// followed in the same kernel by the MPO step into Y (layout of wstep's
// output).  Returns false when the shape is outside the fused kernel.  dl:
// k tiles loaded straight into LDS where gemm_w_dl_fits (same bits).
bool gemm_w_covers(const GemmDesc &d, int P, int Dl, int dph, int Dr, int Q, int nnz);
template <typename T>
bool gemm_w_dl_fits(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, int Dl);
template <typename T>
bool gemm_w_mfma(const GemmDesc &d, const cplx<T> *A, const cplx<T> *B, int P, int Dl, int dph, int Dr, int Q,
                 cplx<T> *Y, int64_t sY, const int *rowptr, const int *col, const c64 *val, int nnz, hipStream_t s,
                 bool dl = false);

template <typename T>
void wstep(const cplx<T> *X, cplx<T> *Y, int P, int I, int J, int Q, int64_t sX, int64_t sY, const int *rowptr,
           const int *col, const c64 *val, const int *active, int B, hipStream_t s);

template <typename T>
void lz_init(const cplx<T> *x, int64_t sx, cplx<T> *V, int64_t sV, int64_t n, double *beta0, int *active, int *nit,
             int *nactive, int nslots, int B, hipStream_t s);
template <typename T>
void lz_dots(const cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j, c64 *part, int kslots, const int *active,
             int B, hipStream_t s);
// Deferred normalisation: the Krylov basis is stored unnormalised, V_q =
// v_q / s_q with s_0 = 1 and s_q = 1 / beta[q]; lz_update scales the matvec
// output by s_j and takes the coefficients of the stored vectors.
template <typename T>
void lz_update(cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, int j, const c64 *part, int kslots, double *partn,
               double *galpha, const int *active, const double *beta, int kmax, int B, hipStream_t s);
// β = ‖w‖ (stored in beta[j+1] when the trajectory goes on) and the
// stopping rule; trajectories that go on are counted in nactive[j] (cleared
// by lz_init for every slot).  With `host_count` (a device pointer to a
// coherent pinned host int) the last workgroup to finish also stores the
// final nactive[j] there; `fin_done` is a device 64-bit word that is zero
// between launches (the last workgroup resets it).
void lz_finalize(int j, int kmax, double dt, double tol, const double *partn, int64_t n, const double *galpha,
                 double *alpha, double *beta, const double *beta0, c64 *coef, int *nit, int *active, int *nactive, int B,
                 unsigned long long *fin_done, int *host_count, hipStream_t s);
template <typename T>
void lz_combine(cplx<T> *x, int64_t sx, const cplx<T> *V, int64_t sV, int64_t nmax, int64_t n, const double *beta0,
                const c64 *coef, const int *nit, const double *beta, int kmax, const int *mask, int B, hipStream_t s);
constexpr int LZ_CHUNK_ELEMS = 2048;

// Hermitian eigen-decomposition of a batch of FP64 Gram matrices (eigen.hip):
// sorted eigenvalues, kept rank under rule (0: T-A, 1: T-C, 2: T-D) and the
// eigenvectors of the kept rank as the first columns of U (row stride n).
// work / sW: per-matrix work area of heev_work_bytes(n, kmax) bytes (kmax:
// most kept columns of any call; null when that is 0), fallback: B ints;
// Z / sZ: n*n + 2n + 4 doubles per matrix; Pg / sP: n(n+1)/2 double2 per
// matrix when n > 128.
int64_t heev_work_bytes(int n, int kmax);
// Returns true when the multi-kernel path ran; with fixup = false the caller
// then runs heev_fixup for the matrices flagged in `fallback` itself (for
// instance only when its readback shows one).
// True when heev reads only the lower triangle of its n x n matrices (the
// wide path up to n = 256 and the single-workgroup solver): a Gram matrix
// for it may skip the tiles above the diagonal.
bool heev_gram_lower_only(int n);
bool heev(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP, void *work,
          int64_t sW, int *fallback, int n, int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
          double thr, int max_bond, const int *mask, int B, hipStream_t s, bool fixup = true);
// The single-workgroup solver for the matrices with fallback[t] != 0 (all
// matrices when fallback is null).
void heev_fixup(const c64 *G, int64_t sG, c64 *U, int64_t sU, double *Z, int64_t sZ, double2 *Pg, int64_t sP,
                const int *fallback, int n, int ns, double *evals, int *perm, int64_t sE, int *keep, int rule,
                double thr, int max_bond, const int *mask, int B, hipStream_t s);
template <typename T>
void write_iso(const c64 *U, int64_t sU, const int *perm, int64_t sE, const int *keep, int n, int kw, int mode,
               cplx<T> *Out, int64_t sO, const int *mask, int B, hipStream_t s);

// CGS2 orthonormalisation of nvec strided vectors (kernels.hip); S / sS:
// scratch of cgs2_scratch_elems(nvec, len) elements per trajectory, fro: B
// doubles.
int64_t cgs2_scratch_elems(int nvec, int len);
template <typename T>
void cgs2(cplx<T> *A, int64_t sA, int nvec, int len, int64_t vs, int64_t es, cplx<T> *Cf, int64_t sC, cplx<T> *S,
          int64_t sS, double *fro, const int *mask, int B, hipStream_t s);

// ρ1 of sites a and a+1 (L x 2 x 2 per trajectory, zero elsewhere) as partial
// traces of the pair's ρ2 (4 x 4 per trajectory).
void ptrace_pair(const c64 *rho2, int64_t s2, int a, c64 *rho1, int64_t s1, int L, int B, hipStream_t s);
template <typename T>
void local_op(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *ops, const int *op_idx, const int *mask, int B,
              hipStream_t s);
template <typename T>
void local_op4(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *M, int B, hipStream_t s);
// local_op4 with the matrix of trajectory t at Mtab + 16 op_idx[t]
// (op_idx[t] < 0: slot untouched).
template <typename T>
void local_op4_sel(cplx<T> *A, int64_t sA, int dl, int dr, const c64 *Mtab, const int *op_idx, int B, hipStream_t s);
template <typename T>
void norm2(const cplx<T> *A, int64_t sA, int64_t n, double *out, const int *mask, int B, hipStream_t s);
template <typename T>
void scale(cplx<T> *A, int64_t sA, int64_t n, const double *f, const int *mask, int B, hipStream_t s);
template <typename TS, typename TD>
void copy2d(const cplx<TS> *S, int64_t sS, int64_t lds, cplx<TD> *D, int64_t sD, int64_t ldd, int rows, int cols,
            const int *mask, int B, hipStream_t s);
template <typename T>
void fill_zero(cplx<T> *A, int64_t sA, int64_t n, int B, hipStream_t s);
// D[r][c] = S[r][c] (S[c][r] with tr) for r, c < n, zero in columns n .. ldd-1
template <typename T>
void copy_bond(const cplx<T> *S, int64_t sS, int lds, bool tr, cplx<T> *D, int64_t sD, int ldd, int n, const int *mask,
               int B, hipStream_t s);
template <typename T>
void zero_col(cplx<T> *Y, int64_t sY, int rows, int ld, int col, const int *mask, int B, hipStream_t s);
template <typename T>
void identity_env(cplx<T> *E, int64_t sE, int dim, int D, int right, int B, hipStream_t s);
template <typename T>
void rho(const c64 *X, int64_t sX, const cplx<T> *A, int64_t sA, int dl, int d, int dr, c64 *out, int64_t sR,
         const int *mask, int B, hipStream_t s);

void jump_decide(const double *norm2v, uint64_t seed, uint64_t traj0, uint64_t *rng_k, int has_procs, int *state,
                 double *fac, int B, hipStream_t s);
// cand_kind[c]: 0 = Tr(L†L ρ1) at cand_site, 1 = Tr(L†L ρ2) of the pair,
// 2 = the per-trajectory factor lrnu[t * ncand + c] (long-range MPO jumps)
void jump_choose(const c64 *rhov, int64_t sR, int L, int ncand, const int *cand_site, const c64 *cand_LdL,
                 const double *cand_gamma, double dt, uint64_t seed, uint64_t traj0, uint64_t *rng_k, int *state,
                 int *chosen, double *wbuf, int B, hipStream_t s, const c64 *rho2 = nullptr, int64_t sR2 = 0,
                 const c64 *cand_LdL4 = nullptr, const int *cand_kind = nullptr, const double *lrnu = nullptr);
// mask[t] = (state[t] == 1 && chosen[t] == c)
void mask_chosen(const int *state, const int *chosen, int c, int *mask, int B, hipStream_t s);
// ν of a general long-range candidate c for the jumping trajectories
// (state 1): with n = Tr ρ of site nsite (rho, stride sR; a site whose ρ is
// the true state's),
// cs = <σ_a τ_b> (val[t]) and cn = <σ†σ_a τ†τ_b> (val[B + t]),
// ‖Oψ‖² = |α|² n + |β|² cn + 2 Re(conj(α) β cs); ν = 1 unless that is at
// most 1e-12 (|α|² n + |β|² cn) (O annihilates ψ), then 0.
void lr_nu(const c64 *val, const c64 *rho, int64_t sR, int nsite, double ar, double ai, double br, double bi, int c,
           int ncand, const int *state, double *lrnu, int B, hipStream_t s);
// Blocks of O ψ for O = α I + β σ_a τ_b on the bond-doubled layout (site
// tensor [l][s][r] with ncol columns, its new layout; masked slots):
// mode 0 (site a): columns R + r <- σ applied to columns r < R, rows l < nl;
// mode 1 (between): rows nl + l, columns R + r <- rows l, columns r (l < nl, r < R);
// mode 2 (site b): rows nl + l <- β τ applied to rows l, rows l <- α rows l
// (l < nl, r < R).  op: σ or τ, 2x2 engine order.
template <typename T>
void mpo_block(cplx<T> *A, int64_t sA, int ncol, int mode, int nl, int R, const c64 *op, double ar, double ai,
               double br, double bi, const int *mask, int B, hipStream_t s);
// Masks of trajectories whose chosen jump is a merged-pair candidate on pair
// (p, p+1) (mode 0, opidx = its operator's index or -1) or beyond p (mode 1).
void jump_pair_mask(const int *state, const int *chosen, const int *cand_site, const int *cand_pair, int p, int mode,
                    int *mask, int *opidx, int B, hipStream_t s);
void jump_opidx(const int *state, const int *chosen, const int *cand_site, const int *cand_op, const int *cand_site2,
                const int *cand_op2, int site, int *opidx, int B, hipStream_t s);
void mask_from_state(const int *state, int v, int *mask, int *count, int B, hipStream_t s);
void observe(const c64 *rho1, int64_t sR1, const c64 *rho2, int64_t sR2, int nobs, const int *obs_site,
             const int *obs_n, const c64 *obs_op, int col, int ncol, int64_t ntraj, int64_t t0, double *out, int B,
             hipStream_t s);
void set_dims(int *dims, int ld, int b, const int *keep, int val, const int *mask, int B, hipStream_t s);
void fill_int(int *p, int v, int n, hipStream_t s);
// NaN/Inf guard before a bond split (src/Algorithms.jl:414-417): a Gram
// matrix G (n x n per trajectory, stride sG) whose diagonal is not finite
// comes from a two-site tensor with a NaN/Inf entry; its trajectory is
// flagged (fail[t] = 1) and the matrix replaced by e0 e0^T, so the eigensolver only ever
// sees finite input.
void guard_gram(c64 *G, int64_t sG, int n, int *fail, const int *mask, int B, hipStream_t s);
// Split of a Gram matrix with kept eigenvalues below tau * lambda_max: per
// trajectory the count above the cut (dcount), a flag when kept ones fall
// below it (at least one above), and the number of flagged trajectories
// added to *count.
void tail_flags(const double *evals, int64_t sE, const int *keep, double tau, int *dcount, int *flag, int *count, int B,
                hipStream_t s);
// D = A - C on n elements per slot (masked).
template <typename T>
void sub(const cplx<T> *A, int64_t sA, const cplx<T> *C, int64_t sC, cplx<T> *D, int64_t sD, int64_t n, const int *mask,
         int B, hipStream_t s);
// Columns (mode 0) / conjugated rows (mode 1) k in [dcount, keep) of an
// isometry written by write_iso, from the leading eigenvectors of U.
template <typename T>
void tail_fix(cplx<T> *Out, int64_t sO, int n, int kw, int mode, const int *dcount, const int *keep, const c64 *U,
              int64_t sU, const int *perm, int64_t sE, const int *mask, int B, hipStream_t s);
// Test hook: stores NaN into element 0 of slot t of A (t < B).
template <typename T>
void poison_slot(cplx<T> *A, int64_t sA, int t, hipStream_t s);

}  // namespace yq
