// ygpu_internal.h -- device context and shared declarations of the HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/yatscf_gpu.h"
#include "md.h"

// --------------------------------------------------------------------------
// error reporting: every entry point returns a status and leaves a message
void ygpu_set_error(int code, const char *fmt, ...);

#define YG_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ygpu_set_error(YGPU_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                           hipGetErrorString(e_), __FILE__, __LINE__);            \
            return YGPU_ERR_HIP;                                                  \
        }                                                                         \
    } while (0)

#define YG_TRY(call)                 \
    do {                             \
        int s_ = (call);             \
        if (s_ != YGPU_OK) return s_; \
    } while (0)

// Newton-Schulz steps of the orthogonal alignment of a difference-density build (from principal
// angles of up to 60 degrees the error falls below 1e-15 in 9 steps)
#define YG_NS_ITERS 10
// Fock-build timers kept per engine (ygpu_get_fock_timings)
#define YG_MAX_FOCK_HIST 4096
// largest n whose density-build eigenproblem runs on the one-workgroup tridiagonal reduction
#define YG_EIG_NMAX_SMALL 272
// largest n of the blocked two-kernel tridiagonal reduction (two rows per thread of its
// one-workgroup column kernel); rocSOLVER dsyevd beyond
#define YG_SYTRD_BLOCKED_NMAX 2048
// largest max |T z - lambda z| / ||T|| accepted from the small path's inverse iteration; a vector
// beyond it sends the density build to the general eigensolver
#define YG_TRIDIAG_RESID 1e-10
// SP2 purification: step limit, idempotency threshold that starts the two closing steps, least
// squared Cholesky pivot of V_prev^T P V_prev accepted for the occupied vectors
#define YG_SP2_KMAX 96
#define YG_SP2_TOL 1e-9
#define YG_SP2_MINPIV 1e-2

// growable device buffer, freed when released or destroyed (not copyable)
// Device buffer.  Blocks of 1 MiB and more come from and go back to a process-wide cache of
// freed device blocks (per device, best fit between n and 2 n bytes), so the large buffers of one
// SCF are reused by the next one on the same device instead of going back to the driver, whose
// allocations of tens of GB after a free of as much took from 0.03 to 2 s at random.  The cache
// holds at most half the device's memory; ygpu_release_cached_memory returns it to the driver,
// and an allocation that fails empties the device's cache and tries once more.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int dev = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    int ensure(size_t n);
    void release();
    template <class T> T *as() const { return (T *) p; }
    void swap(DBuf &o)
    {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(dev, o.dev);
    }
};

struct BasisDev {
    int nshells = 0, nbf = 0, max_am = 0, max_nprim = 0;
    std::vector<int> am, nprim, prim_off, bf_start;
    std::vector<double> xyz, alpha, coef;
    DBuf d_am, d_nprim, d_prim_off, d_bf_start, d_xyz, d_alpha, d_coef;
    DevShells dev() const {
        return DevShells{d_am.as<int>(), d_nprim.as<int>(), d_prim_off.as<int>(), d_bf_start.as<int>(),
                         d_xyz.as<double>(), d_alpha.as<double>(), d_coef.as<double>()};
    }
};

struct ygpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;   // independent work beside the main stream (the J build)
    bool j_beside = false;        // J build on the side stream, beside the K build
    hipEvent_t fork = nullptr, join = nullptr;
    rocblas_handle rb = nullptr;
    hipEvent_t ev[10] = {};
    DBuf boys;                    // Boys table
    BasisDev bas, aux;
    int natoms = 0;
    DBuf d_charge, d_atom_xyz;

    // sparsity (set by ygpu_set_sparsity)
    int have_sparsity = 0;
    int nbf = 0, naux = 0, naux_pad = 0;
    // spherical auxiliary basis (ygpu_set_aux_spherical): naux counts spherical functions; the
    // integrals are made over the naux_cart Cartesian ones and transformed column by column
    int aux_sph = 0, naux_cart = 0, naux_cart_pad = 0, sph_cmax = 1;
    DBuf d_sph_cs, d_sph_nc, d_sph_coef;
    int n_orth = 0;               // orthogonaliser columns in use (0: all); ygpu_set_orth_dim
    long n_up = 0, nnz = 0;
    std::vector<int> h_upmask;    // nbf*nbf -> up row or -1
    std::vector<int> h_mask;      // reference CSR row of (i,j) or -1
    DBuf d_upmask, d_up_i, d_up_j;
    DBuf d_xptr, d_xup, d_xpartner;  // per bf i: up rows and partners (full CSR order)
    int max_row_len = 0;
    // panel-major operand blocks of the two-part BF16 X build: block i holds L_i = (row length
    // of i rounded up to 32) values per row, in 32-deep k panels, and starts kofs[i] * (rows)
    // elements into its part
    DBuf d_kofs;
    DBuf d_ktiles;                // (i, k0) of the 64-deep k tiles of every basis function's block
    int nktiles = 0;
    DBuf d_xorder;                // basis functions by decreasing row length (X build order)
    long ksum = 0;                // sum of L_i

    // screened 3c work list: per (la,lb) class, list of (lid, rid) shell pairs
    std::vector<int> h_sp_lid, h_sp_rid;
    DBuf d_sp_scr, d_df_scr;
    double scrtol2 = 1e-22;
    int num_uniq_sp = 0;
    std::vector<int> h_uniq_lid, h_uniq_rid;
    std::vector<double> h_sp_scrval, h_df_sp_scrval;

    // DF tensor
    int have_df = 0;
    int metric_mode = YGPU_METRIC_CHOLESKY;   // factor of J_PQ used for B (ygpu_set_metric_factor)
    DBuf d_pqA, d_B, d_Jpq;

    // Fock / density work buffers
    DBuf d_D, d_Cocc, d_H, d_J, d_K, d_F, d_tempJ, d_tempJpart, d_jw, d_X, d_Kpart;
    DBuf d_mat0, d_mat1, d_mat2, d_eig, d_info, d_evE;
    DBuf d_gA, d_gB, d_gC;
    DBuf d_S, d_Xo;               // cached overlap and X = S^{-1/2}
    int have_S = 0, have_H = 0;
    DBuf d_Fin, d_Din, d_T1, d_T2, d_diisR, d_diisF, d_dots;   // DIIS work and history
    DBuf d_solver_ws;             // rocBLAS/rocSOLVER workspace owned by the library
    DBuf d_gemm_ws;               // split-K partials of the small GEMMs (main stream only)
    size_t solver_ws_bytes = 0;

    // mixed-precision Fock builds (ygpu_set_fock_precision)
    int fock_mode = YGPU_FOCK_FP64;
    int have_B32 = 0;             // d_B32 holds the FP32 copy of d_B
    int have_Bt = 0;              // d_Bt holds the transposed two-part BF16 blocks of d_B32
    DBuf d_Bt, d_Cp;              // B and C_occ operand blocks of the two-part X build
    int jinc_valid = 0;           // d_Dprev / d_Jprev hold a density and its J (the last FP64 build or FP32-tensor increment)
    DBuf d_B32, d_Dprev, d_Jprev, d_dD, d_Kpart32, d_Cocc32;
    // difference-density exchange (YGPU_FOCK_INCR): the C_occ and FP64 K of the last FP64 build,
    // the aligned C_occ, the polar factor and its Newton-Schulz work, the FP32 [P | M] block
    int kref_valid = 0, kref_nocc = 0;
    DBuf d_Cref, d_Kref, d_C1, d_SC, d_Qa, d_Qb, d_Qt, d_PM32, d_nserr;
    double incr_align_err = 0.0;  // max |Q^T Q - I| of the last difference build
    // F and D of the last full FP64 build (the reference term of a difference build's energy) and
    // the energy terms of the last build (ygpu_fock_energy)
    DBuf d_Fref, d_Dref, d_epart;
    double energy_dhf = 0.0, energy_ref = 0.0;
    int incr_fallback = 0;        // the last YGPU_FOCK_INCR call fell back to a full FP64 build
    int last_mode = YGPU_FOCK_FP64;   // arithmetic the last Fock build ran

    // own eigensolver paths: reflector scalars, tridiagonal workspace, the small route's copy of
    // F' (for the general eigensolver when an inverse-iteration vector fails its residual bound)
    DBuf d_tau, d_tri, d_Fsave;
    int eig_resid_fallbacks = 0;
    // density by purification (ygpu_set_density_method): the occupied vectors V^T (n_occ x nbf,
    // orthogonal basis) of the last density build, the projector, the SP2 partials and logs, and
    // the work of the orthonormalisation V = P V_prev (V_prev^T P V_prev)^{-1/2}
    int density_method = YGPU_DENSITY_EIGEN;
    int have_vt = 0, vt_nocc = 0;
    int density_last_method = YGPU_DENSITY_EIGEN, density_last_steps = 0, sp2_last_steps = 0;
    DBuf d_VT, d_VT2, d_P, d_Y, d_Mo, d_Mw, d_MT, d_sp2part, d_sp2sum, d_sp2log, d_sp2slog, d_sp2misc;

    // device-resident hand-offs between the calls of one SCF iteration (NULL host arguments):
    // d_F, d_D, d_Cocc hold the last Fock build's F and its inputs (fock_valid); after a
    // density build d_D, d_Cocc hold its outputs (density_valid); d_Fcur holds the last
    // DIIS transform or extrapolation (fcur_valid)
    int fock_valid = 0, density_valid = 0, fcur_valid = 0;
    DBuf d_Fcur;
    // SCF hand-off mode (ygpu_set_scf_resident): host inputs that the device already holds are
    // not uploaded again
    int resident = 0;
    // phase timers: every Fock build since the last ygpu_clear_timings, the last DF build
    std::vector<ygpu_fock_timing_t> fock_hist;
    ygpu_df_timing_t df_last = {};
    double diis_err[YGPU_MAX_DIIS] = {};   // max |R| per history slot, from its update
    int diis_err_ok[YGPU_MAX_DIIS] = {};
};

// --------------------------------------------------------------------------
// launchers implemented in the .hip files
int yg_launch_ovlp_hcore(ygpu_ctx *c, double *dS, double *dH);
int yg_launch_schwarz(ygpu_ctx *c, int quirk_a5, double *d_bf_scr, double *d_df_scr);
int yg_launch_eri2(ygpu_ctx *c, double *dJpq, int ld);
// on stream s; returns once the kernels are queued (the host pair lists are copied before)
int yg_launch_eri3(ygpu_ctx *c, double *dpqA, hipStream_t s);

// C[m][n] (ldc) = alpha * sum_k op(A)(m,k) op(B)(k,n) + beta * C, row-major, FP64 MFMA; with a
// workspace ws, products of fewer than wg_target / 4 output tiles split k over about wg_target
// workgroups (partials in ws, added in a fixed order; C must not overlap A, B or ws)
int yg_dgemm(hipStream_t s, int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
             const double *B, int ldb, double beta, double *C, int ldc, DBuf *ws = nullptr, int wg_target = 256);
// C [m][n] = A [m][k] B [n][k]^T with k-contiguous rows (k, lda, ldb multiples of 8), 128x128 tiles;
// lower_b: B lower triangular (its zeros above the diagonal are skipped, half the flops at n = k)
int yg_dgemm_nt128(hipStream_t s, int m, int n, long k, const double *A, long lda, const double *B, long ldb,
                   double *C, long ldc, bool lower_b = false);
// Cholesky A = L L^T (lower, row-major, order n padded to npad, a multiple of 64, with the
// padding rows / columns zero on entry) and W = L^{-1}: L in the lower triangle of A, W (npad x
// npad, leading dimension ldw, zero above the diagonal, identity on the padding); a non-positive
// pivot sets *info (device, zero on entry) to its column + 1; T holds yg_chol_inv_ws(npad)
// doubles of workspace, ws the split-K partials of the inverse's products.  lda and ldw
// multiples of 8
size_t yg_chol_inv_ws(int npad);
int yg_chol_inv_lower(hipStream_t s, int n, int npad, long lda, double *A, double *W, long ldw, int *info, double *T,
                      DBuf *ws);
// X[i] = Cocc(partners of row i)^T * Bup(rows of i) for every basis function i
int yg_xbuild(hipStream_t s, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup,
              const int *xpartner, const double *Cocc, const double *B, double *X, int max_row_len);
// upper-triangle K = X X^T with split-K partials in Kpart, result in K (full, mirrored)
int yg_syrk_upper(hipStream_t s, int nbf, long kdim, const double *X, double *K, DBuf &part);
// the split of yg_syrk_upper: returns the number of k chunks (kchunk_out: their length); the
// partials take nsplit * nbf * nbf doubles
int yg_syrk_upper_split(int nbf, long kdim, long *kchunk_out);
// yg_syrk_upper with the partials in a caller-provided buffer
int yg_syrk_upper_ws(hipStream_t s, int nbf, long kdim, const double *X, double *K, double *part, long kchunk);
// energy terms of a Fock build: yg_energy_blocks() pairs of block partials in part (see ygpu_fock.hip);
// Dr, Fr NULL: no reference term
int yg_energy_blocks();
int yg_energy_terms(hipStream_t s, int nbf, const double *D, const double *H, const double *F, const double *Dr,
                    const double *Fr, double *part);
// the split-K partials of the upper tiles of X X^T only (nbf x nbf each, no reduction)
int yg_syrk_upper_parts(hipStream_t s, int nbf, long kdim, const double *X, double *part, long kchunk);
// SP2 purification (ygpu_purify.hip): X (n x ldx, ldx a multiple of 16) = X_0 from F (rows and
// columns < n_orth), bounds[0..1] = the Gershgorin interval; wsum holds yg_sp2_wsum_doubles(n)
int yg_sp2_begin(hipStream_t s, int n, int n_orth, const double *F, long ldf, double *X, long ldx, double *wsum,
                 double *bounds);
size_t yg_sp2_wsum_doubles(int n);
// steps first .. first + count - 1 (each a SYRK of X into part, ldx / kchunk splits, then the
// choice); elog[k] = tr(X_k) - tr(X_k^2), slog[k] = 1 if step k squared
int yg_sp2_steps(hipStream_t s, int n, long ldx, double *X, int N, double *part, long kchunk, double *wsum,
                 int first, int count, double *elog, int *slog);
// out[0] = min_i L[i][i]^2 over i < n
int yg_min_diag_sq(hipStream_t s, int n, long ld, const double *L, double *out);
// J from density dD (upper triangle of dJ; the rest zero) on stream s, timed by ev[0..2];
// B32 != nullptr reads the FP32 copy
int yg_jbuild(ygpu_ctx *c, hipStream_t s, const double *dD, double *dJ, const float *B32, float *t_tempJ_ms,
              float *t_Jmat_ms);
// row groups of the J build's first pass (its partials take ngroup * naux_pad doubles)
long yg_jbuild_groups(long n_up, long naux_pad, bool fp32);
// yg_jbuild with the pass-1 partials, temp_J and the n_up pair weights in caller-provided
// buffers (nothing allocated)
int yg_jbuild_ws(ygpu_ctx *c, hipStream_t s, const double *dD, double *dJ, const float *B32, double *part,
                 double *tempJ, double *wts, long ngroup);
int yg_fill(hipStream_t s, long n, double *x, double v);                          // x[0..n) = v
// takes the device address of the no-allocation entry point's static scratch array on `device`
// (which must be current); once per device, later calls return at once
int yg_scratch_init(int device);
int yg_delta(hipStream_t s, long n, const double *D, double *Dprev, double *dD);   // dD = D - Dprev; Dprev = D
int yg_symmetrize(hipStream_t s, int n, double *A, long lda);                      // A = (A + A^T) / 2
// tridiagonal reduction of a row-major n x n symmetric matrix (upper triangle read) on one
// workgroup (n <= YG_EIG_NMAX_SMALL); reflectors left in the rows of A
int yg_sytrd_small(hipStream_t s, int n, double *A, double *d, double *e, double *tau);
// blocked tridiagonal reduction of a row-major symmetric n x n matrix (lower triangle read,
// 3 <= n <= YG_SYTRD_BLOCKED_NMAX, leading dimension lda) in LAPACK dsytrd's lower layout of
// the column-major view: d, e, tau, reflectors below the subdiagonal; ws holds
// yg_sytrd_blocked_ws(n) bytes
size_t yg_sytrd_blocked_ws(int n);
int yg_sytrd_blocked(hipStream_t s, int n, long lda, double *A, double *d, double *e, double *tau, double *ws);
// Z <- Q Z for the first nvec columns of Z (column-major, leading dimension ldz) with the
// reflectors yg_sytrd_blocked (or yg_sytrd_small, with lda = n) leaves in A
// (1 <= n <= YG_SYTRD_BLOCKED_NMAX)
int yg_backtransform_cols(hipStream_t s, int n, int nvec, const double *A, long lda, const double *tau, double *Z,
                          long ldz);
// all eigenvalues w (ascending) of the symmetric tridiagonal tridiag(e, d, e) and the
// eigenvectors of the nvec lowest as the rows of Z (nvec x n); ws holds
// n ((nvec + 63) / 64 * 64) doubles; *info (device, zero on entry) = j + 1 for the first vector j
// whose residual exceeds YG_TRIDIAG_RESID ||T||
int yg_tridiag_eig(hipStream_t s, int n, int nvec, const double *d, const double *e, double *w, double *Z,
                   double *ws, int *info);
int yg_accum(hipStream_t s, long n, double *J, double *Jprev);                     // J += Jprev; Jprev = J
int yg_maxabs(hipStream_t s, long n, const double *x, double *out);                // out[0] = max |x|
// mixed precision (BF16-split MFMA, FP32 accumulation)
int yg_to_f32(hipStream_t s, const double *in, float *out, long n);
// X (nbf x n_occ naux_pad) is stored pre-split into nparts (2 or 3) BF16 arrays, one after the
// other (part q of element e at q * nbf * n_occ * naux_pad + e); the X build writes it, the
// SYRK reads it
// dst[q][kofs_i R + (kk / 32) 32 R + 32 r + kk % 32] = BF16 part q (of 2) of src[idx[xptr_i + kk]][r],
// r < R, kk < L_i (zero for kk >= row length): the gathered rows of every basis function i
// transposed into 32-deep k panels
// spherical option: T of degree l (ncart(l) x (2l+1), row-major); the aux column tables and
// naux (spherical) from c->aux; out = in T over the aux columns; A[k][k] += shift, k >= n_orth
int yg_sph_coefficients(int l, double *T);
int yg_aux_sph_setup(ygpu_ctx *c);
int yg_aux_sph_cols(ygpu_ctx *c, hipStream_t s, long rows, const double *in, long ldin, double *out, long ldout);
int yg_shift_tail_diag(hipStream_t s, int n, int n_orth, double *A, double shift);
// the panel-major two-part blocks of src's gathered rows; ktiles: the (i, k0) 64-deep k tiles of every
// basis function (ygpu_ctx::d_ktiles)
int yg_gather_split_t(hipStream_t s, int nktiles, const int2 *ktiles, const float *src, long ld, int R,
                      const int *idx, const int *xptr, const int *kofs, void *dst, size_t dpart);
// X (two BF16 parts, nrow rows per basis function) from the panel-major blocks Cp (nrow rows)
// and Bt (naux_pad rows) made by yg_gather_split_t; order: the basis functions, longest row first
int yg_xbuild_ps(hipStream_t s, int nbf, int nrow, int naux_pad, const int *xptr, const int *kofs, const int *order,
                 const void *Cp, size_t cpart, const void *Bt, size_t btpart, void *X);
int yg_xbuild_bf(hipStream_t s, int nparts, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup,
                  const int *xpartner, const float *Cocc32, const float *B32, void *X);
int yg_syrk_bf(hipStream_t s, int nparts, int nbf, long kdim, const void *X, double *K, DBuf &part);
int yg_fock_assemble(hipStream_t s, int nbf, const double *H, double *J, const double *K, double *F);
// difference-density exchange: K = Kref + (G + G^T) / 2 with G = X^M (X^P)^T, X the pre-split X
// build of the combined [P | M] occupied block (2 n_occ columns per basis function)
int yg_incr_k_bf(hipStream_t s, int nparts, int nbf, int n_occ, int naux_pad, const void *X, const double *Kref,
                 double *K, DBuf &part);
int yg_ns_shift(hipStream_t s, int n, double *T);                                    // T = 3 I - T
int yg_dev_identity(hipStream_t s, int n, const double *T, double *out);             // out[0] = max |T - I|
// PM = [C1 + C0 | C1 - C0] in FP32 (nbf x 2 n_occ)
int yg_pm32(hipStream_t s, int nbf, int n_occ, const double *C1, const double *C0, float *PM);
int yg_scale_rows(hipStream_t s, int n, int ncol, double *A, int lda, const double *w, int mode);
int yg_gather_cols(hipStream_t s, int nbf, int n_occ, const double *C, const int *idx, double *Cocc);
int yg_transpose(hipStream_t s, int rows, int cols, const double *A, double *AT);
int yg_antisym(hipStream_t s, int n, const double *A, double *out);                  // out = A - A^T
// partial <x, y_j> per chunk into part[j][chunk]; with pmax, the chunk maxima of |x| into pmax[chunk]
int yg_dots(hipStream_t s, long len, const double *x, const double *ys, long ystride, int ny, double *part,
            int nchunk, double *pmax = nullptr);
int yg_lincomb(hipStream_t s, long len, const double *ys, long ystride, int ny, const double *coef, double *out);
