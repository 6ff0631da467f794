/*
 * yatscf_gpu.h -- C-ABI of the MI355X (gfx950) DF-RHF hot path.
 *
 * The library libyatscf_gpu.so owns all device state of one TinySCF engine
 * behind an opaque handle.  Host code (the drop-in build_DF_tensor.c,
 * build_Fock.c, build_density.c, DIIS.c and TinySCF.c of
 * yatscf-df_amd/host/) keeps the reference's TinySCF struct and calls these
 * functions with plain host pointers.  All matrices are row-major FP64, as in
 * the reference (utils.h:6-7, TinySCF.h:62-77).  Every function is
 * synchronous at return (host outputs are valid) and returns YGPU_OK or an
 * error code; ygpu_last_error() gives the message of the calling thread's
 * last failure.  A handle is not re-entrant: calls on one handle must not
 * overlap (the reference's engine is not re-entrant either, TinySCF.h:12).
 */
#ifndef YATSCF_GPU_H
#define YATSCF_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YGPU_OK            0
#define YGPU_ERR_INVALID   1   /* bad argument or call order */
#define YGPU_ERR_HIP       2   /* a HIP runtime call failed */
#define YGPU_ERR_NOMEM     3   /* device allocation failed */
#define YGPU_ERR_SOLVER    4   /* rocSOLVER eigensolver failed */
#define YGPU_ERR_NUMERIC   5   /* non-positive metric eigenvalue, NaN */
#define YGPU_ERR_NODEVICE  6   /* no HIP device */

typedef struct ygpu_ctx ygpu_ctx_t;

/* Phase timers in seconds, named after the reference's printouts
 * (build_DF_tensor.c:263-284, build_Fock.c:390-395). */
typedef struct {
    double t_3c, t_2c, t_invsqrt, t_dfgemm;
} ygpu_df_timing_t;

typedef struct {
    double t_tempJ, t_Jmat, t_tempK, t_Kmat, t_symm, t_total;
} ygpu_fock_timing_t;

/* Message of the calling thread's last error ("" if none). */
const char *ygpu_last_error(void);

/* Failure inside a drop-in TinySCF function.  The reference's hot-path
 * functions return void and abort on failure (e.g. the asserts of
 * build_DF_tensor.c:67, :214); the drop-in build_DF_tensor.c, build_Fock.c and
 * build_density.c call ygpu_fail(status, their name) when a call of this
 * library fails.  ygpu_fail passes (status, where, ygpu_last_error()) to the
 * calling thread's handler if one is installed (the handler may leave by
 * longjmp: the host engine's Python bindings do, to turn the failure into an
 * error status), else, or if the handler returns, prints the message and
 * aborts.  ygpu_set_failure_handler(NULL) restores the abort. */
typedef void (*ygpu_failure_handler_t)(int status, const char *where, const char *message);
int ygpu_set_failure_handler(ygpu_failure_handler_t handler);
void ygpu_fail(int status, const char *where);

/* Number of visible HIP devices. */
int ygpu_device_count(int *count);

/* Create / destroy the device engine on `device` (replaces the device-free
 * state set up by init_TinySCF, TinySCF_init_free.c:14-208, and freed by
 * free_TinySCF, TinySCF_init_free.c:211-268). */
int ygpu_create(int device, ygpu_ctx_t **ctx);
int ygpu_destroy(ygpu_ctx_t *ctx);

/* No counterpart in the reference (its buffers are host memory).  Device
 * blocks of 1 MiB and more that a destroyed engine (or a regrown buffer)
 * gives up stay in a process-wide cache and are handed to later engines on
 * the same device (best fit, n to 2 n bytes), so repeated solves do not pay
 * the driver's allocation of tens of GB each time.  The cache holds at most
 * half the device's memory (the largest blocks leave first), and an
 * allocation that fails empties it and tries again.  This returns every
 * cached block to the driver (call it with no engine in use on another
 * thread). */
int ygpu_release_cached_memory(void);

/* Page-locked host memory for the matrices the drop-in exchanges with the
 * device on every SCF iteration (F, D, J, K, C_occ, eigenvalues, X, Hcore, S)
 * and for J^{-1/2}: the reference allocates them with ALIGN64B_MALLOC
 * (TinySCF_init_free.c:131-143, :194) and frees them with ALIGN64B_FREE
 * (:232-246); page-locked, the per-iteration copies run at DMA rate instead of
 * through the runtime's staging buffers.  64-byte aligned or better.  Free
 * with ygpu_host_free only; ygpu_host_free(NULL) is a no-op. */
int ygpu_host_alloc(size_t bytes, void **ptr);
int ygpu_host_free(void *ptr);

/* Shells of the orbital basis (which = 0) or of the density-fitting basis
 * (which = 1), in the reference's shell order (CMS_parse_molecule,
 * CMS_BasisSet.c:192-299) with normalised contraction coefficients (the
 * Simint shells of CMS_createSimint, CMS_Simint.c:79-145).  xyz is 3*nshells
 * Bohr; alpha/coef are indexed by prim_off[shell]. */
int ygpu_set_basis(ygpu_ctx_t *ctx, int which, int nshells, const int *am, const int *nprim,
                   const int *prim_off, const double *xyz, const double *alpha, const double *coef);

/* Spherical-harmonic basis option.  The reference reads a "spherical" flag
 * from the first line of a .gbs file (CMS_BasisSet.c:464-483) and counts 2l+1
 * functions per shell (:277-281), with no integral code for such shells.
 * Here a spherical basis is the span of the real solid harmonics inside the
 * Cartesian one; integrals stay Cartesian.
 *
 * ygpu_sph_coefficients: T of degree l (0..4), (l+1)(l+2)/2 rows (Cartesian
 * components x^a y^b z^c, a = l..0, b = l-a..0, normalised like x^l) by 2l+1
 * columns (real solid harmonics, m = -l..l), row-major, each column of unit
 * norm in the metric of those components; the identity for l <= 1.  Needs no
 * context and no device.
 *
 * ygpu_set_aux_spherical: the density-fitting basis is spherical (on != 0):
 * the columns of (mn|P) and both sides of J_PQ are transformed by T before
 * the metric factor, and every tensor and matrix that counts auxiliary
 * functions (B, Jpq, the packed layout) counts the spherical ones.  Call
 * after ygpu_set_basis(which = 1) and before ygpu_set_sparsity (the layout is
 * cleared).
 *
 * ygpu_set_orth_dim: the orthogonaliser X passed to the density build and the
 * DIIS calls uses only its first n_orth columns (the rest zero), as X =
 * T (T^T S T)^{-1/2} of a spherical orbital basis does; F' = X^T F X then has
 * zero rows and columns beyond n_orth, and the density build adds
 * YGPU_ORTH_SHIFT (Eh) to their diagonal so that those directions, which hold
 * no orbital, are never occupied (their eigenvalues read YGPU_ORTH_SHIFT).
 * n_orth = 0 or nbf: every column (the default). */
#define YGPU_ORTH_SHIFT 1.0e4
int ygpu_sph_coefficients(int l, double *T);
int ygpu_set_aux_spherical(ygpu_ctx_t *ctx, int on);
int ygpu_set_orth_dim(ygpu_ctx_t *ctx, int n_orth);

/* Nuclear charges and positions (Bohr) for the core Hamiltonian. */
int ygpu_set_atoms(ygpu_ctx_t *ctx, int natoms, const double *charge, const double *xyz);

/* S and Hcore = T + V, nbf x nbf (replaces the Simint calls of
 * TinySCF_compute_Hcore_Ovlp_mat, TinySCF.c:20-54). */
int ygpu_ovlp_hcore(ygpu_ctx_t *ctx, double *S, double *Hcore);

/* X = A^{-1/2} = U diag(lambda^{-1/2}) U^T of a symmetric positive definite
 * n x n matrix, FP64 (TinySCF.c:57-76; build_DF_tensor.c:205-237).  min_eig
 * (nullable) receives the smallest eigenvalue; a non-positive one is
 * YGPU_ERR_NUMERIC. */
int ygpu_sym_inv_sqrt(ygpu_ctx_t *ctx, int n, const double *A, double *X, double *min_eig);

/* Schwarz screening values (replaces the Simint calls of
 * TinySCF_compute_sq_Schwarz_scrvals, TinySCF.c:112-177):
 *   bf_pair_scrval[i*nbf+j] = |(MN|MN) element| for i in shell M, j in N,
 *     read at the reference's index when quirk_a5 != 0 (TinySCF.c:153),
 *     at the diagonal (ij|ij) otherwise;
 *   sp_scrval[M*nshells+N] = block maximum;
 *   df_sp_scrval[P] = max_k (P_k|P_k).  Any output may be NULL. */
int ygpu_schwarz(ygpu_ctx_t *ctx, int quirk_a5, double *sp_scrval, double *bf_pair_scrval,
                 double *df_sp_scrval);

/* The screened pair layout (TinySCF_prepare_sparsity, TinySCF.c:228-280)
 * and the unique shell-pair list (TinySCF.c:183-221).  A basis-function pair
 * is kept when either orientation has bf_pair_mask >= 0: the reference's
 * Schwarz index (TinySCF.c:153) can keep (i, j) and drop (j, i), a case its
 * integral copy leaves undefined (a write to row -1, build_DF_tensor.c:36-43).
 * Diagonal pairs must be kept.  The device stores only rows with i <= j ("up"
 * rows), in the reference's CSR order restricted to j >= i. */
int ygpu_set_sparsity(ygpu_ctx_t *ctx, const int *bf_pair_mask, int num_uniq_sp,
                      const int *uniq_sp_lid, const int *uniq_sp_rid, const double *sp_scrval,
                      const double *df_sp_scrval, double shell_scrtol2);

/* Factor W of the Coulomb metric (W^T W = J_PQ^{-1}) that forms the DF tensor
 * B = (mn|P) W^T.  The reference takes W = J^{-1/2} from an eigendecomposition
 * of J_PQ (calc_inverse_sqrt_Jpq, build_DF_tensor.c:205-237):
 * YGPU_METRIC_INVSQRT.  YGPU_METRIC_CHOLESKY (the default) takes W = L^{-1}
 * with J_PQ = L L^T: the fitted integrals B B^T = (mn|P) J^{-1} (Q|kl), hence
 * J, K and F, are the same up to rounding, the factor costs a Cholesky
 * instead of an eigensolve and, W being lower triangular, the contraction
 * half the flops; B's columns and the Jpq output below differ (L^{-1} instead
 * of J^{-1/2}).  A metric that is not positive definite is YGPU_ERR_NUMERIC
 * in both modes. */
#define YGPU_METRIC_CHOLESKY 0
#define YGPU_METRIC_INVSQRT 1
int ygpu_set_metric_factor(ygpu_ctx_t *ctx, int mode);

/* The DF tensor B = (mn|P) W^T on the device (TinySCF_build_DF_tensor,
 * build_DF_tensor.c:256-287), W as set by ygpu_set_metric_factor.
 * Jpq_invsqrt (nullable, naux x naux) receives W (J^{-1/2} as the reference's
 * Jpq ends, :236, in the YGPU_METRIC_INVSQRT mode; L^{-1} in the Cholesky
 * mode) and df_tensor (nullable, the reference's host layout nnz x naux with
 * both triangles, TinySCF.c:268-273) a host copy of B. */
int ygpu_build_df_tensor(ygpu_ctx_t *ctx, double *Jpq_invsqrt, double *df_tensor,
                         ygpu_df_timing_t *timing);

/* As ygpu_build_df_tensor, also returning the 3-centre integrals (mn|P)
 * before the metric contraction (nullable, the reference's pqA layout,
 * build_DF_tensor.c:14-47). */
int ygpu_build_df_tensor_ex(ygpu_ctx_t *ctx, double *pqA, double *Jpq_invsqrt, double *df_tensor,
                            ygpu_df_timing_t *timing);

/* Upload a DF tensor already built on the host in the reference layout
 * (nnz x naux, both triangles) instead of building it on the device. */
int ygpu_set_df_tensor(ygpu_ctx_t *ctx, const double *df_tensor);

/* J, K and F = Hcore + 2J - K from D and C_occ (TinySCF_build_FockMat,
 * build_Fock.c:349-396).  D, Hcore: nbf x nbf (Hcore NULL: the copy cached by
 * ygpu_ovlp_hcore or by an earlier call); Cocc: nbf x n_occ.  D and Cocc both
 * NULL, or the SCF hand-off mode on (ygpu_set_scf_resident) after a
 * ygpu_build_density: that build's D and C_occ, already on the device (no
 * host copy).  J, K (nullable) and F receive full symmetric matrices; J
 * entries of screened pairs are zero.  The call's phase timers (timing,
 * nullable) are also kept by the engine (ygpu_get_fock_timings). */
int ygpu_build_fock(ygpu_ctx_t *ctx, int n_occ, const double *D, const double *Cocc,
                    const double *Hcore, double *J, double *K, double *F,
                    ygpu_fock_timing_t *timing);

/* Arithmetic of the following ygpu_build_fock calls.  The reference builds
 * J and K in FP64 only (build_Fock.c:37-347); YGPU_FOCK_MIXED and
 * YGPU_FOCK_MIXED6 are this library's reduced-precision modes for the early
 * SCF iterations:
 *   K: X = C_occ^T B and K = X X^T as BF16-split products on the matrix cores
 *      with FP32 accumulation, X kept as its BF16 parts, FP64 sums across k
 *      chunks.  YGPU_FOCK_MIXED ("BF16x3"): each FP32 operand split into a
 *      hi and a lo part, products hi*hi + hi*lo + lo*hi (~2^-16 relative).
 *      YGPU_FOCK_MIXED6 ("BF16x6"): three parts, which hold the FP32 value
 *      exactly, and the six products of part-index sum below 3 (~2^-24
 *      relative, FP32-accurate products at twice the BF16x3 cost);
 *   J: incremental, J = J_prev + J(D - D_prev), from an FP32 copy of B with
 *      FP64 arithmetic; the first build after entering a reduced mode from
 *      FP64 is a full FP64 J that starts the increments.
 * YGPU_FOCK_FP64 (the default) is the reference arithmetic; switching between
 * it and a reduced mode restarts the incremental J.  The FP32 copy of B costs
 * n_up x naux_pad x 4 bytes of device memory from the first reduced-mode build
 * until the engine is destroyed (it is refreshed when the DF tensor changes). */
#define YGPU_FOCK_FP64 0
#define YGPU_FOCK_MIXED 1
#define YGPU_FOCK_MIXED6 2
/* YGPU_FOCK_INCR: difference-density build against the last FP64 build on this
 * engine (the reference: its C_occ = C_ref and K_ref, with the same n_occ).
 * The new C_occ is first rotated within its occupied space to C1 = C Q, Q
 * the orthogonal polar factor of C_ref^T S C (Newton-Schulz iteration; S from
 * ygpu_set_overlap), so that M = C1 - C_ref is as small as the change of the
 * density; then D - D_ref = (P M^T + M P^T) / 2 with P = C1 + C_ref and
 *   K = K_ref + (X^M (X^P)^T + X^P (X^M)^T) / 2,  X^Y_i = Y(partners)^T B(rows of i),
 * the difference in BF16x3 products (FP32 accumulation), whose error scales with
 * |M| instead of |K|; J incremental from the FP32 tensor as above.  Without a
 * reference, or when the alignment does not converge (a principal angle near
 * 90 degrees), the call makes a full FP64 build instead, which becomes the new
 * reference.  Every FP64 build records its C_occ and K as the reference
 * (nbf x n_occ + nbf x nbf doubles). */
#define YGPU_FOCK_INCR 3
int ygpu_set_fock_precision(ygpu_ctx_t *ctx, int mode);
/* What the last ygpu_build_fock ran: mode_used (nullable) the YGPU_FOCK_*
 * arithmetic (a YGPU_FOCK_INCR call that had to make a full FP64 build reports
 * YGPU_FOCK_FP64), align_err (nullable) max |Q^T Q - I| of a difference build's
 * alignment (0 otherwise). */
int ygpu_fock_last_info(ygpu_ctx_t *ctx, int *mode_used, double *align_err);
/* Energy terms of the last ygpu_build_fock, summed on the device in a fixed
 * order (the sum TinySCF_calc_energy forms on the host, TinySCF.c:366-372):
 * e_dhf = sum_ij D_ij (Hcore_ij + F_ij).  e_ref: after a YGPU_FOCK_INCR build
 * that ran as one, sum_ij D (F_r - Hcore) - D_r (F - Hcore) with F_r, D_r the
 * Fock matrix and density of the last full FP64 build, each element's terms
 * combined before the sum; 0 otherwise.  With G = F - Hcore symmetric and
 * linear in D, tr(D_r G(D)) = tr(D G(D_r)), so e_dhf + e_ref carries the
 * reduced-precision error of the difference only at second order in D - D_r.
 * Both pointers nullable; valid until the next density build. */
int ygpu_fock_energy(ygpu_ctx_t *ctx, double *e_dhf, double *e_ref);

/* ---- Fock build without device allocation.  No counterpart in the reference
 * (whose Fock build, build_Fock.c:349-396, allocates nothing per call because
 * TinySCF_init_free.c preallocates host buffers): it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.
 * As ygpu_build_fock in FP64 arithmetic (YGPU_FOCK_FP64 must be set), with D,
 * Cocc and Hcore required host inputs and J, K (nullable) and F host outputs;
 * the device work buffers (copies of the matrices, X = C_occ^T B, the split-K
 * and J-pass partials) are carved from one static device array of
 * YGPU_NOALLOC_SCRATCH_BYTES per GPU, taken at the library's load and held until
 * the process ends.  It calls no hipMalloc / hipFree.  A call whose buffers
 * need more than that array (about nbf * n_occ * naux * 8 bytes for X; C10H22
 * in cc-pVDZ / cc-pVDZ-RI fits) returns YGPU_ERR_INVALID.  Every engine on a
 * GPU shares that GPU's array: calls that overlap in time on one GPU race. */
#define YGPU_NOALLOC_SCRATCH_BYTES ((size_t) 128 << 20)
int ygpu_build_fock_noalloc(ygpu_ctx_t *ctx, int n_occ, const double *D, const double *Cocc,
                            const double *Hcore, double *J, double *K, double *F);

/* Density from the orthogonal-basis Fock matrix F_orth = X^T F X
 * (TinySCF_build_DenMat, build_density.c:37-69): eigval ascending, C = X C0,
 * Cocc = the n_occ lowest columns (nbf x n_occ row-major), D = Cocc Cocc^T.
 * X NULL: the copy cached by ygpu_set_overlap.  F_orth NULL: the output of the
 * last ygpu_diis_transform / ygpu_diis_extrapolate, already on the device.  In
 * the SCF hand-off mode (ygpu_set_scf_resident) host F_orth and X are not
 * copied up while those device copies exist.  ev_idx (nullable) receives the
 * order used (identity for the ascending eigenvalues of the eigensolver). */
int ygpu_build_density(ygpu_ctx_t *ctx, int n_occ, const double *F_orth, const double *X,
                       double *D, double *Cocc, double *eigval, int *ev_idx);

/* How ygpu_build_density finds the occupied space (no counterpart in the
 * reference, which always diagonalises; build_density.c:54).
 * YGPU_DENSITY_EIGEN (default): the eigensolver, as above.
 * YGPU_DENSITY_PURIFY: the projector P onto the n_occ lowest eigenvectors of
 * F_orth by SP2 purification (FP64 matrix squares), and the occupied vectors
 * V = P V_prev (V_prev^T P V_prev)^{-1/2} from the previous build's V_prev, so
 * D = X P X^T as from the eigenvectors (C_occ differs by a rotation within the
 * occupied space).  eigval is then not written.  The eigensolver runs instead
 * for the first build, after a change of n_occ, when SP2 needs more than 96
 * steps (a vanishing gap) or when the two occupied spaces are nearly
 * orthogonal. */
#define YGPU_DENSITY_EIGEN 0
#define YGPU_DENSITY_PURIFY 1
int ygpu_set_density_method(ygpu_ctx_t *ctx, int method);
/* What the last ygpu_build_density ran: method_used (nullable) the
 * YGPU_DENSITY_* method, sp2_steps (nullable) the SP2 steps (0 for the
 * eigensolver). */
int ygpu_density_last_info(ygpu_ctx_t *ctx, int *method_used, int *sp2_steps);

/* The RHF density nearest a guess density D_guess (nbf x nbf), for an initial
 * guess that is not idempotent (no counterpart in the reference, whose SAD
 * guess goes to D2Cocc, TinySCF.c:282-322): the n_occ eigenvectors of largest
 * occupation of S^{1/2} D_guess S^{1/2} (S^{1/2} = S X with the cached S and
 * X), by the density build on F_orth = -S^{1/2} D_guess S^{1/2}: D, Cocc and
 * eigval as ygpu_build_density writes them. */
int ygpu_project_guess(ygpu_ctx_t *ctx, int n_occ, const double *D_guess, double *D, double *Cocc,
                       double *eigval);

/* Cache the overlap S and X = S^{-1/2} (nbf x nbf) on the device for the
 * density build and DIIS.  ygpu_ovlp_hcore also caches Hcore; passing NULL
 * for Hcore (ygpu_build_fock) or X (ygpu_build_density) uses the cached
 * device copies. */
int ygpu_set_overlap(ygpu_ctx_t *ctx, const double *S, const double *X);

/* Commutator DIIS with the history on the device (DIIS.c:12-123); the caller
 * keeps the reference's bookkeeping (DIIS_len, the replaced slot, B_mat,
 * DIIS_bmax) and solves the small system.
 *   ygpu_diis_transform: F_orth = X^T F X (iterations 0 and 1, DIIS.c:30-40);
 *     F NULL (or the hand-off mode on): the last ygpu_build_fock's F on the
 *     device.
 *   ygpu_diis_update: residual R = X^T (F D S - S D F) X stored in history
 *     slot `slot`, dots[j] = <R, R_j> for j < len, F_orth = X^T F X stored as
 *     the slot's Fock matrix and returned (DIIS.c:51-110); F or D NULL (or the
 *     hand-off mode on): the last ygpu_build_fock's F and the D it was built
 *     from; F_orth NULL: not returned.  max |R| is formed in the same pass (ygpu_diis_error).
 *   ygpu_diis_extrapolate: F_out = sum_{i<len} coef[i] F_orth_i (DIIS.c:119-122).
 * len <= YGPU_MAX_DIIS. */
#define YGPU_MAX_DIIS 10
int ygpu_diis_transform(ygpu_ctx_t *ctx, const double *F, double *F_orth);
int ygpu_diis_update(ygpu_ctx_t *ctx, int slot, int len, const double *F, const double *D, double *dots,
                     double *F_orth);
int ygpu_diis_extrapolate(ygpu_ctx_t *ctx, int len, const double *coef, double *F_out);
/* max |R| of the residual stored in history slot `slot` (the DIIS error that
 * DIIS.c:57-65 forms; used to leave the mixed-precision phase). */
int ygpu_diis_error(ygpu_ctx_t *ctx, int slot, double *max_abs);

/* Row-major FP64 GEMM on host operands, C = alpha op(A) op(B) + beta C
 * (the cblas_dgemm calls of DIIS.c:34-107). */
int ygpu_dgemm(ygpu_ctx_t *ctx, int transa, int transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc);

/* SCF hand-off mode (no counterpart in the reference, whose matrices all live
 * on the host).  With on != 0, host inputs whose current value the device
 * already holds are not copied up again: D and C_occ after a
 * ygpu_build_density, F and its D after a ygpu_build_fock, F_orth after a DIIS
 * transform or extrapolation, and the cached Hcore and X.  The caller promises
 * not to change those host matrices between the calls of one SCF loop, as the
 * reference's TinySCF_do_SCF does not (TinySCF.c:378-435); the host outputs
 * are written as before.  Off (the default) copies every host input. */
int ygpu_set_scf_resident(ygpu_ctx_t *ctx, int on);

/* Phase timers kept by the engine: of every ygpu_build_fock since the last
 * ygpu_clear_timings, oldest first (count receives how many, at most max are
 * copied; the engine keeps at most 4096), and of the last DF tensor build. */
int ygpu_get_fock_timings(ygpu_ctx_t *ctx, int max, ygpu_fock_timing_t *out, int *count);
int ygpu_get_df_timing(ygpu_ctx_t *ctx, ygpu_df_timing_t *out);
int ygpu_clear_timings(ygpu_ctx_t *ctx);

/* Device-resident pair layout sizes: rows of the packed tensor (i <= j),
 * rows of the reference layout (both triangles) and the padded aux width. */
int ygpu_layout_info(ygpu_ctx_t *ctx, long *n_up, long *nnz, int *naux_pad);

/* ---- learned initial guess: the pair-correction network (no counterpart in
 * the reference, whose guess is SAD only, TinySCF.c:324-364).  FP32 MLP
 * n_in -> nh -> nh -> nh -> n_out with SiLU after the hidden layers; params
 * packed as W1 [n_in][nh], b1 [nh], W2 [nh][nh], b2, W3 [nh][nh], b3,
 * W4 [nh][n_out], b4 (row-major, y = x W + b).  Host pointers throughout. */

/* Y [n][n_out] = network(X [n][n_in]) */
int ygpu_mlp_forward(ygpu_ctx_t *ctx, int n_in, int nh, int n_out, const float *params, int n,
                     const float *X, float *Y);

/* Device-resident trainer: the samples X [n][n_in], targets Y [n][n_out] and
 * 0/1 masks M [n][n_out] are uploaded once; the loss of a batch of B samples
 * is sum M (net(X) - Y)^2 / (B n_out); Adam (0.9, 0.999, 1e-8). */
typedef struct ygpu_mlp_trainer ygpu_mlp_trainer_t;
int ygpu_mlp_trainer_create(ygpu_ctx_t *ctx, int n_in, int nh, int n_out, int n_samples, const float *X,
                            const float *Y, const float *M, const float *params_init,
                            ygpu_mlp_trainer_t **trainer);
/* one pass over the samples in batches, in the order of a shuffle seeded by seed */
int ygpu_mlp_trainer_epoch(ygpu_mlp_trainer_t *trainer, int batch, float lr, unsigned seed, double *mean_loss);
/* loss and gradient (packed like params) on the samples idx[0..n) */
int ygpu_mlp_trainer_grad(ygpu_mlp_trainer_t *trainer, int n, const int *idx, float *grad, double *loss);
int ygpu_mlp_trainer_params(ygpu_mlp_trainer_t *trainer, float *params);
int ygpu_mlp_trainer_destroy(ygpu_mlp_trainer_t *trainer);

#ifdef __cplusplus
}
#endif
#endif
