/*
 * TinySCF.h -- the TinySCF engine struct and lifecycle API (drop-in for the
 * reference's TinySCF.h:12-138).
 *
 * Field names, meanings and layouts are the reference's.  Differences:
 *  - `gpu` (the device engine, include/yatscf_gpu.h) replaces the Simint
 *    handle; integrals, the DF tensor, J/K and the density build run there;
 *  - temp_J / temp_K and the MKL batch-GEMM arrays stay unallocated on the
 *    host (the device owns that work);
 *  - pqA and df_tensor are host copies made only when `host_df_copy` is set
 *    (the device keeps the packed tensor; the host copies are 2 x nnz x naux
 *    doubles in the reference layout);
 *  - switches for reference behaviours that a user may want to change (the
 *    Schwarz element read, D2Cocc, the DIIS replacement rule), the
 *    mixed-precision and learned-guess options and a convergence flag (the
 *    drop-in build_DF_tensor.c, build_Fock.c and build_density.c read none of
 *    these: only the reference's fields and `gpu`; the phase timers live in
 *    the device engine, ygpu_get_fock_timings).
 */
#ifndef _YATSCF_TFOCK_H_
#define _YATSCF_TFOCK_H_

#include "CMS.h"

#define MAX_DIIS 10
#define MIN_DIIS 2

struct TinySCF_struct
{
    int    nthreads;
    int    max_buf_size;
    double *Accum_Fock_buf;

    BasisSet_t basis;
    BasisSet_t df_basis;
    int natoms, nshells;
    int nbasfuncs, n_occ;
    int charge, electron;
    int df_nbf, df_nshells;

    int nshellpairs;
    int num_uniq_sp;
    int mat_size;
    int max_dim;

    int    niters, iter;
    double nuc_energy;
    double HF_energy;
    double ene_tol;

    double shell_scrtol2;
    double max_scrval;
    double max_df_scrval;
    double *sp_scrval;
    double *bf_pair_scrval;
    int    *uniq_sp_lid;
    int    *uniq_sp_rid;
    double *df_sp_scrval;
    int    *bf_pair_mask;
    int    *bf_pair_j;
    int    *bf_pair_diag;
    int    *bf_mask_displs;
    int    num_bf_pair_scr;

    void   *gpu;            /* ygpu_ctx_t* (replaces Simint_t simint) */
    int *shell_bf_sind;
    int *shell_bf_num;
    int *df_shell_bf_sind;
    int *df_shell_bf_num;

    double *Hcore_mat;
    double *S_mat;
    double *F_mat;
    double *D_mat;
    double *J_mat;
    double *K_mat;
    double *X_mat;
    double *Cocc_mat;
    double *eigval;
    int    *ev_idx;
    double *tmp_mat;

    double *pqA, *Jpq, *df_tensor;
    double *temp_J, *temp_K;
    int    df_nbf_16;

    double *F0_mat;
    double *R_mat;
    double *B_mat;
    double *FDS_mat;
    double *DIIS_rhs;
    int    *DIIS_ipiv;
    int    DIIS_len;
    int    DIIS_bmax_id;
    double DIIS_bmax;

    int mat_K_BS, *mat_K_group_size, mat_K_ntiles;
    int *mat_K_transa, *mat_K_transb;
    int *mat_K_m,   *mat_K_n,   *mat_K_k;
    int *mat_K_lda, *mat_K_ldb, *mat_K_ldc;
    double *mat_K_alpha, *mat_K_beta;
    double **mat_K_a, **mat_K_b, **mat_K_c;

    double mem_size, init_time, S_Hcore_time, shell_scr_time;

    /* additions */
    int    quirk_a5;        /* reference Schwarz index (TinySCF.c:153), default 1 */
    int    quirk_a1;        /* reference D2Cocc (TinySCF.c:304-318), default 1 */
    int    quirk_a2;        /* DIIS_bmax never reset (DIIS.c:93-100), default 1; 0 recomputes it
                               over the stored residuals each iteration */
    int    host_df_copy;    /* keep host pqA / df_tensor copies, default 0 */
    int    mixed;           /* 1: reduced-precision Fock builds until the DIIS error < mixed_tol */
    int    mixed_schedule;  /* phases of a mixed run: bit 0 BF16x3 products, bit 1 BF16x6
                               products (FP32-accurate), in that order, then FP64; bit 2: the
                               FP64 phase runs difference-density builds against a full FP64
                               reference build, converged when two energies agree and the later
                               is a full FP64 build or the difference build right after one;
                               default 5 */
    int    mixed_keep_diis; /* keep the DIIS history across a phase change (bit 0: BF16x3 ->
                               BF16x6, bit 1: -> FP64) instead of restarting it; default 2 (the
                               reduced-precision Fock matrices stay in the history of the FP64 and
                               difference builds, which replace them as it fills) */
    double mixed_tol;       /* DIIS error (max |R|) that ends the mixed phase, default 1e-5 */
    double mixed_reref;     /* difference-density phase: a new FP64 reference build once the
                               DIIS error is below this fraction of its value at the last
                               reference; 0 (default): no re-reference */
    int    mixed_confirm;   /* difference-density phase: 1 ends the SCF only on a full FP64 build
                               or the difference build right after one (agreement among later
                               difference builds is confirmed by a full FP64 build); 0 (default):
                               any two agreeing FP64 or difference-build energies, whose
                               reduced-precision error is of second order (see TinySCF_calc_energy) */
    int    purify;          /* 1 (default): density builds of the SCF iterations before the last
                               by SP2 purification (ygpu_set_density_method), the last by the
                               eigensolver; 0: the eigensolver throughout */
    double mixed6_floor;    /* reduced-precision floor of the BF16x6 phase in units of
                               max_i K_ii (see TinySCF_do_SCF), default 1.5 * 2^-21 */
    int    mixed_phase;     /* Fock arithmetic of the current iteration: YGPU_FOCK_FP64,
                               YGPU_FOCK_MIXED or YGPU_FOCK_MIXED6 */
    int    mixed_switch_iter; /* first FP64 iteration of a mixed run (-1: never switched) */
    int    mixed6_switch_iter; /* first BF16x6 iteration of a mixed run (-1: none) */
#define YS_MAX_PHASE_LOG 512
    int    phase_log[YS_MAX_PHASE_LOG]; /* Fock arithmetic (YGPU_FOCK_*) of each SCF iteration */
    double diis_error;      /* max |R| of the newest DIIS residual (iterations >= 2) */
    char   guess_weights[512]; /* learned-guess network file ("" = SAD guess only) */
    int    guess_ml;        /* 1 if the last initial guess used the learned correction */
    int    sph_orb, sph_aux; /* spherical orbital / auxiliary basis (the .gbs flag, or
                               YATSCF_BASIS_TYPE=spherical|cartesian for both): the orbitals
                               span the real solid harmonics among the Cartesian components
                               (X = T (T^T S T)^{-1/2}, zero columns beyond nbf_sph) and the
                               auxiliary functions are the solid harmonics (df_nbf counts them) */
    int    nbf_sph;         /* orbital basis functions of the calculation (nbasfuncs if Cartesian) */
    int    verbose;         /* reference-style printout, default 1 */
    int    converged;
    double energy_delta;
    double t_density, t_diis;
};

typedef struct TinySCF_struct* TinySCF_t;

void init_TinySCF(TinySCF_t TinySCF, char *bas_fname, char *df_bas_fname, char *xyz_fname, const int niters);
void free_TinySCF(TinySCF_t TinySCF);
void TinySCF_init_batch_dgemm_arrays(TinySCF_t TinySCF);
void TinySCF_free_batch_dgemm_arrays(TinySCF_t TinySCF);
void TinySCF_compute_Hcore_Ovlp_mat(TinySCF_t TinySCF);
void TinySCF_compute_sq_Schwarz_scrvals(TinySCF_t TinySCF);
void TinySCF_prepare_sparsity(TinySCF_t TinySCF);
void TinySCF_get_initial_guess(TinySCF_t TinySCF);
void TinySCF_do_SCF(TinySCF_t TinySCF);
void TinySCF_D2Cocc(TinySCF_t TinySCF);

/* Status-returning variant of init_TinySCF (0 = success); init_TinySCF
 * aborts with the message instead, keeping the reference's void contract. */
int  TinySCF_try_init(TinySCF_t TinySCF, char *bas_fname, char *df_bas_fname, char *xyz_fname, const int niters);

#endif
