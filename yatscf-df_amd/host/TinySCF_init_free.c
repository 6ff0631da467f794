/*
 * TinySCF_init_free.c -- engine lifecycle (drop-in for the reference's
 * TinySCF_init_free.c:14-344).
 *
 * Host allocations and constants follow the reference: ene_tol = 1e-11
 * (:75), shell_scrtol2 = 1e-22 (:78), DIIS B matrix initialised to -1 with a
 * zero diagonal (:173-178), bmax = -DBL_MAX (:179-180).  The Simint object
 * (:103) is replaced by the device engine, which receives both shell sets
 * and the nuclei here.  temp_J / temp_K (the reference's per-thread J buffers
 * and the nbf x n_occ x naux K tensor, :186-199) live on the device.
 */
#include <float.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "TinySCF.h"
#include "utils.h"
#include "yatscf_gpu.h"

static void *xalloc(size_t bytes, double *mem)
{
    void *p = ALIGN64B_MALLOC(bytes);
    if (!p) {
        fprintf(stderr, "TinySCF: host allocation of %zu bytes failed\n", bytes);
        abort();
    }
    *mem += (double) bytes;
    return p;
}

/* page-locked: the matrices copied to or from the device on every iteration */
static void *xalloc_pinned(size_t bytes, double *mem)
{
    void *p = NULL;
    if (ygpu_host_alloc(bytes, &p) != YGPU_OK || !p) {
        fprintf(stderr, "TinySCF: page-locked allocation of %zu bytes failed: %s\n", bytes, ygpu_last_error());
        abort();
    }
    *mem += (double) bytes;
    return p;
}

static int env_int(const char *name, int def)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : def;
}

int TinySCF_try_init(TinySCF_t T, char *bas_fname, char *df_bas_fname, char *xyz_fname, const int niters)
{
    if (!T) return -1;
    double st = get_wtime_sec();
    memset(T, 0, sizeof(*T));
    T->quirk_a5 = env_int("YATSCF_QUIRK_A5", 1);
    T->quirk_a1 = env_int("YATSCF_QUIRK_A1", 1);
    T->quirk_a2 = env_int("YATSCF_QUIRK_A2", 1);
    T->mixed = env_int("YATSCF_MIXED", 0);
    T->mixed_tol = 1e-5;
    {
        const char *e = getenv("YATSCF_MIXED_TOL");
        if (e && *e && atof(e) > 0.0) T->mixed_tol = atof(e);
    }
    T->mixed_schedule = env_int("YATSCF_MIXED_SCHEDULE", 5);
    T->mixed_keep_diis = env_int("YATSCF_MIXED_KEEP_DIIS", 2);
    T->mixed6_floor = 1.5 * 4.76837158203125e-07;     /* 1.5 * 2^-21 */
    T->mixed_reref = 0.0;
    T->mixed_confirm = env_int("YATSCF_MIXED_CONFIRM", 0);
    T->purify = env_int("YATSCF_PURIFY", 1);
    {
        const char *e = getenv("YATSCF_MIXED_REREF");
        if (e && *e && atof(e) >= 0.0) T->mixed_reref = atof(e);
    }
    {
        const char *e = getenv("YATSCF_MIXED6_FLOOR");
        if (e && *e && atof(e) > 0.0) T->mixed6_floor = atof(e);
    }
    T->mixed_phase = 0;
    T->mixed_switch_iter = -1;
    T->mixed6_switch_iter = -1;
    T->diis_error = -1.0;
    T->guess_ml = 0;
    T->guess_weights[0] = 0;
    {
        const char *w = getenv("YATSCF_GUESS_WEIGHTS");
        if (w && *w) snprintf(T->guess_weights, sizeof(T->guess_weights), "%s", w);
    }
    T->host_df_copy = env_int("YATSCF_HOST_DF", 0);
    T->verbose = env_int("YATSCF_VERBOSE", 1);

    if (CMS_createBasisSet(&T->basis) != CMS_STATUS_SUCCESS || CMS_createBasisSet(&T->df_basis) != CMS_STATUS_SUCCESS)
        return -1;
    if (CMS_loadChemicalSystem(T->basis, bas_fname, xyz_fname) != CMS_STATUS_SUCCESS) {
        fprintf(stderr, "TinySCF: %s\n", CMS_lastError());
        return -2;
    }
    if (CMS_loadChemicalSystem(T->df_basis, df_bas_fname, xyz_fname) != CMS_STATUS_SUCCESS) {
        fprintf(stderr, "TinySCF: %s\n", CMS_lastError());
        return -2;
    }
    T->natoms = CMS_getNumAtoms(T->basis);
    T->nshells = CMS_getNumShells(T->basis);
    T->nbasfuncs = CMS_getNumFuncs(T->basis);
    T->n_occ = CMS_getNumOccOrb(T->basis);
    T->charge = CMS_getTotalCharge(T->basis);
    T->electron = CMS_getNneutral(T->basis);
    T->df_nbf = CMS_getNumFuncs(T->df_basis);
    T->df_nshells = CMS_getNumShells(T->df_basis);
    {
        /* basis type: the first token of each .gbs file (CMS_BasisSet.c:464-483), or both set by
         * YATSCF_BASIS_TYPE */
        const char *bt = getenv("YATSCF_BASIS_TYPE");
        T->sph_orb = CMS_isSpherical(T->basis);
        T->sph_aux = CMS_isSpherical(T->df_basis);
        if (bt && !strcmp(bt, "spherical")) T->sph_orb = T->sph_aux = 1;
        else if (bt && !strcmp(bt, "cartesian")) T->sph_orb = T->sph_aux = 0;
        const int *am, *np, *po;
        const double *xyz, *al, *co;
        T->nbf_sph = T->nbasfuncs;
        if (T->sph_orb) {
            CMS_getShellArrays(T->basis, &am, &np, &po, &xyz, &al, &co);
            T->nbf_sph = 0;
            for (int s = 0; s < T->nshells; s++) T->nbf_sph += 2 * am[s] + 1;
        }
        if (T->sph_aux) {
            CMS_getShellArrays(T->df_basis, &am, &np, &po, &xyz, &al, &co);
            T->df_nbf = 0;
            for (int s = 0; s < T->df_nshells; s++) T->df_nbf += 2 * am[s] + 1;
        }
    }
    if (T->n_occ <= 0 || T->n_occ > T->nbasfuncs) {
        fprintf(stderr, "TinySCF: %d occupied orbitals for %d basis functions\n", T->n_occ, T->nbasfuncs);
        return -3;
    }
    if (T->verbose) {
        printf("Job information:\n");
        printf("    basis set            = %s\n", basename(bas_fname));
        printf("    DF basis set         = %s\n", basename(df_bas_fname));
        printf("    molecule             = %s\n", basename(xyz_fname));
        printf("    # atoms              = %d\n", T->natoms);
        printf("    # shells             = %d\n", T->nshells);
        printf("    # basis functions    = %d%s\n", T->nbf_sph, T->sph_orb ? " (spherical)" : "");
        printf("    # DF shells          = %d\n", T->df_nshells);
        printf("    # DF basis functions = %d%s\n", T->df_nbf, T->sph_aux ? " (spherical)" : "");
        printf("    # occupied orbits    = %d\n", T->n_occ);
        printf("    # charge             = %d\n", T->charge);
        printf("    # electrons          = %d\n", T->electron);
    }
    int maxAM = CMS_getMaxMomentum(T->basis);
    T->max_dim = (maxAM + 1) * (maxAM + 2) / 2;
    T->nthreads = 1;
    T->max_buf_size = 0;
    T->Accum_Fock_buf = NULL;

    T->nshellpairs = T->nshells * T->nshells;
    T->mat_size = T->nbasfuncs * T->nbasfuncs;
    T->num_uniq_sp = (T->nshells + 1) * T->nshells / 2;
    T->iter = 0;
    T->niters = niters;
    T->ene_tol = 1e-11;
    {
        const char *e = getenv("YATSCF_ENE_TOL");
        if (e && *e && atof(e) > 0.0) T->ene_tol = atof(e);
    }
    T->shell_scrtol2 = 1e-11 * 1e-11;

    const size_t ms = (size_t) T->mat_size;
    double *mem = &T->mem_size;
    T->sp_scrval = (double *) xalloc(DBL_SIZE * T->nshellpairs, mem);
    T->bf_pair_scrval = (double *) xalloc(DBL_SIZE * ms, mem);
    T->df_sp_scrval = (double *) xalloc(DBL_SIZE * T->df_nshells, mem);
    T->uniq_sp_lid = (int *) xalloc(INT_SIZE * T->num_uniq_sp, mem);
    T->uniq_sp_rid = (int *) xalloc(INT_SIZE * T->num_uniq_sp, mem);
    T->bf_pair_mask = (int *) xalloc(INT_SIZE * ms, mem);
    T->bf_pair_j = (int *) xalloc(INT_SIZE * ms, mem);
    T->bf_pair_diag = (int *) xalloc(INT_SIZE * T->nbasfuncs, mem);
    T->bf_mask_displs = (int *) xalloc(INT_SIZE * (T->nbasfuncs + 1), mem);

    T->shell_bf_sind = (int *) xalloc(INT_SIZE * (T->nshells + 1), mem);
    T->shell_bf_num = (int *) xalloc(INT_SIZE * T->nshells, mem);
    for (int i = 0; i < T->nshells; i++) {
        T->shell_bf_sind[i] = CMS_getFuncStartInd(T->basis, i);
        T->shell_bf_num[i] = CMS_getShellDim(T->basis, i);
    }
    T->shell_bf_sind[T->nshells] = T->nbasfuncs;
    T->df_shell_bf_sind = (int *) xalloc(INT_SIZE * (T->df_nshells + 1), mem);
    T->df_shell_bf_num = (int *) xalloc(INT_SIZE * T->df_nshells, mem);
    for (int i = 0, f = 0; i < T->df_nshells; i++) {
        /* spherical auxiliary basis: 2l+1 functions per shell */
        const int dim = CMS_getShellDim(T->df_basis, i);
        const int l = dim == 1 ? 0 : dim == 3 ? 1 : dim == 6 ? 2 : dim == 10 ? 3 : 4;
        T->df_shell_bf_sind[i] = T->sph_aux ? f : CMS_getFuncStartInd(T->df_basis, i);
        T->df_shell_bf_num[i] = T->sph_aux ? 2 * l + 1 : dim;
        f += T->df_shell_bf_num[i];
    }
    T->df_shell_bf_sind[T->df_nshells] = T->df_nbf;

    T->Hcore_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->S_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->F_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->D_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->J_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->K_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->X_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->tmp_mat = (double *) xalloc_pinned(DBL_SIZE * ms, mem);
    T->Cocc_mat = (double *) xalloc_pinned(DBL_SIZE * T->n_occ * T->nbasfuncs, mem);
    T->eigval = (double *) xalloc_pinned(DBL_SIZE * T->nbasfuncs, mem);
    T->ev_idx = (int *) xalloc_pinned(INT_SIZE * T->nbasfuncs, mem);
    memset(T->J_mat, 0, DBL_SIZE * ms);
    memset(T->K_mat, 0, DBL_SIZE * ms);

    T->F0_mat = (double *) xalloc(DBL_SIZE * ms * MAX_DIIS, mem);
    T->R_mat = (double *) xalloc(DBL_SIZE * ms * MAX_DIIS, mem);
    T->B_mat = (double *) xalloc(DBL_SIZE * (MAX_DIIS + 1) * (MAX_DIIS + 1), mem);
    T->FDS_mat = (double *) xalloc(DBL_SIZE * ms, mem);
    T->DIIS_rhs = (double *) xalloc(DBL_SIZE * (MAX_DIIS + 1), mem);
    T->DIIS_ipiv = (int *) xalloc(INT_SIZE * (MAX_DIIS + 1), mem);
    memset(T->F0_mat, 0, DBL_SIZE * ms * MAX_DIIS);
    memset(T->R_mat, 0, DBL_SIZE * ms * MAX_DIIS);
    T->DIIS_len = 0;
    for (int i = 0; i < (MAX_DIIS + 1) * (MAX_DIIS + 1); i++) T->B_mat[i] = -1.0;
    for (int i = 0; i < MAX_DIIS + 1; i++) T->B_mat[i * (MAX_DIIS + 1) + i] = 0.0;
    T->DIIS_bmax_id = 0;
    T->DIIS_bmax = -DBL_MAX;

    T->df_nbf_16 = (T->df_nbf + 15) / 16 * 16;
    T->Jpq = (double *) xalloc_pinned(DBL_SIZE * (size_t) T->df_nbf * T->df_nbf, mem);
    T->temp_J = NULL;
    T->temp_K = NULL;
    T->pqA = NULL;
    T->df_tensor = NULL;

    /* device engine: both shell sets and the nuclei (replaces CMS_createSimint) */
    ygpu_ctx_t *g = NULL;
    int st_c = ygpu_create(env_int("YATSCF_DEVICE", 0), &g);
    if (st_c != YGPU_OK) {
        fprintf(stderr, "TinySCF: %s\n", ygpu_last_error());
        return -4;
    }
    T->gpu = g;
    const int *am, *np, *po;
    const double *xyz, *al, *co, *chg, *axyz;
    CMS_getShellArrays(T->basis, &am, &np, &po, &xyz, &al, &co);
    if (ygpu_set_basis(g, 0, T->nshells, am, np, po, xyz, al, co) != YGPU_OK) goto gpu_fail;
    CMS_getShellArrays(T->df_basis, &am, &np, &po, &xyz, &al, &co);
    if (ygpu_set_basis(g, 1, T->df_nshells, am, np, po, xyz, al, co) != YGPU_OK) goto gpu_fail;
    if (T->sph_aux && ygpu_set_aux_spherical(g, 1) != YGPU_OK) goto gpu_fail;
    CMS_getAtoms(T->basis, &chg, &axyz);
    if (ygpu_set_atoms(g, T->natoms, chg, axyz) != YGPU_OK) goto gpu_fail;
    /* factor of the Coulomb metric: the reference's J^{-1/2} (YATSCF_METRIC=invsqrt) or the
     * Cholesky factor's inverse (the default: the same J and K, a third of the DF build time) */
    {
        const char *m = getenv("YATSCF_METRIC");
        const int mode = (m && !strcmp(m, "invsqrt")) ? YGPU_METRIC_INVSQRT : YGPU_METRIC_CHOLESKY;
        if (ygpu_set_metric_factor(g, mode) != YGPU_OK) goto gpu_fail;
    }

    T->init_time = get_wtime_sec() - st;
    return 0;
gpu_fail:
    fprintf(stderr, "TinySCF: %s\n", ygpu_last_error());
    return -4;
}

void init_TinySCF(TinySCF_t T, char *bas_fname, char *df_bas_fname, char *xyz_fname, const int niters)
{
    int st = TinySCF_try_init(T, bas_fname, df_bas_fname, xyz_fname, niters);
    if (st != 0) {
        fprintf(stderr, "TinySCF: initialisation failed (%d)\n", st);
        abort();
    }
}

void free_TinySCF(TinySCF_t T)
{
    if (!T) return;
    if (T->gpu) ygpu_destroy((ygpu_ctx_t *) T->gpu);
    void *ptrs[] = {T->Accum_Fock_buf, T->sp_scrval, T->bf_pair_scrval, T->df_sp_scrval, T->uniq_sp_lid,
                    T->uniq_sp_rid, T->shell_bf_sind, T->shell_bf_num, T->df_shell_bf_sind, T->df_shell_bf_num,
                    T->bf_pair_mask, T->bf_pair_j, T->bf_pair_diag, T->bf_mask_displs, T->F0_mat, T->R_mat,
                    T->B_mat, T->FDS_mat, T->DIIS_rhs, T->DIIS_ipiv, T->temp_J, T->temp_K, T->pqA, T->df_tensor};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) ALIGN64B_FREE(ptrs[i]);
    void *pinned[] = {T->Hcore_mat, T->S_mat, T->F_mat, T->D_mat, T->J_mat, T->K_mat, T->X_mat,
                      T->tmp_mat, T->Cocc_mat, T->eigval, T->ev_idx, T->Jpq};
    for (size_t i = 0; i < sizeof(pinned) / sizeof(pinned[0]); i++) (void) ygpu_host_free(pinned[i]);
    CMS_destroyBasisSet(T->basis);
    CMS_destroyBasisSet(T->df_basis);
    free(T);   /* the reference frees the struct itself (TinySCF_init_free.c:267) */
}

/* The K build is a device SYRK (no MKL batch arrays); the tile size the
 * reference would use (TinySCF_init_free.c:278-279) is still recorded. */
void TinySCF_init_batch_dgemm_arrays(TinySCF_t T)
{
    int BS = T->nbasfuncs / 10;
    if (BS < 64) BS = 64;
    T->mat_K_BS = BS;
    int nb = (T->nbasfuncs + BS - 1) / BS;
    T->mat_K_ntiles = nb * (nb + 1) / 2;
}

void TinySCF_free_batch_dgemm_arrays(TinySCF_t T) { (void) T; }
