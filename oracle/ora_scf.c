/*
 * ora_scf.c -- CPU oracle restatement of the reference DF-RHF engine
 * (TEST INFRASTRUCTURE ONLY; also the timed CPU baseline of bench.py).
 *
 * Follows, function by function:
 *   init_TinySCF                     TinySCF_init_free.c:14-208
 *   TinySCF_compute_Hcore_Ovlp_mat   TinySCF.c:20-83
 *   TinySCF_compute_sq_Schwarz_scrvals TinySCF.c:112-226 (quirk A5 at :153)
 *   TinySCF_prepare_sparsity         TinySCF.c:228-280
 *   TinySCF_D2Cocc                   TinySCF.c:282-322 (quirk A1)
 *   TinySCF_get_initial_guess        TinySCF.c:324-364
 *   calc_DF_3center_integrals        build_DF_tensor.c:49-150 (+ copy :14-47, A6 fixed)
 *   calc_DF_2center_integrals        build_DF_tensor.c:152-203
 *   calc_inverse_sqrt_Jpq            build_DF_tensor.c:205-237
 *   generate_df_tensor               build_DF_tensor.c:239-254
 *   build_J_mat / reduce_temp_J      build_Fock.c:15-130 (A7: J zeroed)
 *   build_K_mat_Cocc                 build_Fock.c:201-347
 *   TinySCF_build_FockMat            build_Fock.c:349-396
 *   TinySCF_build_DenMat             build_density.c:37-69 (A4: stable order)
 *   TinySCF_DIIS                     DIIS.c:12-123 (A2, A3 reproduced)
 *   TinySCF_calc_energy / do_SCF     TinySCF.c:367-435 (A8)
 *
 * The reference's host layout is kept: pqA and df_tensor hold both triangles
 * (rows = bf_mask_displs CSR), temp_K is nbf x n_occ x df_nbf.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <omp.h>
#include "ora.h"

#define MAX_DIIS 10

struct ora_scf {
    int quirk_a5, quirk_a1, quirk_a2;
    double ene_tol, shell_scrtol2;
    int nthreads;

    int natoms; double *Z, *axyz;
    int charge, electron, n_occ; double nuc_energy;
    int nshells, nbf, df_nshells, df_nbf, mat_size;
    ora_shell_t *shells, *df_shells;
    double *prim_store;
    int *shell_bf_sind, *shell_bf_num, *df_shell_bf_sind, *df_shell_bf_num;
    int df_max_am, *df_am_shell_id, *df_am_shell_spos;

    int niters, iter; double HF_energy, energy_delta;

    double *sp_scrval, *bf_pair_scrval, *df_sp_scrval, max_scrval, max_df_scrval;
    int *uniq_sp_lid, *uniq_sp_rid, num_uniq_sp;
    int *bf_pair_mask, *bf_pair_j, *bf_pair_diag, *bf_mask_displs, num_bf_pair_scr;

    double *Hcore, *S, *F, *D, *J, *K, *X, *Cocc, *eigval, *tmp_mat, *sad;
    int *ev_idx;
    double *pqA, *Jpq, *df_tensor, *temp_J, *temp_K, *Cocc_tmp;
    int df_nbf_16;

    double *F0, *R, *B, *FDS, *DIIS_rhs; int *DIIS_ipiv;
    int DIIS_len, DIIS_bmax_id; double DIIS_bmax;
    int converged;

    /* timings (s) */
    double t_setup, t_3c, t_2c, t_invsqrt, t_dfgemm;
    double t_tempJ, t_Jmat, t_tempK, t_Kmat, t_symm, t_fock, t_energy, t_diis, t_density;
    double t_scf_total;
};

static double wtime(void) { return omp_get_wtime(); }

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: allocation of %zu bytes failed\n", n * sz); abort(); }
    return p;
}

static void copy_basis(const ora_basis_t *b, ora_shell_t **shells_out, double **store_out,
                       int **sind_out, int **num_out, int *nbf_out)
{
    int ns = b->nshells, nprim_tot = 0;
    for (int i = 0; i < ns; i++) nprim_tot += b->nprim[i];
    double *store = (double *) xcalloc(2 * (size_t) nprim_tot, sizeof(double));
    ora_shell_t *sh = (ora_shell_t *) xcalloc(ns, sizeof(ora_shell_t));
    int *sind = (int *) xcalloc(ns + 1, sizeof(int)), *num = (int *) xcalloc(ns, sizeof(int));
    int off = 0, nbf = 0;
    for (int i = 0; i < ns; i++) {
        sh[i].am = b->am[i];
        sh[i].nprim = b->nprim[i];
        for (int x = 0; x < 3; x++) sh[i].xyz[x] = b->xyz[3 * i + x];
        for (int k = 0; k < b->nprim[i]; k++) {
            store[off + k] = b->alpha[b->prim_off[i] + k];
            store[nprim_tot + off + k] = b->coef[b->prim_off[i] + k];
        }
        sh[i].alpha = store + off;
        sh[i].coef = store + nprim_tot + off;
        off += b->nprim[i];
        sind[i] = nbf;
        num[i] = ora_ncart(b->am[i]);
        nbf += num[i];
    }
    sind[ns] = nbf;
    *shells_out = sh; *store_out = store; *sind_out = sind; *num_out = num; *nbf_out = nbf;
}

ora_scf_t *ora_create(const ora_basis_t *bas, const ora_basis_t *aux,
                      int natoms, const double *Z, const double *axyz,
                      int charge, int nelectron, double nuc_energy,
                      const double *sad_blockdiag, int niters)
{
    ora_lapack_init();
    ora_scf_t *s = (ora_scf_t *) xcalloc(1, sizeof(ora_scf_t));
    s->quirk_a5 = 1; s->quirk_a1 = 1; s->quirk_a2 = 1;
    s->ene_tol = 1e-11;                       /* TinySCF_init_free.c:75 */
    s->shell_scrtol2 = 1e-11 * 1e-11;         /* TinySCF_init_free.c:78 */
    s->nthreads = ora_get_threads();
    s->natoms = natoms;
    s->Z = (double *) xcalloc(natoms, sizeof(double));
    s->axyz = (double *) xcalloc(3 * (size_t) natoms, sizeof(double));
    memcpy(s->Z, Z, sizeof(double) * natoms);
    memcpy(s->axyz, axyz, sizeof(double) * 3 * natoms);
    s->charge = charge; s->electron = nelectron;
    s->n_occ = (nelectron - charge) / 2;      /* CMS_BasisSet.c:788-791 */
    s->nuc_energy = nuc_energy;
    double *st1, *st2;
    copy_basis(bas, &s->shells, &st1, &s->shell_bf_sind, &s->shell_bf_num, &s->nbf);
    copy_basis(aux, &s->df_shells, &st2, &s->df_shell_bf_sind, &s->df_shell_bf_num, &s->df_nbf);
    s->prim_store = st1;
    /* keep aux store alive by chaining it onto tmp pointer list: store in sad slot later */
    s->nshells = bas->nshells; s->df_nshells = aux->nshells;
    s->mat_size = s->nbf * s->nbf;
    s->niters = niters;
    int nbf = s->nbf, df_nbf = s->df_nbf;
    size_t ms = (size_t) s->mat_size;

    /* group aux shells by AM (CMS_Simint.c:184-204) */
    s->df_max_am = 0;
    for (int i = 0; i < s->df_nshells; i++) if (s->df_shells[i].am > s->df_max_am) s->df_max_am = s->df_shells[i].am;
    s->df_am_shell_id = (int *) xcalloc(s->df_nshells, sizeof(int));
    s->df_am_shell_spos = (int *) xcalloc(s->df_max_am + 2, sizeof(int));
    {
        int cnt = 0;
        for (int am = 0; am <= s->df_max_am; am++) {
            s->df_am_shell_spos[am] = cnt;
            for (int i = 0; i < s->df_nshells; i++) if (s->df_shells[i].am == am) s->df_am_shell_id[cnt++] = i;
        }
        s->df_am_shell_spos[s->df_max_am + 1] = cnt;
    }

    s->sp_scrval = (double *) xcalloc((size_t) s->nshells * s->nshells, sizeof(double));
    s->bf_pair_scrval = (double *) xcalloc(ms, sizeof(double));
    s->df_sp_scrval = (double *) xcalloc(s->df_nshells, sizeof(double));
    s->uniq_sp_lid = (int *) xcalloc((size_t) s->nshells * (s->nshells + 1) / 2, sizeof(int));
    s->uniq_sp_rid = (int *) xcalloc((size_t) s->nshells * (s->nshells + 1) / 2, sizeof(int));
    s->bf_pair_mask = (int *) xcalloc(ms, sizeof(int));
    s->bf_pair_j = (int *) xcalloc(ms, sizeof(int));
    s->bf_pair_diag = (int *) xcalloc(nbf, sizeof(int));
    s->bf_mask_displs = (int *) xcalloc(nbf + 1, sizeof(int));

    s->Hcore = (double *) xcalloc(ms, sizeof(double));
    s->S = (double *) xcalloc(ms, sizeof(double));
    s->F = (double *) xcalloc(ms, sizeof(double));
    s->D = (double *) xcalloc(ms, sizeof(double));
    s->J = (double *) xcalloc(ms, sizeof(double));
    s->K = (double *) xcalloc(ms, sizeof(double));
    s->X = (double *) xcalloc(ms, sizeof(double));
    s->tmp_mat = (double *) xcalloc(ms, sizeof(double));
    s->sad = (double *) xcalloc(ms, sizeof(double));
    memcpy(s->sad, sad_blockdiag, sizeof(double) * ms);
    s->Cocc = (double *) xcalloc((size_t) nbf * (s->n_occ > 0 ? s->n_occ : 1), sizeof(double));
    s->eigval = (double *) xcalloc(nbf, sizeof(double));
    s->ev_idx = (int *) xcalloc(nbf, sizeof(int));

    s->F0 = (double *) xcalloc(ms * MAX_DIIS, sizeof(double));
    s->R = (double *) xcalloc(ms * MAX_DIIS, sizeof(double));
    s->B = (double *) xcalloc((MAX_DIIS + 1) * (MAX_DIIS + 1), sizeof(double));
    s->FDS = (double *) xcalloc(ms, sizeof(double));
    s->DIIS_rhs = (double *) xcalloc(MAX_DIIS + 1, sizeof(double));
    s->DIIS_ipiv = (int *) xcalloc(MAX_DIIS + 1, sizeof(int));
    s->DIIS_len = 0;
    for (int i = 0; i < (MAX_DIIS + 1) * (MAX_DIIS + 1); i++) s->B[i] = -1.0;   /* TinySCF_init_free.c:175-178 */
    for (int i = 0; i < MAX_DIIS + 1; i++) s->B[i * (MAX_DIIS + 1) + i] = 0.0;
    s->DIIS_bmax_id = 0;
    s->DIIS_bmax = -DBL_MAX;

    s->df_nbf_16 = (df_nbf + 15) / 16 * 16;
    s->Jpq = (double *) xcalloc((size_t) df_nbf * df_nbf, sizeof(double));
    s->temp_J = (double *) xcalloc((size_t) s->df_nbf_16 * s->nthreads, sizeof(double));
    s->temp_K = (double *) xcalloc((size_t) nbf * (s->n_occ > 0 ? s->n_occ : 1) * df_nbf, sizeof(double));
    s->pqA = NULL; s->df_tensor = NULL; s->Cocc_tmp = NULL;
    /* the aux primitive store is owned through df_shells[0].alpha */
    (void) st2;
    return s;
}

void ora_destroy(ora_scf_t *s)
{
    if (!s) return;
    free(s->Z); free(s->axyz);
    if (s->df_nshells > 0) free((void *) s->df_shells[0].alpha);
    free(s->prim_store);
    free(s->shells); free(s->df_shells);
    free(s->shell_bf_sind); free(s->shell_bf_num); free(s->df_shell_bf_sind); free(s->df_shell_bf_num);
    free(s->df_am_shell_id); free(s->df_am_shell_spos);
    free(s->sp_scrval); free(s->bf_pair_scrval); free(s->df_sp_scrval);
    free(s->uniq_sp_lid); free(s->uniq_sp_rid);
    free(s->bf_pair_mask); free(s->bf_pair_j); free(s->bf_pair_diag); free(s->bf_mask_displs);
    free(s->Hcore); free(s->S); free(s->F); free(s->D); free(s->J); free(s->K); free(s->X);
    free(s->tmp_mat); free(s->sad); free(s->Cocc); free(s->eigval); free(s->ev_idx);
    free(s->F0); free(s->R); free(s->B); free(s->FDS); free(s->DIIS_rhs); free(s->DIIS_ipiv);
    free(s->Jpq); free(s->temp_J); free(s->temp_K); free(s->pqA); free(s->df_tensor); free(s->Cocc_tmp);
    free(s);
}

void ora_set_flags(ora_scf_t *s, int qa5, int qa1, double ene_tol, double scrtol)
{
    s->quirk_a5 = qa5; s->quirk_a1 = qa1;
    if (ene_tol > 0) s->ene_tol = ene_tol;
    if (scrtol > 0) s->shell_scrtol2 = scrtol * scrtol;
}

/* quirk A2 off: the replaced slot is the largest residual among the stored ones */
void ora_set_quirk_a2(ora_scf_t *s, int qa2)
{
    s->quirk_a2 = qa2;
}

/* ---------------- setup ---------------- */
static void compute_hcore_ovlp(ora_scf_t *s)
{
    int nbf = s->nbf, ns = s->nshells;
    memset(s->Hcore, 0, sizeof(double) * s->mat_size);
    memset(s->S, 0, sizeof(double) * s->mat_size);
    #pragma omp parallel
    {
        double buf[64 * 64], buf2[64 * 64];
        #pragma omp for schedule(dynamic)
        for (int M = 0; M < ns; M++)
            for (int N = 0; N < ns; N++) {
                int dm = s->shell_bf_num[M], dn = s->shell_bf_num[N];
                int r0 = s->shell_bf_sind[M], c0 = s->shell_bf_sind[N];
                ora_overlap(&s->shells[M], &s->shells[N], buf);
                for (int i = 0; i < dm; i++) for (int j = 0; j < dn; j++)
                    s->S[(size_t)(r0 + i) * nbf + c0 + j] = buf[i * dn + j];
                ora_kinetic(&s->shells[M], &s->shells[N], buf);
                ora_nuclear(&s->shells[M], &s->shells[N], s->natoms, s->Z, s->axyz, buf2);
                for (int i = 0; i < dm; i++) for (int j = 0; j < dn; j++)
                    s->Hcore[(size_t)(r0 + i) * nbf + c0 + j] = buf[i * dn + j] + buf2[i * dn + j];
            }
    }
    /* X = U s^-1/2 U^T (TinySCF.c:57-76); K is scratch as in the reference */
    double *U = s->tmp_mat, *U0 = s->K, *ev = s->eigval;
    memcpy(U, s->S, sizeof(double) * s->mat_size);
    ora_dsyev(nbf, U, ev);
    memcpy(U0, U, sizeof(double) * s->mat_size);
    for (int i = 0; i < nbf; i++) ev[i] = 1.0 / sqrt(ev[i]);
    for (int i = 0; i < nbf; i++) for (int j = 0; j < nbf; j++) U0[(size_t) i * nbf + j] *= ev[j];
    ora_dgemm(0, 1, nbf, nbf, nbf, 1.0, U0, nbf, U, nbf, 0.0, s->X, nbf);
    memset(s->K, 0, sizeof(double) * s->mat_size);
}

static int am_pair_index(const ora_scf_t *s, int P, int Q) { return s->shells[P].am * 5 + s->shells[Q].am; }

static int cmp_pair(int M1, int N1, int M2, int N2) { return (M1 == M2) ? (N1 < N2) : (M1 < M2); }

static void quicksort_pairs(int *M, int *N, int l, int r)     /* TinySCF.c:91-110 */
{
    int i = l, j = r, t;
    int mM = M[(i + j) / 2], mN = N[(i + j) / 2];
    while (i <= j) {
        while (cmp_pair(M[i], N[i], mM, mN)) i++;
        while (cmp_pair(mM, mN, M[j], N[j])) j--;
        if (i <= j) { t = M[i]; M[i] = M[j]; M[j] = t; t = N[i]; N[i] = N[j]; N[j] = t; i++; j--; }
    }
    if (i < r) quicksort_pairs(M, N, i, r);
    if (j > l) quicksort_pairs(M, N, l, j);
}

static void compute_schwarz(ora_scf_t *s)
{
    int nbf = s->nbf, ns = s->nshells;
    double gmax = 0.0;
    #pragma omp parallel for schedule(dynamic) reduction(max:gmax)
    for (int M = 0; M < ns; M++) {
        int dM = s->shell_bf_num[M], m0 = s->shell_bf_sind[M];
        for (int N = 0; N < ns; N++) {
            int dN = s->shell_bf_num[N], n0 = s->shell_bf_sind[N];
            double maxval = 0.0;
            for (int iM = 0; iM < dM; iM++)
                for (int iN = 0; iN < dN; iN++) {
                    int a, b, c, d;
                    if (s->quirk_a5) {
                        /* TinySCF.c:153 reads index iN*(dM*dN*dM+dM) + iM*(dN*dM+1) of the
                         * row-major (MN|MN) block [a][b][c][d] */
                        int idx = iN * (dM * dN * dM + dM) + iM * (dN * dM + 1);
                        d = idx % dN; idx /= dN;
                        c = idx % dM; idx /= dM;
                        b = idx % dN; idx /= dN;
                        a = idx;
                    } else { a = iM; b = iN; c = iM; d = iN; }
                    double v = fabs(ora_eri4_elem(&s->shells[M], &s->shells[N], &s->shells[M], &s->shells[N], a, b, c, d));
                    s->bf_pair_scrval[(size_t)(m0 + iM) * nbf + n0 + iN] = v;
                    if (v > maxval) maxval = v;
                }
            s->sp_scrval[(size_t) M * ns + N] = maxval;
            if (maxval > gmax) gmax = maxval;
        }
    }
    s->max_scrval = gmax;
    /* df screening value: max diagonal of (P|P) (Simint screen_max of the
     * (P,unit) pair; parity unpinned) */
    double gdf = 0.0;
    for (int P = 0; P < s->df_nshells; P++) {
        double buf[16 * 16];
        int dP = s->df_shell_bf_num[P];
        ora_eri2(&s->df_shells[P], &s->df_shells[P], buf);
        double m = 0.0;
        for (int k = 0; k < dP; k++) if (fabs(buf[k * dP + k]) > m) m = fabs(buf[k * dP + k]);
        s->df_sp_scrval[P] = m;
        if (m > gdf) gdf = m;
    }
    s->max_df_scrval = gdf;
    double eta = s->shell_scrtol2 / s->max_df_scrval;
    int nnz = 0;
    for (int M = 0; M < ns; M++)
        for (int N = 0; N < ns; N++) {
            /* a shell pair survives if either orientation passes: with quirk A5 the two
             * orientations can differ, and the union keeps the pair list symmetric */
            if (s->sp_scrval[(size_t) M * ns + N] > eta || s->sp_scrval[(size_t) N * ns + M] > eta) {
                if (N < M) continue;
                int MN_id = am_pair_index(s, M, N), NM_id = am_pair_index(s, N, M);
                if (MN_id > NM_id) { s->uniq_sp_lid[nnz] = M; s->uniq_sp_rid[nnz] = N; }
                else               { s->uniq_sp_lid[nnz] = N; s->uniq_sp_rid[nnz] = M; }
                nnz++;
            }
        }
    s->num_uniq_sp = nnz;
    if (nnz > 0) quicksort_pairs(s->uniq_sp_lid, s->uniq_sp_rid, 0, nnz - 1);
}

static void prepare_sparsity(ora_scf_t *s)
{
    int nbf = s->nbf, nnz = 0;
    double eta = s->shell_scrtol2 / s->max_df_scrval;
    s->bf_mask_displs[0] = 0;
    for (int i = 0; i < nbf; i++) {
        for (int j = 0; j < nbf; j++) {
            /* union of both orientations: the reference's result is undefined when only
             * one orientation survives (quirk A6 writes row -1), the union is symmetric */
            if (s->bf_pair_scrval[(size_t) i * nbf + j] > eta || s->bf_pair_scrval[(size_t) j * nbf + i] > eta) {
                s->bf_pair_mask[(size_t) i * nbf + j] = nnz;
                s->bf_pair_j[nnz] = j;
                nnz++;
            } else s->bf_pair_mask[(size_t) i * nbf + j] = -1;
        }
        s->bf_pair_diag[i] = s->bf_pair_mask[(size_t) i * nbf + i];
        s->bf_mask_displs[i + 1] = nnz;
    }
    s->num_bf_pair_scr = nnz;
    size_t tsz = (size_t) nnz * s->df_nbf;
    free(s->pqA); free(s->df_tensor); free(s->Cocc_tmp);
    s->pqA = (double *) xcalloc(tsz, sizeof(double));
    s->df_tensor = (double *) xcalloc(tsz, sizeof(double));
    s->Cocc_tmp = (double *) xcalloc((size_t) nnz * (s->n_occ > 0 ? s->n_occ : 1), sizeof(double));
}

void ora_d2cocc(ora_scf_t *s)
{
    int nbf = s->nbf, n_occ = s->n_occ, rank = 0;
    double *Chol = s->tmp_mat;
    int *piv = (int *) xcalloc(nbf, sizeof(int));
    memcpy(Chol, s->D, sizeof(double) * s->mat_size);
    ora_dpstrf_lower(nbf, Chol, piv, &rank, 1e-12);                 /* TinySCF.c:295 */
    if (rank < n_occ)
        for (int i = 0; i < nbf; i++) for (int j = rank; j < n_occ; j++) Chol[(size_t) i * nbf + j] = 0.0;
    if (s->quirk_a1) {
        /* pivot ignored and the diagonal dropped for rows < n_occ (TinySCF.c:304-318) */
        for (int i = 0; i < n_occ && i < nbf; i++) {
            for (int j = 0; j < i; j++) s->Cocc[(size_t) i * n_occ + j] = Chol[(size_t) i * nbf + j];
            for (int j = i; j < n_occ; j++) s->Cocc[(size_t) i * n_occ + j] = 0.0;
        }
        for (int i = n_occ; i < nbf; i++)
            memcpy(s->Cocc + (size_t) i * n_occ, Chol + (size_t) i * nbf, sizeof(double) * n_occ);
    } else {
        /* the factor of D itself: C_occ[piv[i]-1][j] = L[i][j], j <= i, j < min(rank,n_occ) */
        memset(s->Cocc, 0, sizeof(double) * (size_t) nbf * n_occ);
        for (int i = 0; i < nbf; i++)
            for (int j = 0; j < n_occ && j <= i && j < rank; j++)
                s->Cocc[(size_t)(piv[i] - 1) * n_occ + j] = Chol[(size_t) i * nbf + j];
    }
    free(piv);
}

static void initial_guess(ora_scf_t *s)
{
    memcpy(s->D, s->sad, sizeof(double) * s->mat_size);
    double R = 1.0;
    if (s->charge != 0 && s->electron != 0) R = (double)(s->electron - s->charge) / (double) s->electron;
    R *= 0.5;                                                   /* TinySCF.c:345-351 */
    for (int i = 0; i < s->mat_size; i++) s->D[i] *= R;
    ora_d2cocc(s);
}

int ora_setup(ora_scf_t *s)
{
    double t0 = wtime();
    compute_hcore_ovlp(s);
    compute_schwarz(s);
    prepare_sparsity(s);
    initial_guess(s);
    s->t_setup = wtime() - t0;
    return 0;
}

/* ---------------- DF tensor ---------------- */
int ora_build_df_tensor(ora_scf_t *s)
{
    int nbf = s->nbf, df_nbf = s->df_nbf, ns = s->nshells;
    double t0 = wtime();
    memset(s->pqA, 0, sizeof(double) * (size_t) s->num_bf_pair_scr * df_nbf);
    #pragma omp parallel
    {
        double *buf = (double *) xcalloc(6 * 6 * 28, sizeof(double));
        #pragma omp for schedule(dynamic)
        for (int iMN = 0; iMN < s->num_uniq_sp; iMN++) {
            int M = s->uniq_sp_lid[iMN], N = s->uniq_sp_rid[iMN];
            int sM = s->shell_bf_sind[M], eM = s->shell_bf_sind[M + 1];
            int sN = s->shell_bf_sind[N], eN = s->shell_bf_sind[N + 1];
            int dN = eN - sN;
            double scr0 = s->sp_scrval[(size_t) M * ns + N];
            for (int iAM = 0; iAM <= s->df_max_am; iAM++)
                for (int iP = s->df_am_shell_spos[iAM]; iP < s->df_am_shell_spos[iAM + 1]; iP++) {
                    int P = s->df_am_shell_id[iP];
                    if (scr0 * s->df_sp_scrval[P] < s->shell_scrtol2) continue;
                    int sP = s->df_shell_bf_sind[P], dP = s->df_shell_bf_num[P];
                    ora_eri3(&s->shells[M], &s->shells[N], &s->df_shells[P], buf);
                    for (int iM = sM; iM < eM; iM++)
                        for (int iN = sN; iN < eN; iN++) {
                            const double *e = buf + ((iM - sM) * dN + (iN - sN)) * dP;
                            int r0 = s->bf_pair_mask[(size_t) iM * nbf + iN];
                            int r1 = s->bf_pair_mask[(size_t) iN * nbf + iM];
                            if (r0 >= 0) memcpy(s->pqA + (size_t) r0 * df_nbf + sP, e, sizeof(double) * dP);
                            if (r1 >= 0) memcpy(s->pqA + (size_t) r1 * df_nbf + sP, e, sizeof(double) * dP);
                        }
                }
        }
        free(buf);
    }
    double t1 = wtime();
    memset(s->Jpq, 0, sizeof(double) * (size_t) df_nbf * df_nbf);
    #pragma omp parallel
    {
        double buf[16 * 16];
        #pragma omp for schedule(dynamic)
        for (int M = 0; M < s->df_nshells; M++)
            for (int N = M; N < s->df_nshells; N++) {
                if (s->df_sp_scrval[M] * s->df_sp_scrval[N] < s->shell_scrtol2) continue;
                ora_eri2(&s->df_shells[M], &s->df_shells[N], buf);
                int sM = s->df_shell_bf_sind[M], dM = s->df_shell_bf_num[M];
                int sN = s->df_shell_bf_sind[N], dN = s->df_shell_bf_num[N];
                for (int i = 0; i < dM; i++) for (int j = 0; j < dN; j++) {
                    double v = buf[i * dN + j];
                    s->Jpq[(size_t)(sM + i) * df_nbf + sN + j] = v;
                    s->Jpq[(size_t)(sN + j) * df_nbf + sM + i] = v;
                }
            }
    }
    double t2 = wtime();
    {
        size_t sz = (size_t) df_nbf * df_nbf;
        double *U = (double *) xcalloc(sz, sizeof(double)), *U1 = (double *) xcalloc(sz, sizeof(double));
        double *ev = (double *) xcalloc(df_nbf, sizeof(double));
        memcpy(U, s->Jpq, sizeof(double) * sz);
        ora_dsyev(df_nbf, U, ev);
        for (int i = 0; i < df_nbf; i++) ev[i] = 1.0 / sqrt(ev[i]);   /* quirk A9: no floor */
        #pragma omp parallel for
        for (int r = 0; r < df_nbf; r++) {
            memcpy(U1 + (size_t) r * df_nbf, U + (size_t) r * df_nbf, sizeof(double) * df_nbf);
            for (int c = 0; c < df_nbf; c++) U[(size_t) r * df_nbf + c] *= ev[c];
        }
        ora_dgemm(0, 1, df_nbf, df_nbf, df_nbf, 1.0, U, df_nbf, U1, df_nbf, 0.0, s->Jpq, df_nbf);
        free(U); free(U1); free(ev);
    }
    double t3 = wtime();
    ora_dgemm(0, 0, s->num_bf_pair_scr, df_nbf, df_nbf, 1.0, s->pqA, df_nbf, s->Jpq, df_nbf, 0.0, s->df_tensor, df_nbf);
    double t4 = wtime();
    s->t_3c = t1 - t0; s->t_2c = t2 - t1; s->t_invsqrt = t3 - t2; s->t_dfgemm = t4 - t3;
    return 0;
}

/* ---------------- Fock ---------------- */
static void build_J(ora_scf_t *s)
{
    int nbf = s->nbf, df_nbf = s->df_nbf, nt = s->nthreads;
    double t0 = wtime(), t1 = 0;
    memset(s->J, 0, sizeof(double) * s->mat_size);               /* quirk A7: zeroed */
    #pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num(), nth = omp_get_num_threads();
        double *tj = s->temp_J + (size_t) s->df_nbf_16 * tid;
        memset(tj, 0, sizeof(double) * df_nbf);
        #pragma omp for schedule(dynamic)
        for (int k = 0; k < nbf; k++) {
            int dk = s->bf_pair_diag[k];
            const double *row = s->df_tensor + (size_t) dk * df_nbf;
            double Dkl = s->D[(size_t) k * nbf + k];
            for (int p = 0; p < df_nbf; p++) tj[p] += Dkl * row[p];
            for (int l_idx = dk + 1; l_idx < s->bf_mask_displs[k + 1]; l_idx++) {
                int l = s->bf_pair_j[l_idx];
                double w = 2.0 * s->D[(size_t) k * nbf + l];
                const double *r2 = s->df_tensor + (size_t) l_idx * df_nbf;
                for (int p = 0; p < df_nbf; p++) tj[p] += w * r2[p];
            }
        }
        /* tree reduction (build_Fock.c:15-33) */
        int n = nth;
        while (n > 1) {
            int mid = (n + 1) / 2, act = n / 2;
            if (tid < act) {
                double *src = s->temp_J + (size_t) s->df_nbf_16 * (tid + mid);
                for (int p = 0; p < df_nbf; p++) tj[p] += src[p];
            }
            #pragma omp barrier
            n = mid;
        }
        #pragma omp master
        t1 = wtime();
        #pragma omp barrier
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < nbf; i++) {
            for (int j_idx = s->bf_pair_diag[i]; j_idx < s->bf_mask_displs[i + 1]; j_idx++) {
                int j = s->bf_pair_j[j_idx];
                const double *row = s->df_tensor + (size_t) j_idx * df_nbf;
                double t = 0.0;
                for (int p = 0; p < df_nbf; p++) t += s->temp_J[p] * row[p];
                s->J[(size_t) i * nbf + j] = t;
            }
        }
    }
    double t2 = wtime();
    s->t_tempJ = t1 - t0; s->t_Jmat = t2 - t1;
}

static void build_K(ora_scf_t *s)
{
    int nbf = s->nbf, df_nbf = s->df_nbf, n_occ = s->n_occ;
    double t0 = wtime();
    int nnz = s->num_bf_pair_scr;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < nnz; r++)
        memcpy(s->Cocc_tmp + (size_t) r * n_occ, s->Cocc + (size_t) s->bf_pair_j[r] * n_occ, sizeof(double) * n_occ);
    /* temp_K(i,:,:) = Cocc_tmp(rows of i)^T * df_tensor(rows of i)  (build_Fock.c:276-285) */
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nbf; i++) {
        int rs = s->bf_mask_displs[i], re = s->bf_mask_displs[i + 1];
        ora_dgemm(1, 0, n_occ, df_nbf, re - rs, 1.0, s->Cocc_tmp + (size_t) rs * n_occ, n_occ,
                  s->df_tensor + (size_t) rs * df_nbf, df_nbf, 0.0,
                  s->temp_K + (size_t) i * n_occ * df_nbf, df_nbf);
    }
    double t1 = wtime();
    int kd = n_occ * df_nbf;
    if (nbf <= 1024) {
        ora_dgemm(0, 1, nbf, nbf, kd, 1.0, s->temp_K, kd, s->temp_K, kd, 0.0, s->K, nbf);
    } else {
        /* upper block tiles, BS = max(nbf/10, 64) (TinySCF_init_free.c:278-279, build_Fock.c:327-341) */
        int BS = nbf / 10; if (BS < 64) BS = 64;
        int nb = (nbf + BS - 1) / BS, ntile = nb * (nb + 1) / 2;
        #pragma omp parallel for schedule(dynamic)
        for (int t = 0; t < ntile; t++) {
            int bi = 0, rem = t;
            while (rem >= nb - bi) { rem -= nb - bi; bi++; }
            int bj = bi + rem;
            int i0 = bi * BS, j0 = bj * BS;
            int il = (nbf - i0 < BS) ? nbf - i0 : BS, jl = (nbf - j0 < BS) ? nbf - j0 : BS;
            ora_dgemm(0, 1, il, jl, kd, 1.0, s->temp_K + (size_t) i0 * kd, kd,
                      s->temp_K + (size_t) j0 * kd, kd, 0.0, s->K + (size_t) i0 * nbf + j0, nbf);
        }
    }
    double t2 = wtime();
    s->t_tempK = t1 - t0; s->t_Kmat = t2 - t1;
}

int ora_build_fock(ora_scf_t *s)
{
    int nbf = s->nbf;
    double t0 = wtime();
    build_J(s);
    build_K(s);
    double t1 = wtime();
    #pragma omp parallel
    {
        #pragma omp for schedule(dynamic)
        for (int i = 1; i < nbf; i++)
            for (int j = 0; j < i; j++) {
                s->J[(size_t) i * nbf + j] = s->J[(size_t) j * nbf + i];
                s->K[(size_t) i * nbf + j] = s->K[(size_t) j * nbf + i];
            }
        #pragma omp for
        for (int i = 0; i < s->mat_size; i++) s->F[i] = s->Hcore[i] + 2.0 * s->J[i] - s->K[i];
    }
    double t2 = wtime();
    s->t_symm = t2 - t1; s->t_fock = t2 - t0;
    return 0;
}

double ora_calc_energy(ora_scf_t *s)
{
    double e = 0.0;
    for (int i = 0; i < s->mat_size; i++) e += s->D[i] * (s->F[i] + s->Hcore[i]);
    s->HF_energy = e;
    return e;
}

int ora_diis(ora_scf_t *s)
{
    int nbf = s->nbf, ms = s->mat_size, ldB = MAX_DIIS + 1;
    double *tmp = s->tmp_mat;
    if (s->iter <= 1) {                                              /* DIIS.c:30-40 (A3) */
        ora_dgemm(1, 0, nbf, nbf, nbf, 1.0, s->X, nbf, s->F, nbf, 0.0, tmp, nbf);
        ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, s->X, nbf, 0.0, s->F, nbf);
        return 0;
    }
    int idx;
    if (s->DIIS_len < MAX_DIIS) { idx = s->DIIS_len; s->DIIS_len++; }
    else idx = s->DIIS_bmax_id;
    ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, s->F, nbf, s->D, nbf, 0.0, tmp, nbf);
    ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, s->S, nbf, 0.0, s->FDS, nbf);
    for (int i = 0; i < nbf; i++) for (int j = 0; j < nbf; j++)
        tmp[(size_t) i * nbf + j] = s->FDS[(size_t) i * nbf + j] - s->FDS[(size_t) j * nbf + i];
    ora_dgemm(1, 0, nbf, nbf, nbf, 1.0, s->X, nbf, tmp, nbf, 0.0, s->FDS, nbf);
    ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, s->FDS, nbf, s->X, nbf, 0.0, tmp, nbf);
    double dot[MAX_DIIS + 1];
    memset(dot, 0, sizeof(dot));
    memcpy(s->R + (size_t) ms * idx, tmp, sizeof(double) * ms);
    const double *Ri = s->R + (size_t) ms * idx;
    for (int j = 0; j < s->DIIS_len; j++) {
        const double *Rj = s->R + (size_t) ms * j;
        double d = 0.0;
        for (int q = 0; q < ms; q++) d += Ri[q] * Rj[q];
        dot[j] = d;
    }
    for (int i = 0; i < s->DIIS_len; i++) { s->B[idx * ldB + i] = dot[i]; s->B[i * ldB + idx] = dot[i]; }
    if (!s->quirk_a2) s->DIIS_bmax = -DBL_MAX;                      /* A2 off: recomputed */
    for (int i = 0; i < s->DIIS_len; i++)                            /* A2: bmax never reset */
        if (s->B[i * ldB + i] > s->DIIS_bmax) { s->DIIS_bmax = s->B[i * ldB + i]; s->DIIS_bmax_id = i; }
    ora_dgemm(1, 0, nbf, nbf, nbf, 1.0, s->X, nbf, s->F, nbf, 0.0, tmp, nbf);
    ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, s->X, nbf, 0.0, s->F, nbf);
    memcpy(s->F0 + (size_t) ms * idx, s->F, sizeof(double) * ms);
    memset(s->DIIS_rhs, 0, sizeof(double) * (MAX_DIIS + 1));
    s->DIIS_rhs[s->DIIS_len] = -1.0;
    double Bc[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    memcpy(Bc, s->B, sizeof(Bc));
    ora_dgesv(s->DIIS_len + 1, Bc, ldB, s->DIIS_ipiv, s->DIIS_rhs);
    memset(s->F, 0, sizeof(double) * ms);
    for (int i = 0; i < s->DIIS_len; i++) {
        double c = s->DIIS_rhs[i];
        const double *f0 = s->F0 + (size_t) i * ms;
        for (int q = 0; q < ms; q++) s->F[q] += c * f0[q];
    }
    return 0;
}

int ora_build_density(ora_scf_t *s)
{
    int nbf = s->nbf, n_occ = s->n_occ;
    memcpy(s->tmp_mat, s->F, sizeof(double) * s->mat_size);
    ora_dsyev(nbf, s->tmp_mat, s->eigval);
    ora_dgemm(0, 0, nbf, nbf, nbf, 1.0, s->X, nbf, s->tmp_mat, nbf, 0.0, s->D, nbf);
    /* A4: the eigenvalues come back ascending; a stable argsort keeps the identity order */
    for (int i = 0; i < nbf; i++) s->ev_idx[i] = i;
    for (int i = 1; i < nbf; i++) {
        int j = i, t = s->ev_idx[i];
        double v = s->eigval[i];
        while (j > 0 && s->eigval[j - 1] > v) { s->eigval[j] = s->eigval[j - 1]; s->ev_idx[j] = s->ev_idx[j - 1]; j--; }
        s->eigval[j] = v; s->ev_idx[j] = t;
    }
    for (int j = 0; j < n_occ; j++)
        for (int i = 0; i < nbf; i++)
            s->Cocc[(size_t) i * n_occ + j] = s->D[(size_t) i * nbf + s->ev_idx[j]];
    ora_dgemm(0, 1, nbf, nbf, n_occ, 1.0, s->Cocc, n_occ, s->Cocc, n_occ, 0.0, s->D, nbf);
    return 0;
}

int ora_do_scf(ora_scf_t *s, int verbose)
{
    double t0 = wtime();
    s->iter = 0;
    double prev = 0.0, delta = 223.0;
    s->t_energy = s->t_diis = s->t_density = 0.0;
    s->converged = 0;
    while (s->iter < s->niters && delta >= s->ene_tol) {
        ora_build_fock(s);
        double ta = wtime();
        ora_calc_energy(s);
        double tb = wtime();
        delta = fabs(s->HF_energy - prev);
        prev = s->HF_energy;
        ora_diis(s);
        double tc = wtime();
        ora_build_density(s);
        double td = wtime();
        s->t_energy += tb - ta; s->t_diis += tc - tb; s->t_density += td - tc;
        if (verbose)
            printf("oracle iter %3d  E = %.10f  dE = %.3e  J %.4f K %.4f s\n", s->iter,
                   s->HF_energy + s->nuc_energy, s->iter > 0 ? delta : 0.0,
                   s->t_tempJ + s->t_Jmat, s->t_tempK + s->t_Kmat);
        if (s->iter == 0) delta = 223.0;
        s->iter++;
    }
    s->energy_delta = delta;
    s->converged = (delta < s->ene_tol);
    s->t_scf_total = wtime() - t0;
    return s->converged ? 0 : 1;
}

/* ---------------- accessors ---------------- */
double *ora_get_dptr(ora_scf_t *s, const char *n, size_t *count)
{
    size_t ms = (size_t) s->mat_size;
#define RET(name, ptr, cnt) if (!strcmp(n, name)) { if (count) *count = (cnt); return (ptr); }
    RET("Hcore_mat", s->Hcore, ms) RET("S_mat", s->S, ms) RET("F_mat", s->F, ms)
    RET("D_mat", s->D, ms) RET("J_mat", s->J, ms) RET("K_mat", s->K, ms) RET("X_mat", s->X, ms)
    RET("Cocc_mat", s->Cocc, (size_t) s->nbf * s->n_occ) RET("eigval", s->eigval, (size_t) s->nbf)
    RET("tmp_mat", s->tmp_mat, ms)
    RET("sp_scrval", s->sp_scrval, (size_t) s->nshells * s->nshells)
    RET("bf_pair_scrval", s->bf_pair_scrval, ms)
    RET("df_sp_scrval", s->df_sp_scrval, (size_t) s->df_nshells)
    RET("pqA", s->pqA, (size_t) s->num_bf_pair_scr * s->df_nbf)
    RET("df_tensor", s->df_tensor, (size_t) s->num_bf_pair_scr * s->df_nbf)
    RET("Jpq", s->Jpq, (size_t) s->df_nbf * s->df_nbf)
    RET("temp_J", s->temp_J, (size_t) s->df_nbf)
    RET("temp_K", s->temp_K, (size_t) s->nbf * s->n_occ * s->df_nbf)
    RET("B_mat", s->B, (size_t)(MAX_DIIS + 1) * (MAX_DIIS + 1))
    RET("R_mat", s->R, ms * MAX_DIIS)
#undef RET
    if (count) *count = 0;
    return NULL;
}

int *ora_get_iptr(ora_scf_t *s, const char *n, size_t *count)
{
#define RET(name, ptr, cnt) if (!strcmp(n, name)) { if (count) *count = (cnt); return (ptr); }
    RET("bf_pair_mask", s->bf_pair_mask, (size_t) s->mat_size)
    RET("bf_pair_j", s->bf_pair_j, (size_t) s->num_bf_pair_scr)
    RET("bf_pair_diag", s->bf_pair_diag, (size_t) s->nbf)
    RET("bf_mask_displs", s->bf_mask_displs, (size_t) s->nbf + 1)
    RET("uniq_sp_lid", s->uniq_sp_lid, (size_t) s->num_uniq_sp)
    RET("uniq_sp_rid", s->uniq_sp_rid, (size_t) s->num_uniq_sp)
    RET("shell_bf_sind", s->shell_bf_sind, (size_t) s->nshells + 1)
    RET("df_shell_bf_sind", s->df_shell_bf_sind, (size_t) s->df_nshells + 1)
    RET("ev_idx", s->ev_idx, (size_t) s->nbf)
#undef RET
    if (count) *count = 0;
    return NULL;
}

int ora_get_int(ora_scf_t *s, const char *n)
{
    if (!strcmp(n, "nbf")) return s->nbf;
    if (!strcmp(n, "df_nbf")) return s->df_nbf;
    if (!strcmp(n, "nshells")) return s->nshells;
    if (!strcmp(n, "df_nshells")) return s->df_nshells;
    if (!strcmp(n, "n_occ")) return s->n_occ;
    if (!strcmp(n, "iter")) return s->iter;
    if (!strcmp(n, "num_uniq_sp")) return s->num_uniq_sp;
    if (!strcmp(n, "num_bf_pair_scr")) return s->num_bf_pair_scr;
    if (!strcmp(n, "DIIS_len")) return s->DIIS_len;
    if (!strcmp(n, "DIIS_bmax_id")) return s->DIIS_bmax_id;
    if (!strcmp(n, "converged")) return s->converged;
    if (!strcmp(n, "nthreads")) return s->nthreads;
    if (!strcmp(n, "mkl")) return ora_lapack_is_mkl();
    return -1;
}

void ora_set_int(ora_scf_t *s, const char *n, int v)
{
    if (!strcmp(n, "iter")) s->iter = v;
    else if (!strcmp(n, "niters")) s->niters = v;
}

double ora_get_double(ora_scf_t *s, const char *n)
{
#define RET(name, v) if (!strcmp(n, name)) return (v);
    RET("HF_energy", s->HF_energy) RET("nuc_energy", s->nuc_energy)
    RET("total_energy", s->HF_energy + s->nuc_energy) RET("energy_delta", s->energy_delta)
    RET("max_scrval", s->max_scrval) RET("max_df_scrval", s->max_df_scrval) RET("DIIS_bmax", s->DIIS_bmax)
    RET("t_setup", s->t_setup) RET("t_3c", s->t_3c) RET("t_2c", s->t_2c)
    RET("t_invsqrt", s->t_invsqrt) RET("t_dfgemm", s->t_dfgemm)
    RET("t_tempJ", s->t_tempJ) RET("t_Jmat", s->t_Jmat) RET("t_tempK", s->t_tempK)
    RET("t_Kmat", s->t_Kmat) RET("t_symm", s->t_symm) RET("t_fock", s->t_fock)
    RET("t_energy", s->t_energy) RET("t_diis", s->t_diis) RET("t_density", s->t_density)
    RET("t_scf_total", s->t_scf_total)
#undef RET
    return NAN;
}
