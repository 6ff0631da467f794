/*
 * TinySCF.c -- setup and SCF driver (drop-in for the reference's
 * TinySCF.c:20-435).  Integrals and eigensolves go to the device engine;
 * the screening logic, the pair layout, the SAD guess, the energy and the
 * iteration loop stay on the host as in the reference.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "DIIS.h"
#include "TinySCF.h"
#include "build_Fock.h"
#include "build_density.h"
#include "guess_ml.h"
#include "utils.h"
#include "yatscf_gpu.h"

/* ---- spherical orbital basis: the real solid harmonics among the Cartesian components
 * (ygpu_sph_coefficients gives T_l per degree l) */
static double ys_dfact(int n)
{
    double r = 1.0;
    for (int i = n; i > 1; i -= 2) r *= i;
    return r;
}

/* P_l = T_l T_l^T M_l (ncart x ncart, row-major), M_l the overlap of the x^l-normalised
 * components of one shell (angular factor (a+a'-1)!!(b+b'-1)!!(c+c'-1)!!/(2l-1)!!): the
 * projector, in that metric, of a shell's Cartesian coefficients onto its solid harmonics */
static void sph_projector(int l, double *P)
{
    const int nc = (l + 1) * (l + 2) / 2, ns = 2 * l + 1;
    double Tl[15 * 9], M[15 * 15], TT[15 * 15];
    int pa[15], pb[15], pc[15], i = 0;
    ygpu_sph_coefficients(l, Tl);
    for (int a = l; a >= 0; a--)
        for (int b = l - a; b >= 0; b--, i++) {
            pa[i] = a;
            pb[i] = b;
            pc[i] = l - a - b;
        }
    for (i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            const int x = pa[i] + pa[j], y = pb[i] + pb[j], z = pc[i] + pc[j];
            M[i * nc + j] = ((x | y | z) & 1) ? 0.0 : ys_dfact(x - 1) * ys_dfact(y - 1) * ys_dfact(z - 1) / ys_dfact(2 * l - 1);
            double t = 0.0;
            for (int q = 0; q < ns; q++) t += Tl[i * ns + q] * Tl[j * ns + q];
            TT[i * nc + j] = t;
        }
    for (i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) {
            double t = 0.0;
            for (int k = 0; k < nc; k++) t += TT[i * nc + k] * M[k * nc + j];
            P[i * nc + j] = t;
        }
}

/* A <- P A P^T with P block-diagonal over the shells (P_l of each shell of degree >= 2) */
static void sph_project_sym(TinySCF_t T, double *A)
{
    const int n = T->nbasfuncs;
    const int *am, *np, *po;
    const double *xyz, *al, *co;
    double P[5][15 * 15], v[15];
    for (int l = 2; l <= 4; l++) sph_projector(l, P[l]);
    CMS_getShellArrays(T->basis, &am, &np, &po, &xyz, &al, &co);
    for (int s = 0; s < T->nshells; s++) {
        const int l = am[s], nc = (l + 1) * (l + 2) / 2, f0 = T->shell_bf_sind[s];
        if (l < 2) continue;
        for (int j = 0; j < n; j++) {          /* rows of the shell: P A */
            for (int c = 0; c < nc; c++) v[c] = A[(size_t) (f0 + c) * n + j];
            for (int c = 0; c < nc; c++) {
                double t = 0.0;
                for (int k = 0; k < nc; k++) t += P[l][c * nc + k] * v[k];
                A[(size_t) (f0 + c) * n + j] = t;
            }
        }
        for (int i = 0; i < n; i++) {          /* columns of the shell: A P^T */
            double *row = A + (size_t) i * n + f0;
            for (int c = 0; c < nc; c++) v[c] = row[c];
            for (int c = 0; c < nc; c++) {
                double t = 0.0;
                for (int k = 0; k < nc; k++) t += P[l][c * nc + k] * v[k];
                row[c] = t;
            }
        }
    }
}

/* X = T (T^T S T)^{-1/2} (nbf x nbf_sph) in the first nbf_sph columns of X_mat, zero beyond */
static void sph_orthogonaliser(TinySCF_t T)
{
    ygpu_ctx_t *g = (ygpu_ctx_t *) T->gpu;
    const int n = T->nbasfuncs, m = T->nbf_sph;
    const int *am, *np, *po;
    const double *xyz, *al, *co;
    double Tl[15 * 9];
    double *Ts = (double *) calloc((size_t) n * m, DBL_SIZE), *W = (double *) malloc(DBL_SIZE * (size_t) n * m),
           *Ss = (double *) malloc(DBL_SIZE * (size_t) m * m), *Xs = (double *) malloc(DBL_SIZE * (size_t) m * m);
    if (!Ts || !W || !Ss || !Xs) {
        free(Ts); free(W); free(Ss); free(Xs);
        ys_check_gpu(YGPU_ERR_NOMEM, "spherical orthogonaliser (host allocation)");
    }
    CMS_getShellArrays(T->basis, &am, &np, &po, &xyz, &al, &co);
    for (int s = 0, q0 = 0; s < T->nshells; s++) {
        const int l = am[s], nc = (l + 1) * (l + 2) / 2, ns = 2 * l + 1, f0 = T->shell_bf_sind[s];
        ygpu_sph_coefficients(l, Tl);
        for (int c = 0; c < nc; c++)
            for (int q = 0; q < ns; q++) Ts[(size_t) (f0 + c) * m + q0 + q] = Tl[c * ns + q];
        q0 += ns;
    }
    YS_GPU(ygpu_dgemm(g, 0, 0, n, m, n, 1.0, T->S_mat, n, Ts, m, 0.0, W, m));      /* S T */
    YS_GPU(ygpu_dgemm(g, 1, 0, m, m, n, 1.0, Ts, m, W, m, 0.0, Ss, m));            /* T^T S T */
    YS_GPU(ygpu_sym_inv_sqrt(g, m, Ss, Xs, NULL));
    YS_GPU(ygpu_dgemm(g, 0, 0, n, m, m, 1.0, Ts, m, Xs, m, 0.0, W, m));            /* T (T^T S T)^{-1/2} */
    for (int i = 0; i < n; i++) {
        memcpy(T->X_mat + (size_t) i * n, W + (size_t) i * m, DBL_SIZE * m);
        memset(T->X_mat + (size_t) i * n + m, 0, DBL_SIZE * (n - m));
    }
    YS_GPU(ygpu_set_orth_dim(g, m));
    free(Ts); free(W); free(Ss); free(Xs);
}

void TinySCF_compute_Hcore_Ovlp_mat(TinySCF_t T)
{
    double st = get_wtime_sec();
    ygpu_ctx_t *g = (ygpu_ctx_t *) T->gpu;
    YS_GPU(ygpu_ovlp_hcore(g, T->S_mat, T->Hcore_mat));
    /* X = S^{-1/2} = U s^{-1/2} U^T (TinySCF.c:57-76); a spherical basis confines X's columns
     * to the solid harmonics */
    if (T->sph_orb) sph_orthogonaliser(T);
    else YS_GPU(ygpu_sym_inv_sqrt(g, T->nbasfuncs, T->S_mat, T->X_mat, NULL));
    YS_GPU(ygpu_set_overlap(g, T->S_mat, T->X_mat));
    T->S_Hcore_time = get_wtime_sec() - st;
    if (T->verbose)
        printf("TinySCF precompute Hcore, S, and X matrices over,  elapsed time = %.3lf (s)\n", T->S_Hcore_time);
}

static int before(int M1, int N1, int M2, int N2) { return M1 == M2 ? N1 < N2 : M1 < M2; }

/* quicksort of (M,N) pairs by M then N (TinySCF.c:91-110) */
static void sort_pairs(int *M, int *N, int l, int r)
{
    int i = l, j = r;
    const int pm = M[(l + r) / 2], pn = N[(l + r) / 2];
    while (i <= j) {
        while (before(M[i], N[i], pm, pn)) i++;
        while (before(pm, pn, M[j], N[j])) j--;
        if (i <= j) {
            int t = M[i]; M[i] = M[j]; M[j] = t;
            t = N[i]; N[i] = N[j]; N[j] = t;
            i++;
            j--;
        }
    }
    if (l < j) sort_pairs(M, N, l, j);
    if (i < r) sort_pairs(M, N, i, r);
}

void TinySCF_compute_sq_Schwarz_scrvals(TinySCF_t T)
{
    double st = get_wtime_sec();
    const int ns = T->nshells;
    YS_GPU(ygpu_schwarz((ygpu_ctx_t *) T->gpu, T->quirk_a5, T->sp_scrval, T->bf_pair_scrval, T->df_sp_scrval));
    double gmax = 0.0, dfmax = 0.0;
    for (int i = 0; i < ns * ns; i++) if (T->sp_scrval[i] > gmax) gmax = T->sp_scrval[i];
    for (int i = 0; i < T->df_nshells; i++) if (T->df_sp_scrval[i] > dfmax) dfmax = T->df_sp_scrval[i];
    T->max_scrval = gmax;
    T->max_df_scrval = dfmax;
    /* unique shell pairs surviving eta = tol^2 / max_P (P|P), AM(lid) >= AM(rid)
     * (TinySCF.c:183-221); either orientation passing keeps the pair, so the list
     * stays symmetric when the A5 index makes the two orientations differ */
    const double eta = T->shell_scrtol2 / T->max_df_scrval;
    int nnz = 0;
    for (int M = 0; M < ns; M++)
        for (int N = M; N < ns; N++) {
            if (!(T->sp_scrval[M * ns + N] > eta || T->sp_scrval[N * ns + M] > eta)) continue;
            const int amM = CMS_getShellDim(T->basis, M), amN = CMS_getShellDim(T->basis, N);
            if (amM > amN) { T->uniq_sp_lid[nnz] = M; T->uniq_sp_rid[nnz] = N; }
            else           { T->uniq_sp_lid[nnz] = N; T->uniq_sp_rid[nnz] = M; }
            nnz++;
        }
    T->num_uniq_sp = nnz;
    if (nnz > 0) sort_pairs(T->uniq_sp_lid, T->uniq_sp_rid, 0, nnz - 1);
    T->shell_scr_time = get_wtime_sec() - st;
    if (T->verbose)
        printf("TinySCF precompute shell screening info over,      elapsed time = %.3lf (s)\n", T->shell_scr_time);
}

void TinySCF_prepare_sparsity(TinySCF_t T)
{
    double st = get_wtime_sec();
    const int nbf = T->nbasfuncs;
    const double eta = T->shell_scrtol2 / T->max_df_scrval;
    int nnz = 0;
    T->bf_mask_displs[0] = 0;
    for (int i = 0; i < nbf; i++) {
        for (int j = 0; j < nbf; j++) {
            /* union of both orientations (see TinySCF_compute_sq_Schwarz_scrvals) */
            if (T->bf_pair_scrval[(size_t) i * nbf + j] > eta || T->bf_pair_scrval[(size_t) j * nbf + i] > eta) {
                T->bf_pair_mask[(size_t) i * nbf + j] = nnz;
                T->bf_pair_j[nnz] = j;
                nnz++;
            } else {
                T->bf_pair_mask[(size_t) i * nbf + j] = -1;
            }
        }
        T->bf_pair_diag[i] = T->bf_pair_mask[(size_t) i * nbf + i];
        T->bf_mask_displs[i + 1] = nnz;
    }
    T->num_bf_pair_scr = nnz;
    YS_GPU(ygpu_set_sparsity((ygpu_ctx_t *) T->gpu, T->bf_pair_mask, T->num_uniq_sp, T->uniq_sp_lid, T->uniq_sp_rid,
                             T->sp_scrval, T->df_sp_scrval, T->shell_scrtol2));
    if (T->host_df_copy) {
        size_t tsz = (size_t) nnz * T->df_nbf * DBL_SIZE;
        T->pqA = (double *) ALIGN64B_MALLOC(tsz);
        T->df_tensor = (double *) ALIGN64B_MALLOC(tsz);
        if (!T->pqA || !T->df_tensor) {
            fprintf(stderr, "TinySCF: host DF tensor copies (%zu bytes each) could not be allocated\n", tsz);
            abort();
        }
        T->mem_size += 2.0 * (double) tsz;
    }
    if (T->verbose) {
        printf("TinySCF handling shell pair sparsity over,         elapsed time = %.3lf (s)\n", get_wtime_sec() - st);
        printf("TinySCF regular + density fitting memory usage = %.2lf MB \n", T->mem_size / 1048576.0);
        printf("#### Sparsity of basis function pairs = %lf, %lf\n", (double) T->num_uniq_sp / T->nshellpairs,
               (double) nnz / T->mat_size);
    }
}

/*
 * Pivoted Cholesky D = P L L^T P^T, lower, row-major, with LAPACK dpstf2
 * semantics (first maximal diagonal is the pivot, stop when it is <= tol).
 * On return A's lower triangle holds L in pivoted order, piv is 1-based.
 */
static int pivoted_cholesky(int n, double *A, int *piv, double tol)
{
    double *work = (double *) calloc(n, sizeof(double));
    for (int i = 0; i < n; i++) piv[i] = i + 1;
    int rank = n;
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = -HUGE_VAL;
        for (int i = j; i < n; i++) {
            if (j > 0) work[i] += A[(size_t) i * n + j - 1] * A[(size_t) i * n + j - 1];
            const double v = A[(size_t) i * n + i] - work[i];
            if (v > best) { best = v; p = i; }
        }
        if (!(best > tol)) { A[(size_t) j * n + j] = best; rank = j; break; }
        if (p != j) {
            A[(size_t) p * n + p] = A[(size_t) j * n + j];
            for (int c = 0; c < j; c++) {
                double t = A[(size_t) j * n + c]; A[(size_t) j * n + c] = A[(size_t) p * n + c]; A[(size_t) p * n + c] = t;
            }
            for (int r = p + 1; r < n; r++) {
                double t = A[(size_t) r * n + j]; A[(size_t) r * n + j] = A[(size_t) r * n + p]; A[(size_t) r * n + p] = t;
            }
            for (int c = j + 1; c < p; c++) {
                double t = A[(size_t) c * n + j]; A[(size_t) c * n + j] = A[(size_t) p * n + c]; A[(size_t) p * n + c] = t;
            }
            double t = work[j]; work[j] = work[p]; work[p] = t;
            int ip = piv[j]; piv[j] = piv[p]; piv[p] = ip;
        }
        const double ajj = sqrt(best);
        A[(size_t) j * n + j] = ajj;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t) i * n + j];
            for (int c = 0; c < j; c++) s -= A[(size_t) i * n + c] * A[(size_t) j * n + c];
            A[(size_t) i * n + j] = s / ajj;
        }
    }
    free(work);
    return rank;
}

void TinySCF_D2Cocc(TinySCF_t T)
{
    const int nbf = T->nbasfuncs, n_occ = T->n_occ;
    double *L = T->tmp_mat;
    int *piv = (int *) malloc(sizeof(int) * nbf);
    memcpy(L, T->D_mat, DBL_SIZE * T->mat_size);
    const int rank = pivoted_cholesky(nbf, L, piv, 1e-12);            /* TinySCF.c:295 */
    if (rank < n_occ)
        for (int i = 0; i < nbf; i++)
            for (int j = rank; j < n_occ; j++) L[(size_t) i * nbf + j] = 0.0;
    if (T->quirk_a1) {
        /* the reference ignores the pivot and drops the diagonal of rows < n_occ
         * (TinySCF.c:304-318): this C_occ only feeds the first K build */
        for (int i = 0; i < n_occ && i < nbf; i++) {
            for (int j = 0; j < i; j++) T->Cocc_mat[(size_t) i * n_occ + j] = L[(size_t) i * nbf + j];
            for (int j = i; j < n_occ; j++) T->Cocc_mat[(size_t) i * n_occ + j] = 0.0;
        }
        for (int i = n_occ; i < nbf; i++) memcpy(T->Cocc_mat + (size_t) i * n_occ, L + (size_t) i * nbf, DBL_SIZE * n_occ);
    } else {
        memset(T->Cocc_mat, 0, DBL_SIZE * (size_t) nbf * n_occ);
        for (int i = 0; i < nbf; i++)
            for (int j = 0; j < n_occ && j <= i && j < rank; j++)
                T->Cocc_mat[(size_t) (piv[i] - 1) * n_occ + j] = L[(size_t) i * nbf + j];
    }
    free(piv);
}

void TinySCF_get_initial_guess(TinySCF_t T)
{
    const int nbf = T->nbasfuncs;
    memset(T->D_mat, 0, DBL_SIZE * T->mat_size);
    double t0 = get_wtime_sec();
    for (int a = 0; a < T->natoms; a++) {
        double *guess;
        int spos, epos;
        CMS_getInitialGuess(T->basis, a, &guess, &spos, &epos);
        const int ldg = epos - spos + 1;
        copy_matrix_block(T->D_mat + (size_t) spos * nbf + spos, nbf, guess, ldg, ldg, ldg);
    }
    double R = 1.0;
    if (T->charge != 0 && T->electron != 0) R = (double) (T->electron - T->charge) / (double) T->electron;
    R *= 0.5;
    for (int i = 0; i < T->mat_size; i++) T->D_mat[i] *= R;
    /* a spherical basis keeps the solid-harmonic part of each shell's block */
    if (T->sph_orb) sph_project_sym(T, T->D_mat);
    T->nuc_energy = CMS_getNucEnergy(T->basis);
    double t1 = get_wtime_sec();
    T->guess_ml = 0;
    if (T->guess_weights[0]) {
        /* learned pair correction, which forms its own C_occ (the occupied eigenvectors of the
         * corrected density); on any failure the SAD guess stays and D2Cocc below factors it */
        double *sad = (double *) malloc(DBL_SIZE * T->mat_size);
        if (sad) {
            memcpy(sad, T->D_mat, DBL_SIZE * T->mat_size);
            const int st = TinySCF_learned_guess(T, T->guess_weights);
            T->guess_ml = (st == 0);
            if (!T->guess_ml) memcpy(T->D_mat, sad, DBL_SIZE * T->mat_size);
            free(sad);
            if (T->verbose)
                printf("TinySCF learned guess correction: %s (%lf s)\n", st == 0 ? "applied" : "not applied, SAD kept",
                       get_wtime_sec() - t1);
        }
    }
    double t2 = get_wtime_sec();
    if (!T->guess_ml) TinySCF_D2Cocc(T);
    if (T->verbose && !T->guess_ml)
        printf("TinySCF SAD initial guess: D mat / Cocc mat used %lf, %lf (s)\n", t1 - t0, get_wtime_sec() - t2);
}

/* E = sum D (Hcore + F) (TinySCF.c:367-376), summed on the device.  Energies of a difference-density
 * phase: with G = F - H = 2J - K the energy is 2 tr(D H) + tr(D G(D)).  A difference build forms
 * G(D) = G_r + G~(D - D_r) from the full FP64 build's G_r = G(D_r) and a reduced-precision G~
 * whose error e is of first order in the distance D - D_r, and tr(D G) carries
 * tr(D_r e) + tr((D - D_r) e).  The Coulomb and exchange forms are symmetric, tr(A G(B)) =
 * tr(B G(A)), so tr(D_r G(D - D_r)) = tr((D - D_r) G_r): the energy taken as
 * tr(D (H + F)) + tr(D G_r) - tr(D_r G) is exact up to tr((D - D_r) e), which is of second
 * order (ygpu_fock_energy's e_ref, 0 for other builds). */
static void TinySCF_calc_energy(TinySCF_t T)
{
    double e = 0.0, eref = 0.0;
    YS_GPU(ygpu_fock_energy((ygpu_ctx_t *) T->gpu, &e, &eref));
    T->HF_energy = e + eref;
}

/* the DIIS history as TinySCF_init_free.c:173-180 leaves it */
static void diis_restart(TinySCF_t T)
{
    T->DIIS_len = 0;
    T->DIIS_bmax_id = 0;
    T->DIIS_bmax = -DBL_MAX;
    /* B: -1 off the diagonal (the constraint row and column), 0 on it */
    for (int i = 0; i < (MAX_DIIS + 1) * (MAX_DIIS + 1); i++) T->B_mat[i] = -1.0;
    for (int i = 0; i < MAX_DIIS + 1; i++) T->B_mat[i * (MAX_DIIS + 1) + i] = 0.0;
}

/* the reduced-precision phase of a mixed run after `phase` (FP64 when none is left):
 * BF16x3 products, then BF16x6 products, as mixed_schedule selects */
static int next_reduced_phase(const TinySCF_t T, int phase)
{
    if (phase == YGPU_FOCK_FP64 && (T->mixed_schedule & 1)) return YGPU_FOCK_MIXED;
    if ((phase == YGPU_FOCK_FP64 || phase == YGPU_FOCK_MIXED) && (T->mixed_schedule & 2)) return YGPU_FOCK_MIXED6;
    return YGPU_FOCK_FP64;
}

void TinySCF_do_SCF(TinySCF_t T)
{
    if (T->verbose) {
        printf("TinySCF SCF iteration started...\n");
        printf("Nuclear energy = %.10lf\n", T->nuc_energy);
    }
    T->iter = 0;
    ygpu_ctx_t *g = (ygpu_ctx_t *) T->gpu;
    YS_GPU(ygpu_set_scf_resident(g, 0));        /* iteration 0 starts from the host's guess */
    YS_GPU(ygpu_clear_timings(g));
    T->converged = 0;
    T->diis_error = -1.0;
    /* Precision phases of a mixed run (mixed_schedule): reduced-precision Fock builds (BF16x3
     * products, then BF16x6) while the DIIS error is large; then FP64: with bit 2 set, one full
     * FP64 build becomes the reference of difference-density builds (K = K_ref + K(D - D_ref),
     * YGPU_FOCK_INCR) until the energy converges, and a last full FP64 build confirms it;
     * without it, full FP64 builds to convergence.  Energies of reduced-precision builds never
     * end the SCF. */
    const int use_incr = T->mixed && (T->mixed_schedule & 4);
    T->mixed_phase = T->mixed ? next_reduced_phase(T, YGPU_FOCK_FP64) : YGPU_FOCK_FP64;
    T->mixed_switch_iter = (T->mixed && T->mixed_phase == YGPU_FOCK_FP64) ? 0 : -1;
    T->mixed6_switch_iter = T->mixed_phase == YGPU_FOCK_MIXED6 ? 0 : -1;
    double prev = 0.0, delta = 223.0, best_err = -1.0, diis_at_ref = DBL_MAX;
    int prev_exact = 0, stall = 0, since_ref = 0;
    while (T->iter < T->niters && delta >= T->ene_tol) {
        double st0 = get_wtime_sec();
        if (T->verbose) printf("--------------- Iteration %d ---------------\n", T->iter);
        const int phase = T->mixed_phase;
        double floor_est = -1.0;                /* the reduced-precision floor of this build's DIIS error */
        if (T->mixed) YS_GPU(ygpu_set_fock_precision(g, phase));
        TinySCF_build_FockMat(T);
        int used = phase;
        if (T->mixed) YS_GPU(ygpu_fock_last_info(g, &used, NULL));   /* a difference build may fall back */
        if (T->iter < YS_MAX_PHASE_LOG) T->phase_log[T->iter] = used;
        TinySCF_calc_energy(T);
        delta = fabs(T->HF_energy - prev);
        prev = T->HF_energy;
        double st1 = get_wtime_sec();
        TinySCF_DIIS(T);
        double st2 = get_wtime_sec();
        const double delta_raw = delta;
        /* an energy from full or difference FP64 builds is compared only with another one */
        const int exact = phase == YGPU_FOCK_FP64 || phase == YGPU_FOCK_INCR;
        if (T->iter == 0 || !exact || !prev_exact) delta = 223.0;   /* iteration 0 never converges (TinySCF.c:429) */
        prev_exact = exact;
        if (phase == YGPU_FOCK_MIXED || phase == YGPU_FOCK_MIXED6) {
            /* a phase ends once the DIIS error is below mixed_tol or the phase's reduced-precision
             * floor, or stops falling */
            if (T->iter >= 2 && T->diis_error >= 0.0) {
                /* no new best by 30 % in 3 iterations: the floor lies above the estimate */
                if (best_err < 0.0 || T->diis_error < 0.7 * best_err) {
                    best_err = T->diis_error;
                    stall = 0;
                } else {
                    stall++;
                }
                /* the reduced-precision floor: products with a relative error e give
                 * |dK_ij| <= e sum_k |X_ik||X_jk| <= e max_i K_ii (K_ii = sum_k X_ik^2), and the
                 * DIIS error cannot fall much below that.  BF16x3: e ~ 2^-16 (representation; measured
                 * floors 0.4-1.8 x 1.5 2^-16 max K_ii on C10H22, C54H18, C96H24).  BF16x6: the
                 * products are FP32-accurate and the FP32 accumulation sets e */
                double kdiag = 0.0;
                for (int i = 0; i < T->nbasfuncs; i++)
                    kdiag = fmax(kdiag, fabs(T->K_mat[(size_t) i * T->nbasfuncs + i]));
                floor_est = (phase == YGPU_FOCK_MIXED ? 1.5 * ldexp(1.0, -16) : T->mixed6_floor) * kdiag;
                /* within twice the estimate the build is in the floor region (the measured floors
                 * reach 1.8 x it): the next reduced-precision iteration gains little, and on C96H24
                 * the one after a DIIS error of 1.3 x the estimate left the FP64 part as long */
                if (T->diis_error < fmax(T->mixed_tol, 2.0 * floor_est) || stall >= 3) {
                    const int next = next_reduced_phase(T, phase);
                    const int keep = (next == YGPU_FOCK_FP64) ? (T->mixed_keep_diis & 2) : (T->mixed_keep_diis & 1);
                    /* the stored Fock matrices carry the reduced-precision error: unless kept, the
                     * next phase restarts the DIIS history from its own Fock matrices */
                    if (!keep) diis_restart(T);
                    if (next == YGPU_FOCK_FP64) T->mixed_switch_iter = T->iter + 1;
                    else T->mixed6_switch_iter = T->iter + 1;
                    T->mixed_phase = next;
                    best_err = -1.0;
                    stall = 0;
                }
            }
        } else if (use_incr) {
            /* FP64 and difference builds: a full FP64 build is the reference of the difference
             * builds that follow, whose energies are exact to second order in the distance from
             * it (TinySCF_calc_energy), so two agreeing energies end the SCF.  Options:
             * mixed_reref > 0 takes a new reference once the DIIS error has fallen below that
             * fraction of its value at the last one; mixed_confirm ends the SCF only on a full
             * FP64 build or the difference build right after one, and confirms agreement among
             * later difference builds by a full FP64 build */
            if (used == YGPU_FOCK_FP64) {
                /* a reference build, or a confirming one whose energy disagreed: difference
                 * builds follow (an agreeing confirmation ends the loop) */
                diis_at_ref = T->diis_error >= 0.0 ? T->diis_error : DBL_MAX;
                since_ref = 0;
                if (delta >= T->ene_tol) T->mixed_phase = YGPU_FOCK_INCR;
            } else if (delta < T->ene_tol) {
                if (T->mixed_confirm && ++since_ref > 1) {
                    T->mixed_phase = YGPU_FOCK_FP64;
                    delta = 223.0;
                }
            } else {
                since_ref++;
                if (T->mixed_reref > 0.0 && T->diis_error >= 0.0 && T->diis_error < T->mixed_reref * diis_at_ref)
                    T->mixed_phase = YGPU_FOCK_FP64;
            }
        }
        /* the density of an iteration that continues comes from purification of F' (its occupied
         * space only); the last one, which leaves the orbitals and their energies, from the
         * eigensolver, as the reference's */
        const int last = T->iter + 1 >= T->niters || delta < T->ene_tol;
        YS_GPU(ygpu_set_density_method(g, (last || !T->purify) ? YGPU_DENSITY_EIGEN : YGPU_DENSITY_PURIFY));
        TinySCF_build_DenMat(T);
        /* D, C_occ, F now stay on the device between the calls of the loop (the host copies are
         * still written, as the reference's are) */
        if (T->iter == 0) YS_GPU(ygpu_set_scf_resident(g, 1));
        double st3 = get_wtime_sec();
        T->t_diis = st2 - st1;
        T->t_density = st3 - st2;
        if (T->verbose) {
            static const char *tag[] = {"", " [BF16x3]", " [BF16x6]", " [FP64 + difference]"};
            printf("* DIIS procedure        : %.3lf (s)\n", st2 - st1);
            int dm = YGPU_DENSITY_EIGEN, dsteps = 0;
            YS_GPU(ygpu_density_last_info(g, &dm, &dsteps));
            if (dm == YGPU_DENSITY_PURIFY) printf("* Build density matrix  : %.3lf (s) [SP2, %d steps]\n", st3 - st2, dsteps);
            else printf("* Build density matrix  : %.3lf (s)\n", st3 - st2);
            printf("* Iteration runtime     = %.3lf (s)\n", st3 - st0);
            printf("* Energy = %.10lf (%.10lf)", T->HF_energy + T->nuc_energy, T->HF_energy);
            if (T->iter > 0) printf(", delta = %e", delta_raw);
            if (T->iter >= 2) printf(", DIIS error = %.3e", T->diis_error);
            if (floor_est >= 0.0) printf(" (floor %.3e)", floor_est);
            printf("%s\n", tag[phase]);
        }
        T->iter++;
    }
    YS_GPU(ygpu_set_scf_resident(g, 0));        /* calls outside the loop copy their host inputs */
    if (T->mixed) YS_GPU(ygpu_set_fock_precision(g, YGPU_FOCK_FP64));
    T->energy_delta = delta;
    T->converged = delta < T->ene_tol;
    if (T->verbose) printf("--------------- SCF iterations finished ---------------\n");
}
