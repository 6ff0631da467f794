/*
 * DIIS.c -- commutator DIIS, drop-in for the reference's DIIS.c:12-123.
 *
 * FP64 throughout.  The residuals and transformed Fock matrices of the history
 * live on the device (ygpu_diis_update / ygpu_diis_extrapolate), so the host's
 * R_mat / F0_mat stay unused; the (DIIS_len+1)^2 linear system is solved here
 * by Gaussian elimination with partial pivoting (the role of LAPACKE_dgesv at
 * DIIS.c:117).  Iterations 0
 * and 1 only transform F (:30-40); the slot of the largest stored residual is
 * replaced once MAX_DIIS vectors exist, and that maximum is never reset
 * (:93-100), as in the reference, unless quirk_a2 is 0: then the maximum is
 * taken over the stored residuals each iteration.
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "DIIS.h"
#include "utils.h"
#include "yatscf_gpu.h"

static int solve_small(int n, double *A, int lda, double *b)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        double m = fabs(A[k * lda + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * lda + k]) > m) { m = fabs(A[i * lda + k]); p = i; }
        if (m == 0.0) return k + 1;
        if (p != k) {
            for (int c = 0; c < n; c++) { double t = A[k * lda + c]; A[k * lda + c] = A[p * lda + c]; A[p * lda + c] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            const double f = A[i * lda + k] / A[k * lda + k];
            for (int c = k; c < n; c++) A[i * lda + c] -= f * A[k * lda + c];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int c = i + 1; c < n; c++) s -= A[i * lda + c] * b[c];
        b[i] = s / A[i * lda + i];
    }
    return 0;
}

void TinySCF_DIIS(TinySCF_t T)
{
    ygpu_ctx_t *g = (ygpu_ctx_t *) T->gpu;
    const int ldB = MAX_DIIS + 1;
    double st = get_wtime_sec();
    if (T->iter <= 1) {
        /* F = X^T F X only (DIIS.c:30-40), in place: on the device stream the copy of F up (when
         * it is copied) precedes the copy of X^T F X down */
        YS_GPU(ygpu_diis_transform(g, T->F_mat, T->F_mat));
        T->t_diis = get_wtime_sec() - st;
        return;
    }
    int idx;
    if (T->DIIS_len < MAX_DIIS) idx = T->DIIS_len++;
    else idx = T->DIIS_bmax_id;
    /* device: residual X^T (FDS - FDS^T) X into history slot idx, its dot products with
     * the stored residuals, and X^T F X stored as the slot's Fock matrix */
    double dot[MAX_DIIS + 1];
    /* X^T F X is not copied out: the extrapolation below overwrites F_mat (inside the SCF loop
     * the engine's hand-off mode also keeps F and D from being copied up again) */
    YS_GPU(ygpu_diis_update(g, idx, T->DIIS_len, T->F_mat, T->D_mat, dot, NULL));
    YS_GPU(ygpu_diis_error(g, idx, &T->diis_error));
    for (int i = 0; i < T->DIIS_len; i++) {
        T->B_mat[idx * ldB + i] = dot[i];
        T->B_mat[i * ldB + idx] = dot[i];
    }
    if (!T->quirk_a2) T->DIIS_bmax = -DBL_MAX;   /* recomputed over the stored residuals */
    for (int i = 0; i < T->DIIS_len; i++)
        if (T->B_mat[i * ldB + i] > T->DIIS_bmax) {
            T->DIIS_bmax = T->B_mat[i * ldB + i];
            T->DIIS_bmax_id = i;
        }
    memset(T->DIIS_rhs, 0, DBL_SIZE * (MAX_DIIS + 1));
    T->DIIS_rhs[T->DIIS_len] = -1.0;
    double Bc[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    memcpy(Bc, T->B_mat, sizeof(Bc));
    if (solve_small(T->DIIS_len + 1, Bc, ldB, T->DIIS_rhs) != 0) {
        /* singular system (e.g. two identical residuals): keep the newest Fock matrix */
        memset(T->DIIS_rhs, 0, DBL_SIZE * (MAX_DIIS + 1));
        T->DIIS_rhs[idx] = 1.0;
    }
    YS_GPU(ygpu_diis_extrapolate(g, T->DIIS_len, T->DIIS_rhs, T->F_mat));
    T->t_diis = get_wtime_sec() - st;
}
