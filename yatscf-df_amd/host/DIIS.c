/*
 * DIIS.c -- commutator DIIS, drop-in for the reference's DIIS.c:12-123.
 *
 * FP64 throughout.  The nbf^3 products run on the device (ygpu_dgemm); the
 * (DIIS_len+1)^2 linear system is solved here by Gaussian elimination with
 * partial pivoting (the role of LAPACKE_dgesv at DIIS.c:117).  Iterations 0
 * and 1 only transform F (:30-40); the slot of the largest stored residual is
 * replaced once MAX_DIIS vectors exist, and that maximum is never reset
 * (:93-100), as in the reference.
 */
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
    const int nbf = T->nbasfuncs, ldB = MAX_DIIS + 1;
    const size_t ms = (size_t) T->mat_size;
    double *tmp = T->tmp_mat;
    double st = get_wtime_sec();
    if (T->iter <= 1) {
        YS_GPU(ygpu_dgemm(g, 1, 0, nbf, nbf, nbf, 1.0, T->X_mat, nbf, T->F_mat, nbf, 0.0, tmp, nbf));
        YS_GPU(ygpu_dgemm(g, 0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, T->X_mat, nbf, 0.0, T->F_mat, nbf));
        T->t_diis = get_wtime_sec() - st;
        return;
    }
    int idx;
    if (T->DIIS_len < MAX_DIIS) idx = T->DIIS_len++;
    else idx = T->DIIS_bmax_id;
    /* FDS = F D S; residual X^T (FDS - FDS^T) X */
    YS_GPU(ygpu_dgemm(g, 0, 0, nbf, nbf, nbf, 1.0, T->F_mat, nbf, T->D_mat, nbf, 0.0, tmp, nbf));
    YS_GPU(ygpu_dgemm(g, 0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, T->S_mat, nbf, 0.0, T->FDS_mat, nbf));
    for (int i = 0; i < nbf; i++)
        for (int j = 0; j < nbf; j++)
            tmp[(size_t) i * nbf + j] = T->FDS_mat[(size_t) i * nbf + j] - T->FDS_mat[(size_t) j * nbf + i];
    YS_GPU(ygpu_dgemm(g, 1, 0, nbf, nbf, nbf, 1.0, T->X_mat, nbf, tmp, nbf, 0.0, T->FDS_mat, nbf));
    YS_GPU(ygpu_dgemm(g, 0, 0, nbf, nbf, nbf, 1.0, T->FDS_mat, nbf, T->X_mat, nbf, 0.0, tmp, nbf));
    double dot[MAX_DIIS + 1];
    memcpy(T->R_mat + ms * idx, tmp, DBL_SIZE * ms);
    const double *Ri = T->R_mat + ms * idx;
    for (int j = 0; j < T->DIIS_len; j++) {
        const double *Rj = T->R_mat + ms * j;
        double d = 0.0;
        for (size_t q = 0; q < ms; q++) d += Ri[q] * Rj[q];
        dot[j] = d;
    }
    for (int i = 0; i < T->DIIS_len; i++) {
        T->B_mat[idx * ldB + i] = dot[i];
        T->B_mat[i * ldB + idx] = dot[i];
    }
    for (int i = 0; i < T->DIIS_len; i++)
        if (T->B_mat[i * ldB + i] > T->DIIS_bmax) {
            T->DIIS_bmax = T->B_mat[i * ldB + i];
            T->DIIS_bmax_id = i;
        }
    YS_GPU(ygpu_dgemm(g, 1, 0, nbf, nbf, nbf, 1.0, T->X_mat, nbf, T->F_mat, nbf, 0.0, tmp, nbf));
    YS_GPU(ygpu_dgemm(g, 0, 0, nbf, nbf, nbf, 1.0, tmp, nbf, T->X_mat, nbf, 0.0, T->F_mat, nbf));
    memcpy(T->F0_mat + ms * idx, T->F_mat, DBL_SIZE * ms);
    memset(T->DIIS_rhs, 0, DBL_SIZE * (MAX_DIIS + 1));
    T->DIIS_rhs[T->DIIS_len] = -1.0;
    double Bc[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    memcpy(Bc, T->B_mat, sizeof(Bc));
    solve_small(T->DIIS_len + 1, Bc, ldB, T->DIIS_rhs);
    memset(T->F_mat, 0, DBL_SIZE * ms);
    for (int i = 0; i < T->DIIS_len; i++) {
        const double c = T->DIIS_rhs[i];
        const double *f0 = T->F0_mat + ms * i;
        for (size_t q = 0; q < ms; q++) T->F_mat[q] += c * f0[q];
    }
    T->t_diis = get_wtime_sec() - st;
}
