/*
 * build_Fock.c -- drop-in for the reference's build_Fock.c:349-396.
 *
 * One device call builds J (the bandwidth-bound two-pass contraction of the
 * packed DF tensor with D), K (X = C_occ^T B per basis function, K = X X^T on
 * FP64 MFMA) and F = Hcore + 2J - K; the host receives full symmetric J, K
 * and F.  The phase timers keep the reference's labels.
 */
#include <stdio.h>

#include "build_Fock.h"
#include "utils.h"
#include "yatscf_gpu.h"

void TinySCF_build_FockMat(TinySCF_t T)
{
    ygpu_fock_timing_t t;
    double st = get_wtime_sec();
    /* Hcore: the device copy made by TinySCF_compute_Hcore_Ovlp_mat */
    const int res = T->dev_resident;            /* D, C_occ of the last density build on the device */
    YS_GPU(ygpu_build_fock((ygpu_ctx_t *) T->gpu, T->n_occ, res ? NULL : T->D_mat, res ? NULL : T->Cocc_mat, NULL,
                           T->J_mat, T->K_mat, T->F_mat, &t));
    T->t_fock[0] = t.t_tempJ;
    T->t_fock[1] = t.t_Jmat;
    T->t_fock[2] = t.t_tempK;
    T->t_fock[3] = t.t_Kmat;
    T->t_fock[4] = t.t_symm;
    T->t_fock[5] = t.t_total;
    if (T->n_hist < YS_MAX_HIST) {
        for (int k = 0; k < 6; k++) T->t_hist[T->n_hist][k] = T->t_fock[k];
        T->n_hist++;
    }
    if (T->verbose) {
        printf("* Build Fock matrix     : %.3lf (s)\n", get_wtime_sec() - st);
        printf("|---- aux. J tensor     |---- %.3lf (s)\n", t.t_tempJ);
        printf("|---- J matrix          |---- %.3lf (s)\n", t.t_Jmat);
        printf("|---- aux. K tensor     |---- %.3lf (s)\n", t.t_tempK);
        printf("|---- K matrix          |---- %.3lf (s)\n", t.t_Kmat);
        printf("|---- J K symmetrizing  |---- %.3lf (s)\n", t.t_symm);
    }
}
