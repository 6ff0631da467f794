/*
 * build_density.c -- drop-in for the reference's build_density.c:12-69.
 *
 * F_mat already holds X^T F X (written by TinySCF_DIIS).  The device solves
 * the FP64 eigenproblem, forms C = X C0, picks the n_occ lowest orbitals and
 * D = C_occ C_occ^T; the host receives D_mat, Cocc_mat, eigval and ev_idx as
 * the reference leaves them.  Only fields of the reference's TinySCF struct
 * (TinySCF.h:12-100) plus the device handle `gpu` are read; inside an SCF loop
 * that set the engine's hand-off mode (ygpu_set_scf_resident), the F_orth and
 * X already on the device are not copied up again.
 */
#include <stdio.h>
#include <stdlib.h>

#include "build_density.h"
#include "yatscf_gpu.h"

/* Sorts eigval ascending and carries ev_idx along.  Same role and signature
 * as the reference's quickSort_eigval (build_density.c:12-35), whose index
 * swap writes ev_idx[j] = ev_idx[i] instead of the saved value; here the swap
 * is a real swap. */
void quickSort_eigval(double *eigval, int *ev_idx, int l, int r)
{
    int i = l, j = r;
    const double mid = eigval[(l + r) / 2];
    while (i <= j) {
        while (eigval[i] < mid) i++;
        while (eigval[j] > mid) j--;
        if (i <= j) {
            int it = ev_idx[i]; ev_idx[i] = ev_idx[j]; ev_idx[j] = it;
            double dt = eigval[i]; eigval[i] = eigval[j]; eigval[j] = dt;
            i++;
            j--;
        }
    }
    if (i < r) quickSort_eigval(eigval, ev_idx, i, r);
    if (j > l) quickSort_eigval(eigval, ev_idx, l, j);
}

void TinySCF_build_DenMat(TinySCF_t T)
{
    const int status = ygpu_build_density((ygpu_ctx_t *) T->gpu, T->n_occ, T->F_mat, T->X_mat, T->D_mat,
                                          T->Cocc_mat, T->eigval, T->ev_idx);
    if (status != YGPU_OK) ygpu_fail(status, "TinySCF_build_DenMat");
}
