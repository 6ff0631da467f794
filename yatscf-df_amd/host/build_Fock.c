/*
 * build_Fock.c -- drop-in for the reference's build_Fock.c:349-396.
 *
 * One device call builds J (the bandwidth-bound two-pass contraction of the
 * packed DF tensor with D), K (X = C_occ^T B per basis function, K = X X^T on
 * the matrix cores) and F = Hcore + 2J - K; the host receives full symmetric
 * J, K and F, as the reference leaves them.  The file reads only fields of the
 * reference's TinySCF struct (TinySCF.h:12-100) plus the device handle `gpu`;
 * inside an SCF loop that set the engine's hand-off mode
 * (ygpu_set_scf_resident), D, C_occ and Hcore already on the device are not
 * copied up again.  The phase timers keep the reference's labels; the device
 * engine also keeps them (ygpu_get_fock_timings).  YATSCF_VERBOSE=0 silences
 * the printout.
 */
#include <stdio.h>
#include <stdlib.h>

#include "build_Fock.h"
#include "utils.h"
#include "yatscf_gpu.h"

static int fock_verbose(void)
{
    const char *v = getenv("YATSCF_VERBOSE");
    return !(v && v[0] == '0');
}

void TinySCF_build_FockMat(TinySCF_t T)
{
    ygpu_fock_timing_t t;
    const double st = get_wtime_sec();
    const int status = ygpu_build_fock((ygpu_ctx_t *) T->gpu, T->n_occ, T->D_mat, T->Cocc_mat, T->Hcore_mat,
                                       T->J_mat, T->K_mat, T->F_mat, &t);
    if (status != YGPU_OK) ygpu_fail(status, "TinySCF_build_FockMat");
    if (fock_verbose()) {
        printf("* Build Fock matrix     : %.3lf (s)\n", get_wtime_sec() - st);
        printf("|---- aux. J tensor     |---- %.3lf (s)\n", t.t_tempJ);
        printf("|---- J matrix          |---- %.3lf (s)\n", t.t_Jmat);
        printf("|---- aux. K tensor     |---- %.3lf (s)\n", t.t_tempK);
        printf("|---- K matrix          |---- %.3lf (s)\n", t.t_Kmat);
        printf("|---- J K symmetrizing  |---- %.3lf (s)\n", t.t_symm);
    }
}
