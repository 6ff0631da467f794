/*
 * build_DF_tensor.c -- drop-in for the reference's build_DF_tensor.c:256-287.
 *
 * The four phases (3-centre integrals, 2-centre metric, J^{-1/2}, the DF
 * GEMM) run on the device; the tensor stays there in packed form.  On return
 * Jpq holds J^{-1/2} as in the reference (:236); pqA and df_tensor are
 * filled only when the host copies were requested (host_df_copy).
 */
#include <stdio.h>

#include "build_DF_tensor.h"
#include "utils.h"
#include "yatscf_gpu.h"

void TinySCF_build_DF_tensor(TinySCF_t T)
{
    ygpu_df_timing_t t;
    if (T->verbose) printf("---------- DF tensor construction ----------\n");
    YS_GPU(ygpu_build_df_tensor_ex((ygpu_ctx_t *) T->gpu, T->pqA, T->Jpq, T->df_tensor, &t));
    T->t_df[0] = t.t_3c;
    T->t_df[1] = t.t_2c;
    T->t_df[2] = t.t_invsqrt;
    T->t_df[3] = t.t_dfgemm;
    if (T->verbose) {
        printf("* 3-center integral : %.3lf (s)\n", t.t_3c);
        printf("* 2-center integral : %.3lf (s)\n", t.t_2c);
        printf("* matrix inv-sqrt   : %.3lf (s)\n", t.t_invsqrt);
        printf("* build DF tensor   : %.3lf (s)\n", t.t_dfgemm);
        printf("---------- DF tensor construction finished ----------\n");
    }
}
