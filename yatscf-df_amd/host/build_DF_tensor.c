/*
 * build_DF_tensor.c -- drop-in for the reference's build_DF_tensor.c:256-287.
 *
 * The four phases (3-centre integrals, 2-centre metric, the metric factor,
 * the DF GEMM) run on the device; the tensor stays there in packed form.  On
 * return Jpq holds the metric factor W of B = (mn|P) W^T (J^{-1/2} as in the
 * reference, :236, when the engine is set to YGPU_METRIC_INVSQRT; L^{-1} in
 * the default Cholesky mode); pqA and df_tensor receive host copies in the
 * reference's layout when they are allocated (the reference allocates both in
 * TinySCF_prepare_sparsity, TinySCF.c:268-273; NULL skips the copy).  Only
 * fields of the reference's TinySCF struct (TinySCF.h:12-100) plus the device
 * handle `gpu` are read.  YATSCF_VERBOSE=0 silences the printout.
 */
#include <stdio.h>
#include <stdlib.h>

#include "build_DF_tensor.h"
#include "yatscf_gpu.h"

void TinySCF_build_DF_tensor(TinySCF_t T)
{
    const char *v = getenv("YATSCF_VERBOSE");
    const int verbose = !(v && v[0] == '0');
    ygpu_df_timing_t t;
    if (verbose) printf("---------- DF tensor construction ----------\n");
    const int status = ygpu_build_df_tensor_ex((ygpu_ctx_t *) T->gpu, T->pqA, T->Jpq, T->df_tensor, &t);
    if (status != YGPU_OK) ygpu_fail(status, "TinySCF_build_DF_tensor");
    if (verbose) {
        printf("* 3-center integral : %.3lf (s)\n", t.t_3c);
        printf("* 2-center integral : %.3lf (s)\n", t.t_2c);
        printf("* matrix inv-sqrt   : %.3lf (s)\n", t.t_invsqrt);
        printf("* build DF tensor   : %.3lf (s)\n", t.t_dfgemm);
        printf("---------- DF tensor construction finished ----------\n");
    }
}
