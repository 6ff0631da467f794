/*
 * main.c -- command-line driver with the reference's call sequence
 * (main.c:16-49): TinySCF.exe <basis> <denfit_basis> <xyz> <niter>.
 */
#include <stdio.h>
#include <stdlib.h>

#include "TinySCF.h"
#include "build_DF_tensor.h"
#include "utils.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        printf("Usage: %s <basis> <denfit_basis> <xyz> <niter>\n", argv[0]);
        return 255;
    }
    TinySCF_t T = (TinySCF_t) malloc(sizeof(struct TinySCF_struct));
    if (!T) return 1;
    double t0 = get_wtime_sec();
    init_TinySCF(T, argv[1], argv[2], argv[3], atoi(argv[4]));
    double t1 = get_wtime_sec();
    TinySCF_init_batch_dgemm_arrays(T);
    TinySCF_compute_Hcore_Ovlp_mat(T);
    TinySCF_compute_sq_Schwarz_scrvals(T);
    TinySCF_prepare_sparsity(T);
    TinySCF_get_initial_guess(T);
    TinySCF_build_DF_tensor(T);
    TinySCF_do_SCF(T);
    double t2 = get_wtime_sec();
    printf("Final energy = %.10lf Eh, %d iterations, %s, time to converged energy %.3lf s (after input %.3lf s)\n",
           T->HF_energy + T->nuc_energy, T->iter, T->converged ? "converged" : "NOT converged", t2 - t1, t1 - t0);
    int ok = T->converged;
    TinySCF_free_batch_dgemm_arrays(T);
    free_TinySCF(T);
    return ok ? 0 : 2;
}
