/*
 * tinyscf_pyapi.c -- accessors that let Python (tests, bench.py) drive the
 * TinySCF host engine through ctypes without mirroring the struct layout.
 * Not part of the reference API.
 */
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "DIIS.h"
#include "TinySCF.h"
#include "build_DF_tensor.h"
#include "build_Fock.h"
#include "build_density.h"
#include "guess_ml.h"
#include "utils.h"
#include "yatscf_gpu.h"

/* Runs one TinySCF API function by name; a failing device call returns its
 * status here (message from tinyscf_error) instead of aborting the caller. */
int tinyscf_call(TinySCF_t T, const char *name)
{
    static const struct { const char *n; void (*f)(TinySCF_t); } fns[] = {
        {"TinySCF_init_batch_dgemm_arrays", TinySCF_init_batch_dgemm_arrays},
        {"TinySCF_compute_Hcore_Ovlp_mat", TinySCF_compute_Hcore_Ovlp_mat},
        {"TinySCF_compute_sq_Schwarz_scrvals", TinySCF_compute_sq_Schwarz_scrvals},
        {"TinySCF_prepare_sparsity", TinySCF_prepare_sparsity},
        {"TinySCF_get_initial_guess", TinySCF_get_initial_guess},
        {"TinySCF_build_DF_tensor", TinySCF_build_DF_tensor},
        {"TinySCF_do_SCF", TinySCF_do_SCF},
        {"TinySCF_build_FockMat", TinySCF_build_FockMat},
        {"TinySCF_build_DenMat", TinySCF_build_DenMat},
        {"TinySCF_DIIS", TinySCF_DIIS},
        {"TinySCF_D2Cocc", TinySCF_D2Cocc},
    };
    for (size_t k = 0; k < sizeof(fns) / sizeof(fns[0]); k++) {
        if (strcmp(name, fns[k].n) != 0) continue;
        jmp_buf jb;
        int st = setjmp(jb);
        if (st == 0) {
            ys_set_error_jump(&jb);
            fns[k].f(T);
        }
        ys_set_error_jump(NULL);
        return st;
    }
    return -1;
}

const char *tinyscf_error(void) { return ys_error_message(); }

TinySCF_t tinyscf_alloc(void) { return (TinySCF_t) calloc(1, sizeof(struct TinySCF_struct)); }

double *tinyscf_dptr(TinySCF_t T, const char *n, size_t *count)
{
    const size_t ms = (size_t) T->mat_size;
#define RET(name, ptr, cnt) if (!strcmp(n, name)) { *count = (ptr) ? (cnt) : 0; return (ptr); }
    RET("Hcore_mat", T->Hcore_mat, ms) RET("S_mat", T->S_mat, ms) RET("F_mat", T->F_mat, ms)
    RET("D_mat", T->D_mat, ms) RET("J_mat", T->J_mat, ms) RET("K_mat", T->K_mat, ms)
    RET("X_mat", T->X_mat, ms) RET("tmp_mat", T->tmp_mat, ms)
    RET("Cocc_mat", T->Cocc_mat, (size_t) T->nbasfuncs * T->n_occ) RET("eigval", T->eigval, (size_t) T->nbasfuncs)
    RET("sp_scrval", T->sp_scrval, (size_t) T->nshellpairs) RET("bf_pair_scrval", T->bf_pair_scrval, ms)
    RET("df_sp_scrval", T->df_sp_scrval, (size_t) T->df_nshells)
    RET("Jpq", T->Jpq, (size_t) T->df_nbf * T->df_nbf)
    RET("pqA", T->pqA, (size_t) T->num_bf_pair_scr * T->df_nbf)
    RET("df_tensor", T->df_tensor, (size_t) T->num_bf_pair_scr * T->df_nbf)
    RET("B_mat", T->B_mat, (size_t) (MAX_DIIS + 1) * (MAX_DIIS + 1))
#undef RET
    /* phase timers kept by the device engine: the last DF build, every Fock build of the last
     * SCF (six values each, the fields of ygpu_fock_timing_t) */
    static double t_df[4];
    static double t_hist[4096 * 6];
    if (!strcmp(n, "t_df")) {
        ygpu_df_timing_t t;
        if (!T->gpu || ygpu_get_df_timing((ygpu_ctx_t *) T->gpu, &t) != YGPU_OK) { *count = 0; return NULL; }
        t_df[0] = t.t_3c; t_df[1] = t.t_2c; t_df[2] = t.t_invsqrt; t_df[3] = t.t_dfgemm;
        *count = 4;
        return t_df;
    }
    if (!strcmp(n, "t_hist")) {
        int cnt = 0;
        if (!T->gpu || ygpu_get_fock_timings((ygpu_ctx_t *) T->gpu, 4096, (ygpu_fock_timing_t *) t_hist, &cnt) != YGPU_OK) {
            *count = 0;
            return NULL;
        }
        *count = (size_t) 6 * (cnt < 4096 ? cnt : 4096);
        return t_hist;
    }
    *count = 0;
    return NULL;
}

int *tinyscf_iptr(TinySCF_t T, const char *n, size_t *count)
{
#define RET(name, ptr, cnt) if (!strcmp(n, name)) { *count = (ptr) ? (cnt) : 0; return (ptr); }
    RET("bf_pair_mask", T->bf_pair_mask, (size_t) T->mat_size)
    RET("phase_log", T->phase_log, (size_t) (T->iter < YS_MAX_PHASE_LOG ? T->iter : YS_MAX_PHASE_LOG))
    RET("bf_pair_j", T->bf_pair_j, (size_t) T->num_bf_pair_scr)
    RET("bf_pair_diag", T->bf_pair_diag, (size_t) T->nbasfuncs)
    RET("bf_mask_displs", T->bf_mask_displs, (size_t) T->nbasfuncs + 1)
    RET("uniq_sp_lid", T->uniq_sp_lid, (size_t) T->num_uniq_sp)
    RET("uniq_sp_rid", T->uniq_sp_rid, (size_t) T->num_uniq_sp)
    RET("shell_bf_sind", T->shell_bf_sind, (size_t) T->nshells + 1)
    RET("ev_idx", T->ev_idx, (size_t) T->nbasfuncs)
#undef RET
    *count = 0;
    return NULL;
}

int tinyscf_geti(TinySCF_t T, const char *n)
{
    if (!strcmp(n, "nbasfuncs")) return T->nbasfuncs;
    if (!strcmp(n, "nbf_sph")) return T->nbf_sph;
    if (!strcmp(n, "sph_orb")) return T->sph_orb;
    if (!strcmp(n, "sph_aux")) return T->sph_aux;
    if (!strcmp(n, "df_nbf")) return T->df_nbf;
    if (!strcmp(n, "nshells")) return T->nshells;
    if (!strcmp(n, "df_nshells")) return T->df_nshells;
    if (!strcmp(n, "n_occ")) return T->n_occ;
    if (!strcmp(n, "natoms")) return T->natoms;
    if (!strcmp(n, "iter")) return T->iter;
    if (!strcmp(n, "niters")) return T->niters;
    if (!strcmp(n, "num_uniq_sp")) return T->num_uniq_sp;
    if (!strcmp(n, "num_bf_pair_scr")) return T->num_bf_pair_scr;
    if (!strcmp(n, "converged")) return T->converged;
    if (!strcmp(n, "DIIS_len")) return T->DIIS_len;
    if (!strcmp(n, "charge")) return T->charge;
    if (!strcmp(n, "quirk_a1")) return T->quirk_a1;
    if (!strcmp(n, "quirk_a2")) return T->quirk_a2;
    if (!strcmp(n, "quirk_a5")) return T->quirk_a5;
    if (!strcmp(n, "mixed")) return T->mixed;
    if (!strcmp(n, "mixed_switch_iter")) return T->mixed_switch_iter;
    if (!strcmp(n, "mixed6_switch_iter")) return T->mixed6_switch_iter;
    if (!strcmp(n, "mixed_schedule")) return T->mixed_schedule;
    if (!strcmp(n, "mixed_keep_diis")) return T->mixed_keep_diis;
    if (!strcmp(n, "mixed_confirm")) return T->mixed_confirm;
    if (!strcmp(n, "purify")) return T->purify;
    if (!strcmp(n, "guess_ml")) return T->guess_ml;
    if (!strcmp(n, "electron")) return T->electron;
    return -1;
}

void tinyscf_seti(TinySCF_t T, const char *n, int v)
{
    if (!strcmp(n, "iter")) T->iter = v;
    else if (!strcmp(n, "niters")) T->niters = v;
    else if (!strcmp(n, "verbose")) T->verbose = v;
    else if (!strcmp(n, "quirk_a1")) T->quirk_a1 = v;
    else if (!strcmp(n, "quirk_a2")) T->quirk_a2 = v;
    else if (!strcmp(n, "quirk_a5")) T->quirk_a5 = v;
    else if (!strcmp(n, "mixed")) T->mixed = v;
    else if (!strcmp(n, "mixed_schedule")) T->mixed_schedule = v;
    else if (!strcmp(n, "mixed_keep_diis")) T->mixed_keep_diis = v;
    else if (!strcmp(n, "mixed_confirm")) T->mixed_confirm = v;
    else if (!strcmp(n, "purify")) T->purify = v;
}

double tinyscf_getd(TinySCF_t T, const char *n)
{
    if (!strcmp(n, "HF_energy")) return T->HF_energy;
    if (!strcmp(n, "nuc_energy")) return T->nuc_energy;
    if (!strcmp(n, "total_energy")) return T->HF_energy + T->nuc_energy;
    if (!strcmp(n, "energy_delta")) return T->energy_delta;
    if (!strcmp(n, "max_df_scrval")) return T->max_df_scrval;
    if (!strcmp(n, "t_density")) return T->t_density;
    if (!strcmp(n, "t_diis")) return T->t_diis;
    if (!strcmp(n, "ene_tol")) return T->ene_tol;
    if (!strcmp(n, "diis_error")) return T->diis_error;
    if (!strcmp(n, "mixed_tol")) return T->mixed_tol;
    return 0.0;
}

void tinyscf_setd(TinySCF_t T, const char *n, double v)
{
    if (!strcmp(n, "ene_tol")) T->ene_tol = v;
    else if (!strcmp(n, "mixed_tol")) T->mixed_tol = v;
    else if (!strcmp(n, "mixed6_floor")) T->mixed6_floor = v;
    else if (!strcmp(n, "mixed_reref")) T->mixed_reref = v;
}

void *tinyscf_gpu(TinySCF_t T) { return T->gpu; }

/* the engine's molecule for the pair model; ss receives natoms + 1 shell starts */
void tinyscf_fill_mol(TinySCF_t T, ysml_mol_t *M, int *ss)
{
    ysml_mol_of(T, M);
    for (int a = 0; a < T->natoms; a++) ss[a] = CMS_getAtomStartInd(T->basis, a);
    ss[T->natoms] = T->nshells;
    M->atom_shell_start = ss;
}

void tinyscf_set_guess_weights(TinySCF_t T, const char *path)
{
    snprintf(T->guess_weights, sizeof(T->guess_weights), "%s", path ? path : "");
}

/* TinySCF_D2Cocc on a caller-given density (nbf x nbf, row-major) without an engine: the
 * host-only step between the SAD guess and the first Fock build, exposed so that it can be
 * compared with the oracle's dpstrf-based factor on the CPU. */
int tinyscf_d2cocc(int nbf, int n_occ, int quirk_a1, const double *D, double *Cocc)
{
    struct TinySCF_struct T;
    memset(&T, 0, sizeof(T));
    T.nbasfuncs = nbf;
    T.n_occ = n_occ;
    T.mat_size = nbf * nbf;
    T.quirk_a1 = quirk_a1;
    T.D_mat = (double *) malloc(sizeof(double) * (size_t) nbf * nbf);
    T.tmp_mat = (double *) malloc(sizeof(double) * (size_t) nbf * nbf);
    T.Cocc_mat = Cocc;
    if (!T.D_mat || !T.tmp_mat) {
        free(T.D_mat);
        free(T.tmp_mat);
        return -1;
    }
    memcpy(T.D_mat, D, sizeof(double) * (size_t) nbf * nbf);
    TinySCF_D2Cocc(&T);
    free(T.D_mat);
    free(T.tmp_mat);
    return 0;
}
