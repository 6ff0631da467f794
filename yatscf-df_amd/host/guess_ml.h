/*
 * guess_ml.h -- learned per-atom-pair correction to the SAD initial density
 * (an extension of TinySCF_get_initial_guess, TinySCF.c:324-364).
 */
#ifndef _YATSCF_GUESS_ML_H_
#define _YATSCF_GUESS_ML_H_

#include "TinySCF.h"

#define YSML_NCEN 32                                   /* radial Gaussian centres */
#define YSML_RCUT 6.0                                  /* cutoff (Angstrom) */
#define YSML_NEL 2                                     /* element channels: H, C */
#define YSML_NTYPES 5                                  /* C-C, C-H, H-H, C self, H self */
#define YSML_NIN (2 * YSML_NEL * YSML_NCEN + YSML_NCEN + YSML_NTYPES)
#define YSML_NOUT 29                                   /* coefficients of a C-C pair */
#define YSML_T_CC 0
#define YSML_T_CH 1
#define YSML_T_HH 2
#define YSML_T_CSELF 3
#define YSML_T_HSELF 4

typedef struct ysml_model ysml_model_t;

/* The molecule as the pair model sees it: atoms (charge, Bohr coordinates),
 * shells per atom [atom_shell_start[a], atom_shell_start[a+1]), each shell's
 * angular momentum and first basis function. */
typedef struct {
    int natoms, nbf;
    const double *Z, *xyz;
    const int *atom_shell_start, *am, *shell_bf_start;
} ysml_mol_t;

/* 1 if every atom is H or C with s/p/d shells the pair model covers */
int ysml_supported(const ysml_mol_t *M);
/* pair list (A, B) within the cutoff, self pairs included; returns the count
 * (pairs may be NULL to count only) */
int ysml_pairs(const ysml_mol_t *M, int *pairs, int max_pairs);
/* YSML_NIN features per pair (unnormalised) */
int ysml_features(const ysml_mol_t *M, int npairs, const int *pairs, float *X);
/* YSML_NOUT coefficients and their mask per pair: the projection of dD */
int ysml_targets(const ysml_mol_t *M, int npairs, const int *pairs, const double *dD, float *Y, float *Mask);
/* dD (nbf x nbf) rebuilt from coefficients */
int ysml_assemble(const ysml_mol_t *M, int npairs, const int *pairs, const float *Y, double *dD);
/* D_mat, Cocc_mat from the n_occ leading eigenvectors of S^1/2 Din S^1/2 */
int ysml_project_density(TinySCF_t T, const double *Din);

ysml_model_t *ysml_load(const char *path);
void ysml_free(ysml_model_t *m);
/* predicted dD with the network evaluated on device context gpu */
int ysml_predict(const ysml_mol_t *M, void *gpu, const ysml_model_t *m, double *dD);
/* the engine's molecule (atom_shell_start left NULL: filled by the caller) */
void ysml_mol_of(TinySCF_t T, ysml_mol_t *M);

/* D_mat (holding the SAD guess) += predicted correction, then projected; nonzero
 * return leaves D_mat untouched (SAD fallback) */
int TinySCF_learned_guess(TinySCF_t T, const char *weights_path);

#endif
