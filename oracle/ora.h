/*
 * ora.h -- CPU oracle for the DF-RHF hot path of YATSCF-DF.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (yatscf-df_amd/) may link,
 * load or call this code.  It is used by tests/, by __graft_entry__.smoke()
 * as the checker, and by bench.py's cpu_baseline leg (timed CPU restatement).
 *
 * It restates, in plain C + OpenMP (+ MKL loaded with dlopen when present):
 *   - the reference's integral provider (Simint, not vendored, version
 *     unpinned) by the McMurchie-Davidson scheme: ora_ints.c
 *   - TinySCF.c / TinySCF_init_free.c setup, build_DF_tensor.c, build_Fock.c,
 *     build_density.c, DIIS.c and the SCF loop: ora_scf.c
 *
 * Pinning (see tests/): Boys function and every integral class against an
 * independent mpmath Obara-Saika implementation and closed forms; DF tensor
 * against J^-1 identities; J/K/F against a 4-index DF rebuild; density
 * against idempotency/trace identities; SCF against the converged-energy
 * invariants.  Parity unpinned (no reference vector exists): Simint's
 * FASTSCHWARZ primitive screening (1e-14, CMS_Simint.c:147-148) and its
 * screen_max definition for the (P,unit) aux pairs (CMS_Simint.c:632-637).
 */
#ifndef ORA_H
#define ORA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A contracted Cartesian shell.  coef[] already holds normalised
 * contraction coefficients (x^l-normalised convention, the one Simint applies
 * in simint_normalize_shells, CMS_Simint.c:144-145). */
typedef struct {
    int am, nprim;
    double xyz[3];
    const double *alpha;
    const double *coef;
} ora_shell_t;

/* ---- integrals (ora_ints.c) ---- */
void   ora_boys(int nmax, double T, double *F);
int    ora_ncart(int l);
void   ora_cart_powers(int l, int idx, int *lx, int *ly, int *lz);
void   ora_overlap(const ora_shell_t *A, const ora_shell_t *B, double *out);
void   ora_kinetic(const ora_shell_t *A, const ora_shell_t *B, double *out);
void   ora_nuclear(const ora_shell_t *A, const ora_shell_t *B, int natoms,
                   const double *Z, const double *axyz, double *out);
void   ora_eri3(const ora_shell_t *A, const ora_shell_t *B, const ora_shell_t *C,
                double *out);
void   ora_eri2(const ora_shell_t *C, const ora_shell_t *D, double *out);
double ora_eri4_elem(const ora_shell_t *A, const ora_shell_t *B,
                     const ora_shell_t *C, const ora_shell_t *D,
                     int ia, int ib, int ic, int id);

/* ---- BLAS/LAPACK provider (ora_lapack.c) ---- */
int  ora_lapack_init(void);          /* 1 = MKL loaded, 0 = built-in fallback */
int  ora_lapack_is_mkl(void);
void ora_dgemm(int transa, int transb, int m, int n, int k, double alpha,
               const double *A, int lda, const double *B, int ldb, double beta,
               double *C, int ldc);   /* row-major */
int  ora_dsyev(int n, double *A, double *w);     /* row-major, 'V','U' */
int  ora_dpstrf_lower(int n, double *A, int *piv, int *rank, double tol);
int  ora_dgesv(int n, double *A, int lda, int *ipiv, double *b);
void ora_set_threads(int nt);
int  ora_get_threads(void);

/* ---- SCF engine (ora_scf.c) ---- */
typedef struct ora_scf ora_scf_t;

typedef struct {
    int nshells;
    const int *am, *nprim, *prim_off;   /* per shell */
    const double *xyz;                  /* 3 * nshells, Bohr */
    const double *alpha, *coef;         /* normalised, indexed by prim_off */
} ora_basis_t;

ora_scf_t *ora_create(const ora_basis_t *bas, const ora_basis_t *aux,
                      int natoms, const double *Z, const double *axyz,
                      int charge, int nelectron, double nuc_energy,
                      const double *sad_blockdiag, int niters);
void   ora_destroy(ora_scf_t *s);
void   ora_set_flags(ora_scf_t *s, int quirk_a5_schwarz, int quirk_a1_d2cocc,
                     double ene_tol, double scrtol);
void   ora_set_quirk_a2(ora_scf_t *s, int quirk_a2_diis_bmax);
int    ora_setup(ora_scf_t *s);            /* S,H,X; Schwarz; sparsity; SAD; D2Cocc */
int    ora_build_df_tensor(ora_scf_t *s);
int    ora_build_fock(ora_scf_t *s);
int    ora_build_density(ora_scf_t *s);
int    ora_diis(ora_scf_t *s);
double ora_calc_energy(ora_scf_t *s);
int    ora_do_scf(ora_scf_t *s, int verbose);
void   ora_d2cocc(ora_scf_t *s);

/* accessors: returns pointer to the named array and its element count */
double *ora_get_dptr(ora_scf_t *s, const char *name, size_t *count);
int    *ora_get_iptr(ora_scf_t *s, const char *name, size_t *count);
int     ora_get_int(ora_scf_t *s, const char *name);
double  ora_get_double(ora_scf_t *s, const char *name);
void    ora_set_int(ora_scf_t *s, const char *name, int v);

#ifdef __cplusplus
}
#endif
#endif
