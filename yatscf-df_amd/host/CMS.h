/*
 * CMS.h -- basis-set / molecule input for the TinySCF host engine.
 *
 * Same entry points and conventions as the reference's libCMS
 * (libCMS/CMS_BasisSet.h:82-129): .xyz with charge on line 2 and Angstrom
 * coordinates (A2BOHR = 1/0.52917720859), .gbs with SP splitting and the
 * Cartesian default, SAD guesses <El>.dat next to the .gbs file.  The
 * difference from the reference is that every parser failure is reported
 * through the returned status (the reference's init ignores it,
 * TinySCF_init_free.c:29-30, and continues with partial data).
 */
#ifndef YATSCF_CMS_H
#define YATSCF_CMS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CMS_STATUS_SUCCESS = 0,
    CMS_STATUS_INVALID_VALUE = -1,
    CMS_STATUS_FILEIO_FAILED = -2,
    CMS_STATUS_ALLOC_FAILED = -3
} CMSStatus_t;

typedef struct BasisSet *BasisSet_t;

CMSStatus_t CMS_createBasisSet(BasisSet_t *basis);
CMSStatus_t CMS_destroyBasisSet(BasisSet_t basis);
CMSStatus_t CMS_loadChemicalSystem(BasisSet_t basis, char *bsfile, char *xyzfile);
const char *CMS_lastError(void);

int    CMS_getNumAtoms(BasisSet_t basis);
int    CMS_getNumShells(BasisSet_t basis);
int    CMS_getNumFuncs(BasisSet_t basis);
int    CMS_getNumOccOrb(BasisSet_t basis);
int    CMS_getFuncStartInd(BasisSet_t basis, int shellid);
int    CMS_getFuncEndInd(BasisSet_t basis, int shellid);
int    CMS_getShellDim(BasisSet_t basis, int shellid);
int    CMS_getMaxShellDim(BasisSet_t basis);
int    CMS_getAtomStartInd(BasisSet_t basis, int atomid);
int    CMS_getAtomEndInd(BasisSet_t basis, int atomid);
int    CMS_getTotalCharge(BasisSet_t basis);
int    CMS_getNneutral(BasisSet_t basis);
int    CMS_getMaxMomentum(BasisSet_t basis);
/* 1 if the .gbs file's first token is "spherical" (CMS_BasisSet.c:464-483); the
 * shells are stored with their Cartesian components in both cases */
int    CMS_isSpherical(BasisSet_t basis);
double CMS_getNucEnergy(BasisSet_t basis);
void   CMS_getInitialGuess(BasisSet_t basis, int atomid, double **guess, int *spos, int *epos);
void   CMS_getShellxyz(BasisSet_t basis, int shellid, double *x, double *y, double *z);

/* Shell arrays for the device: am, nprim, prim_off per shell, xyz (3 per
 * shell, Bohr), alpha and normalised coef (x^l-normalised contracted
 * Cartesian shells, the convention of Simint's normalisation,
 * CMS_Simint.c:144-145).  Arrays are owned by the basis set. */
void CMS_getShellArrays(BasisSet_t basis, const int **am, const int **nprim, const int **prim_off,
                        const double **xyz, const double **alpha, const double **coef);
/* Atom charges (double) and coordinates (Bohr, 3 per atom). */
void CMS_getAtoms(BasisSet_t basis, const double **charge, const double **xyz);

#ifdef __cplusplus
}
#endif
#endif
