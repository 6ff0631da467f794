/*
 * cms_basis.c -- .xyz / .gbs / SAD input for the TinySCF host engine.
 *
 * Conventions follow the reference's libCMS so that the same files describe
 * the same chemical system:
 *   xyz: atom count line, charge = atof(line 2) stored as int, "El x y z" in
 *        Angstrom scaled by 1/0.52917720859 (CMS_BasisSet.c:39, 310-426);
 *        a blank line after the atoms is an error (the reference's sscanf then
 *        re-reads the previous symbol and fails, :365-381); nuclear repulsion
 *        by the closed form (:409-421).
 *   gbs: first token "spherical" selects spherical, anything else Cartesian
 *        (:468-483); the shells are stored with their Cartesian components
 *        either way (a spherical basis is the span of the solid harmonics
 *        among them, see CMS_isSpherical); element blocks end at "****"; "<letters> <nexp> <scale>"
 *        shell lines, one shell per letter sharing the exponents (SP), scale
 *        ignored; primitive lines read with one persistent coefficient array
 *        (:556-616).
 *   SAD: <El>.dat next to the .gbs, nfunc^2 values, one-sided symmetry test
 *        g[j][k]-g[k][j] > 1e-12, zero block when missing or bad (:657-746).
 * Shells are ordered atom by atom, each element's shells in file order
 * (:192-299).
 */
#include <ctype.h>
#include <libgen.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "CMS.h"

#define NELEM 50
static const char *ELEMENTS[NELEM] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
    "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
    "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn"};
static const char AM_LETTERS[] = "SPDFG";
static const double ANG_TO_BOHR = 1.0 / 0.52917720859;

typedef struct {
    int am, nprim;
    double *exps, *cc;
} gbs_shell_t;

typedef struct {
    int present, nshell;
    gbs_shell_t *shells;
    double *sad;      /* nf x nf */
    int nf;
} gbs_elem_t;

struct BasisSet {
    int natoms, *eid, Q, nelectrons;
    double *charge, *axyz, ene_nuc;
    gbs_elem_t elem[NELEM];
    int cartesian;
    int nshells, nfunctions, maxdim, max_momentum;
    int *am, *nprim, *prim_off, *f_start, *s_start;
    double *xyz, *alpha, *coef;
};

static char cms_err[512];

const char *CMS_lastError(void) { return cms_err; }

static CMSStatus_t fail(CMSStatus_t st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(cms_err, sizeof(cms_err), fmt, ap);
    va_end(ap);
    return st;
}

static int elem_index(const char *s)
{
    for (int i = 0; i < NELEM; i++)
        if (strcmp(s, ELEMENTS[i]) == 0) return i;
    return -1;
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

CMSStatus_t CMS_createBasisSet(BasisSet_t *basis)
{
    *basis = (BasisSet_t) calloc(1, sizeof(struct BasisSet));
    return *basis ? CMS_STATUS_SUCCESS : fail(CMS_STATUS_ALLOC_FAILED, "allocation failed");
}

CMSStatus_t CMS_destroyBasisSet(BasisSet_t b)
{
    if (!b) return CMS_STATUS_SUCCESS;
    for (int e = 0; e < NELEM; e++) {
        for (int s = 0; s < b->elem[e].nshell; s++) {
            free(b->elem[e].shells[s].exps);
            free(b->elem[e].shells[s].cc);
        }
        free(b->elem[e].shells);
        free(b->elem[e].sad);
    }
    free(b->eid); free(b->charge); free(b->axyz);
    free(b->am); free(b->nprim); free(b->prim_off); free(b->f_start); free(b->s_start);
    free(b->xyz); free(b->alpha); free(b->coef);
    free(b);
    return CMS_STATUS_SUCCESS;
}

static CMSStatus_t read_xyz(BasisSet_t b, const char *path)
{
    FILE *fp = fopen(path, "r");
    char line[1024], sym[1024];
    if (!fp) return fail(CMS_STATUS_FILEIO_FAILED, "cannot open molecule file %s", path);
    if (!fgets(line, sizeof(line), fp) || sscanf(line, "%d", &b->natoms) != 1 || b->natoms <= 0) {
        fclose(fp);
        return fail(CMS_STATUS_FILEIO_FAILED, "%s: bad atom count line", path);
    }
    if (!fgets(line, sizeof(line), fp)) {
        fclose(fp);
        return fail(CMS_STATUS_FILEIO_FAILED, "%s: missing charge line", path);
    }
    b->Q = (int) atof(line);
    b->eid = (int *) calloc(b->natoms, sizeof(int));
    b->charge = (double *) calloc(b->natoms, sizeof(double));
    b->axyz = (double *) calloc(3 * (size_t) b->natoms, sizeof(double));
    int n = 0;
    b->nelectrons = 0;
    while (fgets(line, sizeof(line), fp)) {
        double x, y, z;
        int got = sscanf(line, "%1023s %lf %lf %lf", sym, &x, &y, &z);
        if (got <= 0) {
            if (n > 0) { fclose(fp); return fail(CMS_STATUS_INVALID_VALUE, "%s: blank line after atom %d", path, n); }
            continue;
        }
        if (!isalpha((unsigned char) sym[0])) continue;
        int e = elem_index(sym);
        if (strlen(sym) > 2 || e < 0) { fclose(fp); return fail(CMS_STATUS_INVALID_VALUE, "%s: atom %s is not supported", path, sym); }
        if (got != 4) { fclose(fp); return fail(CMS_STATUS_INVALID_VALUE, "%s: short atom line for %s", path, sym); }
        if (n >= b->natoms) { n++; continue; }
        b->eid[n] = e + 1;
        b->charge[n] = (double) (e + 1);
        b->axyz[3 * n] = x * ANG_TO_BOHR;
        b->axyz[3 * n + 1] = y * ANG_TO_BOHR;
        b->axyz[3 * n + 2] = z * ANG_TO_BOHR;
        b->nelectrons += e + 1;
        n++;
    }
    fclose(fp);
    if (n != b->natoms)
        return fail(CMS_STATUS_FILEIO_FAILED, "%s: %d atoms found but the header says %d", path, n, b->natoms);
    double ene = 0.0;
    for (int a = 0; a < n; a++)
        for (int c = a + 1; c < n; c++) {
            double dx = b->axyz[3 * a] - b->axyz[3 * c], dy = b->axyz[3 * a + 1] - b->axyz[3 * c + 1];
            double dz = b->axyz[3 * a + 2] - b->axyz[3 * c + 2];
            ene += b->charge[a] * b->charge[c] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    b->ene_nuc = ene;
    return CMS_STATUS_SUCCESS;
}

static CMSStatus_t read_gbs(BasisSet_t b, const char *path)
{
    FILE *fp = fopen(path, "r");
    char line[1024], tok[1024];
    if (!fp) return fail(CMS_STATUS_FILEIO_FAILED, "cannot open basis file %s", path);
    if (!fgets(line, sizeof(line), fp)) { fclose(fp); return fail(CMS_STATUS_FILEIO_FAILED, "%s: empty", path); }
    tok[0] = 0;
    sscanf(line, "%1023s", tok);
    b->cartesian = strcmp(tok, "spherical") != 0;
    double cc[3] = {0.0, 0.0, 0.0};
    while (fgets(line, sizeof(line), fp)) {
        if (!isalpha((unsigned char) line[0])) continue;
        sscanf(line, "%1023s", tok);
        int e = elem_index(tok);
        if (e < 0) { fclose(fp); return fail(CMS_STATUS_INVALID_VALUE, "%s: element %s is not supported", path, tok); }
        gbs_elem_t *el = &b->elem[e];
        el->present = 1;
        while (fgets(line, sizeof(line), fp)) {
            if (line[0] == '*') break;
            if (!isalpha((unsigned char) line[0])) continue;
            int nexp = 0;
            double scale = 0.0;
            if (sscanf(line, "%1023s %d %lf", tok, &nexp, &scale) < 2 || nexp <= 0 || strlen(tok) > 3) {
                fclose(fp);
                return fail(CMS_STATUS_INVALID_VALUE, "%s: invalid shell line '%s'", path, tok);
            }
            int ns = (int) strlen(tok);
            double *ex = (double *) calloc(nexp, sizeof(double)), *cv = (double *) calloc(3 * (size_t) nexp, sizeof(double));
            for (int k = 0; k < nexp; k++) {
                if (!fgets(line, sizeof(line), fp) || line[0] == '*' || isalpha((unsigned char) line[0])) {
                    free(ex); free(cv); fclose(fp);
                    return fail(CMS_STATUS_FILEIO_FAILED, "%s: truncated shell", path);
                }
                sscanf(line, "%lf %lf %lf %lf", &ex[k], &cc[0], &cc[1], &cc[2]);
                for (int q = 0; q < 3; q++) cv[3 * k + q] = cc[q];
            }
            for (int q = 0; q < ns; q++) {
                const char *p = strchr(AM_LETTERS, tok[q]);
                if (!p || !*p) { free(ex); free(cv); fclose(fp); return fail(CMS_STATUS_INVALID_VALUE, "%s: shell %s is not supported", path, tok); }
                el->shells = (gbs_shell_t *) realloc(el->shells, sizeof(gbs_shell_t) * (el->nshell + 1));
                gbs_shell_t *sh = &el->shells[el->nshell++];
                sh->am = (int) (p - AM_LETTERS);
                sh->nprim = nexp;
                sh->exps = (double *) malloc(sizeof(double) * nexp);
                sh->cc = (double *) malloc(sizeof(double) * nexp);
                memcpy(sh->exps, ex, sizeof(double) * nexp);
                for (int k = 0; k < nexp; k++) sh->cc[k] = cv[3 * k + q];
            }
            free(ex);
            free(cv);
        }
    }
    fclose(fp);
    return CMS_STATUS_SUCCESS;
}

static void read_sad(BasisSet_t b, const char *gbs_path)
{
    char *tmp = strdup(gbs_path);
    char *dir = dirname(tmp);
    char fname[2048], line[1024];
    for (int e = 0; e < NELEM; e++) {
        gbs_elem_t *el = &b->elem[e];
        if (!el->present) continue;
        int nf = 0;
        for (int s = 0; s < el->nshell; s++) nf += ncart(el->shells[s].am);
        el->nf = nf;
        el->sad = (double *) calloc((size_t) nf * nf, sizeof(double));
        snprintf(fname, sizeof(fname), "%s/%s.dat", dir, ELEMENTS[e]);
        FILE *fp = fopen(fname, "r");
        int bad = (fp == NULL);
        for (int k = 0; !bad && k < nf * nf; k++) {
            if (!fgets(line, sizeof(line), fp)) bad = 1;
            else el->sad[k] = atof(line);
        }
        if (fp) fclose(fp);
        for (int j = 0; !bad && j < nf; j++)
            for (int k = 0; k < nf; k++)
                if (el->sad[j * nf + k] - el->sad[k * nf + j] > 1e-12) bad = 1;
        if (bad) memset(el->sad, 0, sizeof(double) * nf * nf);
    }
    free(tmp);
}

static double dfact(int n)
{
    double r = 1.0;
    for (; n > 1; n -= 2) r *= n;
    return r;
}

static CMSStatus_t build_shells(BasisSet_t b)
{
    int ns = 0, np = 0;
    for (int a = 0; a < b->natoms; a++) {
        gbs_elem_t *el = &b->elem[b->eid[a] - 1];
        if (!el->present) return fail(CMS_STATUS_INVALID_VALUE, "element %s missing from the basis set", ELEMENTS[b->eid[a] - 1]);
        ns += el->nshell;
        for (int s = 0; s < el->nshell; s++) np += el->shells[s].nprim;
    }
    b->nshells = ns;
    b->am = (int *) calloc(ns, sizeof(int));
    b->nprim = (int *) calloc(ns, sizeof(int));
    b->prim_off = (int *) calloc(ns, sizeof(int));
    b->f_start = (int *) calloc(ns + 1, sizeof(int));
    b->s_start = (int *) calloc(b->natoms + 1, sizeof(int));
    b->xyz = (double *) calloc(3 * (size_t) ns, sizeof(double));
    b->alpha = (double *) calloc(np, sizeof(double));
    b->coef = (double *) calloc(np, sizeof(double));
    int si = 0, pi = 0, nf = 0;
    b->maxdim = 0;
    b->max_momentum = 0;
    for (int a = 0; a < b->natoms; a++) {
        gbs_elem_t *el = &b->elem[b->eid[a] - 1];
        b->s_start[a] = si;
        for (int s = 0; s < el->nshell; s++, si++) {
            const gbs_shell_t *sh = &el->shells[s];
            const int l = sh->am;
            b->am[si] = l;
            b->nprim[si] = sh->nprim;
            b->prim_off[si] = pi;
            for (int x = 0; x < 3; x++) b->xyz[3 * si + x] = b->axyz[3 * a + x];
            b->f_start[si] = nf;
            nf += ncart(l);
            if (ncart(l) > b->maxdim) b->maxdim = ncart(l);
            if (l > b->max_momentum) b->max_momentum = l;
            /* x^l-normalised contraction: primitive norms (2a/pi)^{3/4} (4a)^{l/2} / sqrt((2l-1)!!),
             * then the contracted axis component scaled to unit self-overlap */
            const double df = dfact(2 * l - 1);
            double norm2 = 0.0;
            for (int i = 0; i < sh->nprim; i++) {
                const double ai = sh->exps[i];
                b->alpha[pi + i] = ai;
                b->coef[pi + i] = sh->cc[i] * pow(2.0 * ai / M_PI, 0.75) * pow(4.0 * ai, 0.5 * l) / sqrt(df);
            }
            for (int i = 0; i < sh->nprim; i++)
                for (int j = 0; j < sh->nprim; j++) {
                    const double p = sh->exps[i] + sh->exps[j];
                    norm2 += b->coef[pi + i] * b->coef[pi + j] * df / pow(2.0 * p, l) * pow(M_PI / p, 1.5);
                }
            for (int i = 0; i < sh->nprim; i++) b->coef[pi + i] /= sqrt(norm2);
            pi += sh->nprim;
        }
    }
    b->s_start[b->natoms] = si;
    b->f_start[ns] = nf;
    b->nfunctions = nf;
    return CMS_STATUS_SUCCESS;
}

CMSStatus_t CMS_loadChemicalSystem(BasisSet_t b, char *bsfile, char *xyzfile)
{
    CMSStatus_t st;
    cms_err[0] = 0;
    if ((st = read_xyz(b, xyzfile)) != CMS_STATUS_SUCCESS) return st;
    if ((st = read_gbs(b, bsfile)) != CMS_STATUS_SUCCESS) return st;
    if ((st = build_shells(b)) != CMS_STATUS_SUCCESS) return st;
    read_sad(b, bsfile);
    return CMS_STATUS_SUCCESS;
}

int CMS_getNumAtoms(BasisSet_t b) { return b->natoms; }
int CMS_getNumShells(BasisSet_t b) { return b->nshells; }
int CMS_getNumFuncs(BasisSet_t b) { return b->nfunctions; }
int CMS_getNumOccOrb(BasisSet_t b) { return (b->nelectrons - b->Q) / 2; }
int CMS_getFuncStartInd(BasisSet_t b, int s) { return b->f_start[s]; }
int CMS_getFuncEndInd(BasisSet_t b, int s) { return b->f_start[s + 1] - 1; }
int CMS_getShellDim(BasisSet_t b, int s) { return b->f_start[s + 1] - b->f_start[s]; }
int CMS_getMaxShellDim(BasisSet_t b) { return b->maxdim; }
int CMS_getAtomStartInd(BasisSet_t b, int a) { return b->s_start[a]; }
int CMS_getAtomEndInd(BasisSet_t b, int a) { return b->s_start[a + 1] - 1; }
int CMS_getTotalCharge(BasisSet_t b) { return b->Q; }
int CMS_getNneutral(BasisSet_t b) { return b->nelectrons; }
int CMS_getMaxMomentum(BasisSet_t b) { return b->max_momentum; }
int CMS_isSpherical(BasisSet_t b) { return !b->cartesian; }
double CMS_getNucEnergy(BasisSet_t b) { return b->ene_nuc; }

void CMS_getInitialGuess(BasisSet_t b, int atomid, double **guess, int *spos, int *epos)
{
    gbs_elem_t *el = &b->elem[b->eid[atomid] - 1];
    *guess = el->sad;
    *spos = b->f_start[b->s_start[atomid]];
    *epos = b->f_start[b->s_start[atomid + 1]] - 1;
}

void CMS_getShellxyz(BasisSet_t b, int s, double *x, double *y, double *z)
{
    *x = b->xyz[3 * s];
    *y = b->xyz[3 * s + 1];
    *z = b->xyz[3 * s + 2];
}

void CMS_getShellArrays(BasisSet_t b, const int **am, const int **nprim, const int **prim_off,
                        const double **xyz, const double **alpha, const double **coef)
{
    *am = b->am; *nprim = b->nprim; *prim_off = b->prim_off;
    *xyz = b->xyz; *alpha = b->alpha; *coef = b->coef;
}

void CMS_getAtoms(BasisSet_t b, const double **charge, const double **xyz)
{
    *charge = b->charge;
    *xyz = b->axyz;
}
