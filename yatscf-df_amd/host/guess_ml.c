/*
 * guess_ml.c -- learned per-atom-pair correction to the SAD initial density.
 *
 * The reference's initial guess is the superposition of atomic densities
 * (TinySCF_get_initial_guess, TinySCF.c:324-364) followed by D2Cocc.  Here an
 * MLP, evaluated on the GPU (ygpu_mlp_forward), predicts a correction dD to
 * the SAD density for every atom pair within YSML_RCUT:
 *
 *   inputs  (YSML_NIN): radial-Gaussian descriptors of both atoms' environments
 *           (YSML_NCEN centres on [0.5, 6] A, smooth cosine cutoff at 6 A, one
 *           channel per neighbour element H / C), Gaussians of the pair
 *           distance, and a one-hot pair type (C-C, C-H, H-H, C self, H self);
 *   outputs (YSML_NOUT): invariant coefficients from which the pair block is
 *           rebuilt along the unit vector e from atom A to atom B:
 *             s_a s_b          dD = c
 *             s_a p_b,i        dD = c e_i
 *             p_a,i s_b        dD = c e_i
 *             p_a,i p_b,j      dD = c_sigma e_i e_j + c_pi (delta_ij - e_i e_j)
 *           and for the same atom: s_a s_b (symmetrised), p_a,i p_b,j = c delta_ij.
 *           d shells keep the SAD value (zero).
 * The block rebuild is covariant under rotations, so the model sees only
 * rotation- and translation-invariant inputs.  The corrected density is then
 * made a valid RHF density: C_occ = the n_occ leading eigenvectors of
 * S^1/2 D S^1/2, back-transformed with X = S^-1/2, and D = C_occ C_occ^T.
 *
 * Only H and C atoms are modelled; any other element, a missing or malformed
 * weights file, or a failed device call leaves the SAD guess (and the
 * reference's D2Cocc) in place.
 *
 * Weights file (little-endian): "YSML", int32 version 1, n_in, n_hidden,
 * n_layers (3), n_out, n_types; float32 in_mean[n_in], in_std[n_in],
 * out_mean[n_types*n_out], out_std[n_types*n_out], then the network
 * parameters in the order of ygpu_mlp_forward.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "TinySCF.h"
#include "guess_ml.h"
#include "utils.h"
#include "yatscf_gpu.h"

#define BOHR_TO_ANG 0.52917720859

static const double YSML_MU0 = 0.5;

static int elem_index(double Z)
{
    if (fabs(Z - 1.0) < 1e-9) return 0;
    if (fabs(Z - 6.0) < 1e-9) return 1;
    return -1;
}

static double centre(int k) { return YSML_MU0 + k * (YSML_RCUT - YSML_MU0) / (YSML_NCEN - 1); }

static double eta(void)
{
    const double h = (YSML_RCUT - YSML_MU0) / (YSML_NCEN - 1);
    return 0.5 / (h * h);
}

static double fcut(double r) { return r < YSML_RCUT ? 0.5 * (cos(M_PI * r / YSML_RCUT) + 1.0) : 0.0; }

/* shells of one atom split by angular momentum; returns 0 if unsupported */
typedef struct {
    int ns, np, s_bf[4], p_bf[4];
} AtomShells;

static int atom_shells(const ysml_mol_t *M, int a, AtomShells *out)
{
    memset(out, 0, sizeof(*out));
    for (int sh = M->atom_shell_start[a]; sh < M->atom_shell_start[a + 1]; sh++) {
        if (M->am[sh] == 0) {
            if (out->ns >= 4) return 0;
            out->s_bf[out->ns++] = M->shell_bf_start[sh];
        } else if (M->am[sh] == 1) {
            if (out->np >= 4) return 0;
            out->p_bf[out->np++] = M->shell_bf_start[sh];
        } else if (M->am[sh] > 2) {
            return 0;
        }
    }
    return 1;
}

/* pair type from the (canonically ordered) element indices */
static int pair_type(int ea, int eb, int self)
{
    if (self) return ea == 1 ? YSML_T_CSELF : YSML_T_HSELF;
    if (ea == 1 && eb == 1) return YSML_T_CC;
    if (ea == 1 && eb == 0) return YSML_T_CH;
    return YSML_T_HH;
}

int ysml_supported(const ysml_mol_t *M)
{
    const double *Z = M->Z;
    for (int a = 0; a < M->natoms; a++) {
        AtomShells sh;
        if (elem_index(Z[a]) < 0 || !atom_shells(M, a, &sh)) return 0;
        if (sh.ns > 3 || sh.np > 2) return 0;
    }
    return 1;
}

/* pairs (A, B): every atom with itself, and every A != B within the cutoff once, ordered
 * C before H and by index within one element.  pairs[2k], pairs[2k+1]. */
int ysml_pairs(const ysml_mol_t *M, int *pairs, int max_pairs)
{
    const double *Z = M->Z, *xyz = M->xyz;
    int n = 0;
    for (int a = 0; a < M->natoms; a++)
        for (int b = a; b < M->natoms; b++) {
            const double dx = xyz[3 * b] - xyz[3 * a], dy = xyz[3 * b + 1] - xyz[3 * a + 1],
                         dz = xyz[3 * b + 2] - xyz[3 * a + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz) * BOHR_TO_ANG;
            if (a != b && r >= YSML_RCUT) continue;
            int A = a, B = b;
            if (elem_index(Z[a]) == 0 && elem_index(Z[b]) == 1) { A = b; B = a; }
            if (pairs && n < max_pairs) { pairs[2 * n] = A; pairs[2 * n + 1] = B; }
            n++;
        }
    return n;
}

static void env_desc(const ysml_mol_t *M, int a, double *g)
{
    const double *Z = M->Z, *xyz = M->xyz;
    const double et = eta();
    memset(g, 0, sizeof(double) * YSML_NEL * YSML_NCEN);
    for (int b = 0; b < M->natoms; b++) {
        if (b == a) continue;
        const double dx = xyz[3 * b] - xyz[3 * a], dy = xyz[3 * b + 1] - xyz[3 * a + 1], dz = xyz[3 * b + 2] - xyz[3 * a + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz) * BOHR_TO_ANG;
        const int e = elem_index(Z[b]);
        if (r >= YSML_RCUT || e < 0) continue;
        const double fc = fcut(r);
        for (int k = 0; k < YSML_NCEN; k++) {
            const double d = r - centre(k);
            g[e * YSML_NCEN + k] += exp(-et * d * d) * fc;
        }
    }
}

int ysml_features(const ysml_mol_t *M, int npairs, const int *pairs, float *X)
{
    const double *Z = M->Z, *xyz = M->xyz;
    const int na = M->natoms, ng = YSML_NEL * YSML_NCEN;
    double *G = (double *) malloc(sizeof(double) * (size_t) na * ng);
    if (!G) return 1;
    for (int a = 0; a < na; a++) env_desc(M, a, G + (size_t) a * ng);
    const double et = eta();
    for (int p = 0; p < npairs; p++) {
        const int A = pairs[2 * p], B = pairs[2 * p + 1];
        float *x = X + (size_t) p * YSML_NIN;
        for (int k = 0; k < ng; k++) {
            x[k] = (float) G[(size_t) A * ng + k];
            x[ng + k] = (float) G[(size_t) B * ng + k];
        }
        const double dx = xyz[3 * B] - xyz[3 * A], dy = xyz[3 * B + 1] - xyz[3 * A + 1], dz = xyz[3 * B + 2] - xyz[3 * A + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz) * BOHR_TO_ANG;
        for (int k = 0; k < YSML_NCEN; k++) {
            const double d = r - centre(k);
            x[2 * ng + k] = (float) exp(-et * d * d);
        }
        for (int t = 0; t < YSML_NTYPES; t++) x[2 * ng + YSML_NCEN + t] = 0.f;
        x[2 * ng + YSML_NCEN + pair_type(elem_index(Z[A]), elem_index(Z[B]), A == B)] = 1.f;
    }
    free(G);
    return 0;
}

/* unit vector from A to B (zero for A == B) */
static void pair_axis(const ysml_mol_t *M, int A, int B, double e[3])
{
    const double *xyz = M->xyz;
    double d[3] = {xyz[3 * B] - xyz[3 * A], xyz[3 * B + 1] - xyz[3 * A + 1], xyz[3 * B + 2] - xyz[3 * A + 2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; i++) e[i] = (A == B || r == 0.0) ? 0.0 : d[i] / r;
}

/* Visit every coefficient of pair (A, B): for coefficient q, op(q, kind, bf_a, bf_b, ...).
 * kind: 0 ss, 1 sp, 2 ps, 3 pp sigma, 4 pp pi, 5 self ss, 6 self pp */
typedef void (*coef_fn)(void *ctx, int q, int kind, int bfa, int bfb, const double e[3]);

static int visit_coefs(const ysml_mol_t *M, int A, int B, coef_fn fn, void *ctx)
{
    AtomShells sa, sb;
    if (!atom_shells(M, A, &sa) || !atom_shells(M, B, &sb)) return -1;
    double e[3];
    pair_axis(M, A, B, e);
    int q = 0;
    if (A == B) {
        for (int i = 0; i < sa.ns; i++)
            for (int j = 0; j < sa.ns; j++) fn(ctx, q++, 5, sa.s_bf[i], sa.s_bf[j], e);
        for (int i = 0; i < sa.np; i++)
            for (int j = 0; j < sa.np; j++) fn(ctx, q++, 6, sa.p_bf[i], sa.p_bf[j], e);
        return q;
    }
    for (int i = 0; i < sa.ns; i++)
        for (int j = 0; j < sb.ns; j++) fn(ctx, q++, 0, sa.s_bf[i], sb.s_bf[j], e);
    for (int i = 0; i < sa.ns; i++)
        for (int j = 0; j < sb.np; j++) fn(ctx, q++, 1, sa.s_bf[i], sb.p_bf[j], e);
    for (int i = 0; i < sa.np; i++)
        for (int j = 0; j < sb.ns; j++) fn(ctx, q++, 2, sa.p_bf[i], sb.s_bf[j], e);
    for (int i = 0; i < sa.np; i++)
        for (int j = 0; j < sb.np; j++) {
            fn(ctx, q++, 3, sa.p_bf[i], sb.p_bf[j], e);
            fn(ctx, q++, 4, sa.p_bf[i], sb.p_bf[j], e);
        }
    return q;
}

typedef struct {
    const double *M;     /* nbf x nbf matrix (dD) */
    int nbf;
    float *y, *m;        /* coefficients and mask of this pair */
} ProjCtx;

static void project_one(void *vc, int q, int kind, int a, int b, const double e[3])
{
    ProjCtx *c = (ProjCtx *) vc;
    const int n = c->nbf;
    const double *M = c->M;
    double v = 0.0;
    switch (kind) {
    case 0: case 5: v = M[(size_t) a * n + b]; break;
    case 1: for (int i = 0; i < 3; i++) v += M[(size_t) a * n + b + i] * e[i]; break;
    case 2: for (int i = 0; i < 3; i++) v += M[(size_t) (a + i) * n + b] * e[i]; break;
    case 3:
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) v += M[(size_t) (a + i) * n + b + j] * e[i] * e[j];
        break;
    case 4: {
        double tr = 0.0, sg = 0.0;
        for (int i = 0; i < 3; i++) tr += M[(size_t) (a + i) * n + b + i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) sg += M[(size_t) (a + i) * n + b + j] * e[i] * e[j];
        v = 0.5 * (tr - sg);
        break;
    }
    case 6: for (int i = 0; i < 3; i++) v += M[(size_t) (a + i) * n + b + i]; v /= 3.0; break;
    }
    c->y[q] = (float) v;
    c->m[q] = 1.f;
}

int ysml_targets(const ysml_mol_t *M, int npairs, const int *pairs, const double *dD, float *Y, float *Mask)
{
    ProjCtx c = {dD, M->nbf, NULL, NULL};
    for (int p = 0; p < npairs; p++) {
        c.y = Y + (size_t) p * YSML_NOUT;
        c.m = Mask + (size_t) p * YSML_NOUT;
        for (int q = 0; q < YSML_NOUT; q++) { c.y[q] = 0.f; c.m[q] = 0.f; }
        if (visit_coefs(M, pairs[2 * p], pairs[2 * p + 1], project_one, &c) < 0) return 1;
    }
    return 0;
}

typedef struct {
    double *M;
    int nbf;
    const float *y;
} AsmCtx;

static void assemble_one(void *vc, int q, int kind, int a, int b, const double e[3])
{
    AsmCtx *c = (AsmCtx *) vc;
    const int n = c->nbf;
    double *M = c->M;
    const double v = c->y[q];
    /* every block is written together with its transpose (dD symmetric) */
#define ADD(i, j, x) do { M[(size_t) (i) * n + (j)] += (x); if ((i) != (j)) M[(size_t) (j) * n + (i)] += (x); } while (0)
    switch (kind) {
    case 0: ADD(a, b, v); break;
    case 5: ADD(a, b, 0.5 * v); if (a == b) M[(size_t) a * n + a] += 0.5 * v; break;
    case 1: for (int i = 0; i < 3; i++) ADD(a, b + i, v * e[i]); break;
    case 2: for (int i = 0; i < 3; i++) ADD(a + i, b, v * e[i]); break;
    case 3:
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ADD(a + i, b + j, v * e[i] * e[j]);
        break;
    case 4:
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ADD(a + i, b + j, v * ((i == j ? 1.0 : 0.0) - e[i] * e[j]));
        break;
    case 6:
        for (int i = 0; i < 3; i++) {
            const int r = a + i, s = b + i;
            if (r == s) M[(size_t) r * n + r] += v;
            else { M[(size_t) r * n + s] += 0.5 * v; M[(size_t) s * n + r] += 0.5 * v; }
        }
        break;
    }
#undef ADD
}

int ysml_assemble(const ysml_mol_t *M, int npairs, const int *pairs, const float *Y, double *dD)
{
    AsmCtx c = {dD, M->nbf, NULL};
    memset(dD, 0, sizeof(double) * (size_t) M->nbf * M->nbf);
    for (int p = 0; p < npairs; p++) {
        c.y = Y + (size_t) p * YSML_NOUT;
        if (visit_coefs(M, pairs[2 * p], pairs[2 * p + 1], assemble_one, &c) < 0) return 1;
    }
    return 0;
}

/* C_occ and D = C_occ C_occ^T from the n_occ leading eigenvectors of S^1/2 D_in S^1/2 */
int ysml_project_density(TinySCF_t T, const double *Din)
{
    /* the n_occ leading eigenvectors of S^1/2 D S^1/2, on the device */
    return ygpu_project_guess((ygpu_ctx_t *) T->gpu, T->n_occ, Din, T->D_mat, T->Cocc_mat, T->eigval);
}

/* ---- weights */
struct ysml_model {
    int n_in, n_hidden, n_layers, n_out, n_types;
    float *in_mean, *in_std, *out_mean, *out_std, *params;
    size_t n_params;
};

static size_t n_params_of(int n_in, int nh, int n_out)
{
    return (size_t) n_in * nh + nh + 2 * ((size_t) nh * nh + nh) + (size_t) nh * n_out + n_out;
}

void ysml_free(ysml_model_t *m)
{
    if (!m) return;
    free(m->in_mean); free(m->in_std); free(m->out_mean); free(m->out_std); free(m->params);
    free(m);
}

ysml_model_t *ysml_load(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    char magic[4];
    int hdr[6];
    ysml_model_t *m = NULL;
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "YSML", 4) != 0 || fread(hdr, sizeof(int), 6, f) != 6 ||
        hdr[0] != 1 || hdr[1] != YSML_NIN || hdr[3] != 3 || hdr[4] != YSML_NOUT || hdr[5] != YSML_NTYPES ||
        hdr[2] <= 0 || hdr[2] > 4096)
        goto fail;
    m = (ysml_model_t *) calloc(1, sizeof(*m));
    if (!m) goto fail;
    m->n_in = hdr[1]; m->n_hidden = hdr[2]; m->n_layers = hdr[3]; m->n_out = hdr[4]; m->n_types = hdr[5];
    m->n_params = n_params_of(m->n_in, m->n_hidden, m->n_out);
    m->in_mean = (float *) malloc(sizeof(float) * m->n_in);
    m->in_std = (float *) malloc(sizeof(float) * m->n_in);
    m->out_mean = (float *) malloc(sizeof(float) * m->n_types * m->n_out);
    m->out_std = (float *) malloc(sizeof(float) * m->n_types * m->n_out);
    m->params = (float *) malloc(sizeof(float) * m->n_params);
    if (!m->in_mean || !m->in_std || !m->out_mean || !m->out_std || !m->params) goto fail;
    if (fread(m->in_mean, sizeof(float), m->n_in, f) != (size_t) m->n_in ||
        fread(m->in_std, sizeof(float), m->n_in, f) != (size_t) m->n_in ||
        fread(m->out_mean, sizeof(float), (size_t) m->n_types * m->n_out, f) != (size_t) m->n_types * m->n_out ||
        fread(m->out_std, sizeof(float), (size_t) m->n_types * m->n_out, f) != (size_t) m->n_types * m->n_out ||
        fread(m->params, sizeof(float), m->n_params, f) != m->n_params)
        goto fail;
    fclose(f);
    return m;
fail:
    fclose(f);
    ysml_free(m);
    return NULL;
}

/* dD predicted for the engine's molecule (nbf x nbf, symmetric) */
int ysml_predict(const ysml_mol_t *M, void *gpu, const ysml_model_t *m, double *dD)
{
    const double *Z = M->Z;
    const int np = ysml_pairs(M, NULL, 0);
    int *pairs = (int *) malloc(sizeof(int) * 2 * (size_t) np);
    float *X = (float *) malloc(sizeof(float) * (size_t) np * YSML_NIN);
    float *Y = (float *) malloc(sizeof(float) * (size_t) np * YSML_NOUT);
    int st = (pairs && X && Y) ? 0 : 1;
    if (!st) {
        ysml_pairs(M, pairs, np);
        st = ysml_features(M, np, pairs, X);
    }
    if (!st) {
        for (size_t p = 0; p < (size_t) np; p++)
            for (int k = 0; k < YSML_NIN; k++)
                X[p * YSML_NIN + k] = (X[p * YSML_NIN + k] - m->in_mean[k]) / m->in_std[k];
        st = ygpu_mlp_forward((ygpu_ctx_t *) gpu, m->n_in, m->n_hidden, m->n_out, m->params, np, X, Y);
    }
    if (!st) {
        for (size_t p = 0; p < (size_t) np; p++) {
            const int A = pairs[2 * p], B = pairs[2 * p + 1];
            const int t = pair_type(elem_index(Z[A]), elem_index(Z[B]), A == B);
            for (int q = 0; q < YSML_NOUT; q++)
                Y[p * YSML_NOUT + q] = Y[p * YSML_NOUT + q] * m->out_std[t * YSML_NOUT + q] + m->out_mean[t * YSML_NOUT + q];
        }
        st = ysml_assemble(M, np, pairs, Y, dD);
    }
    free(pairs); free(X); free(Y);
    return st;
}

void ysml_mol_of(TinySCF_t T, ysml_mol_t *M)
{
    const int *am, *nprim, *poff;
    const double *sxyz, *alpha, *coef, *Z, *xyz;
    CMS_getShellArrays(T->basis, &am, &nprim, &poff, &sxyz, &alpha, &coef);
    CMS_getAtoms(T->basis, &Z, &xyz);
    M->natoms = T->natoms;
    M->Z = Z;
    M->xyz = xyz;
    M->am = am;
    M->shell_bf_start = T->shell_bf_sind;
    M->nbf = T->nbasfuncs;
    M->atom_shell_start = NULL;
}

int TinySCF_learned_guess(TinySCF_t T, const char *path)
{
    ysml_mol_t M;
    ysml_mol_of(T, &M);
    int *ss = (int *) malloc(sizeof(int) * (T->natoms + 1));
    if (!ss) return 4;
    for (int a = 0; a < T->natoms; a++) ss[a] = CMS_getAtomStartInd(T->basis, a);
    ss[T->natoms] = T->nshells;
    M.atom_shell_start = ss;
    if (!ysml_supported(&M)) { free(ss); return 1; }
    ysml_model_t *m = ysml_load(path);
    if (!m) { free(ss); return 2; }
    const size_t ms = (size_t) T->mat_size;
    double *dD = (double *) malloc(DBL_SIZE * ms);
    int st = dD ? ysml_predict(&M, T->gpu, m, dD) : 3;
    if (!st) {
        for (size_t i = 0; i < ms; i++) dD[i] += T->D_mat[i];
        st = ysml_project_density(T, dD);
    }
    free(dD);
    free(ss);
    ysml_free(m);
    return st;
}
