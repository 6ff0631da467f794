/*
 * ora_ints.c -- CPU oracle integrals (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the arithmetic the reference obtains from Simint
 * (simint_compute_eri at CMS_Simint.c:363,527,595; simint_compute_overlap/ke/
 * potential at CMS_Simint.c:411,441,444) with the McMurchie-Davidson scheme
 * (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
 * sections 9.5, 9.7, 9.9):
 *   Hermite expansion coefficients E^{ij}_t  (eq. 9.5.6-9.5.7)
 *   Hermite Coulomb integrals R^{n}_{tuv}     (eq. 9.9.18-9.9.20)
 *   (ab|cd) = 2 pi^{5/2}/(p q sqrt(p+q)) sum E^{ab}_{tuv} (-1)^{tau+nu+phi}
 *             E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}       (eq. 9.9.33)
 * The 3-centre class (ab|P) uses the reference's unit-shell trick
 * (CMS_Simint.c:131-141): P is paired with an s shell of exponent 0 and
 * coefficient 1, i.e. a single-centre Hermite expansion of P.
 *
 * Boys function F_n(T): 8-term Taylor expansion about a grid of spacing 0.05
 * on [0,40] whose nodes come from the convergent series
 *   F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
 * evaluated in long double; for T >= 40 the asymptotic F_0 = sqrt(pi/T)/2 and
 * upward recursion.  Pinned against mpmath in tests/test_oracle_ints.py.
 *
 * Cartesian component order inside a shell: lexicographic in descending x
 * then y power (xx, xy, xz, yy, yz, zz).  Parity unpinned: the reference's
 * Simint component order is assumed to be the same (CMS_Simint.c cannot be
 * inspected without Simint).
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "ora.h"

#define BOYS_NMAX  32
#define BOYS_TMAX  40.0
#define BOYS_H     0.05
#define BOYS_NPTS  802
#define MAXL_R     12           /* maximum total L handled by the R recursion */

static double boys_tab[BOYS_NPTS][BOYS_NMAX + 1];
static int boys_ready = 0;

static long double boys_series(int n, long double T)
{
    long double term = 1.0L / (2 * n + 1), sum = term;
    for (int k = 1; k < 4000; k++) {
        term *= 2.0L * T / (long double)(2 * n + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return expl(-T) * sum;
}

static void boys_init(void)
{
    if (boys_ready) return;
    #pragma omp critical(ora_boys_init)
    {
        if (!boys_ready) {
            for (int k = 0; k < BOYS_NPTS; k++) {
                long double T = (long double)k * (long double)BOYS_H;
                for (int n = 0; n <= BOYS_NMAX; n++)
                    boys_tab[k][n] = (double) boys_series(n, T);
            }
            boys_ready = 1;
        }
    }
}

void ora_boys(int nmax, double T, double *F)
{
    boys_init();
    if (T < BOYS_TMAX) {
        int k = (int)(T / BOYS_H + 0.5);
        double d = k * BOYS_H - T;   /* F_n(T) = sum_j F_{n+j}(T_k) (T_k - T)^j / j! */
        for (int n = 0; n <= nmax; n++) {
            const double *f = &boys_tab[k][n];
            double s = f[7];
            for (int j = 6; j >= 0; j--) s = f[j] + s * d / (double)(j + 1);
            F[n] = s;
        }
    } else {
        double eT = exp(-T);
        F[0] = 0.5 * sqrt(M_PI / T);
        for (int n = 0; n < nmax; n++) F[n + 1] = ((2 * n + 1) * F[n] - eT) / (2.0 * T);
    }
}

int ora_ncart(int l) { return (l + 1) * (l + 2) / 2; }

void ora_cart_powers(int l, int idx, int *lx, int *ly, int *lz)
{
    int c = 0;
    for (int i = l; i >= 0; i--)
        for (int j = l - i; j >= 0; j--) {
            if (c == idx) { *lx = i; *ly = j; *lz = l - i - j; return; }
            c++;
        }
    *lx = *ly = *lz = -1;
}

/* Hermite expansion coefficients for one Cartesian direction.
 * E[(i*(lb+1)+j)*(la+lb+1)+t] = E^{ij}_t, including exp(-mu X_AB^2). */
static void md_E(int la, int lb, double a, double b, double XAB, double *E)
{
    const int nt = la + lb + 1;
    const double p = a + b, mu = a * b / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB, oo2p = 0.5 / p;
    memset(E, 0, sizeof(double) * (la + 1) * (lb + 1) * nt);
#define EI(i, j, t) E[((i) * (lb + 1) + (j)) * nt + (t)]
    EI(0, 0, 0) = exp(-mu * XAB * XAB);
    for (int i = 0; i <= la; i++) {
        if (i > 0)
            for (int t = 0; t <= i; t++) {
                double v = XPA * EI(i - 1, 0, t);
                if (t > 0) v += oo2p * EI(i - 1, 0, t - 1);
                if (t + 1 <= i - 1) v += (t + 1) * EI(i - 1, 0, t + 1);
                EI(i, 0, t) = v;
            }
        for (int j = 1; j <= lb; j++)
            for (int t = 0; t <= i + j; t++) {
                double v = (t <= i + j - 1) ? XPB * EI(i, j - 1, t) : 0.0;
                if (t > 0) v += oo2p * EI(i, j - 1, t - 1);
                if (t + 1 <= i + j - 1) v += (t + 1) * EI(i, j - 1, t + 1);
                EI(i, j, t) = v;
            }
    }
#undef EI
}

/* Hermite Coulomb integrals R_{tuv} = R^{0}_{tuv} for t+u+v <= L.
 * Output R[(t*(L+1)+u)*(L+1)+v]. */
static void md_R(int L, double alpha, double X, double Y, double Z, double *R)
{
    double F[MAXL_R + 1];
    const int d = L + 1;
    double T = alpha * (X * X + Y * Y + Z * Z);
    ora_boys(L, T, F);
    /* Rn[n][t][u][v] */
    static __thread double Rn[(MAXL_R + 1) * (MAXL_R + 1) * (MAXL_R + 1) * (MAXL_R + 1)];
#define RN(n, t, u, v) Rn[(((n) * d + (t)) * d + (u)) * d + (v)]
    double m2a = -2.0 * alpha, f = 1.0;
    for (int n = 0; n <= L; n++) { RN(n, 0, 0, 0) = f * F[n]; f *= m2a; }
    for (int N = 1; N <= L; N++)
        for (int t = 0; t <= N; t++)
            for (int u = 0; u <= N - t; u++) {
                int v = N - t - u;
                for (int n = 0; n <= L - N; n++) {
                    double r;
                    if (t > 0) {
                        r = X * RN(n + 1, t - 1, u, v);
                        if (t > 1) r += (t - 1) * RN(n + 1, t - 2, u, v);
                    } else if (u > 0) {
                        r = Y * RN(n + 1, t, u - 1, v);
                        if (u > 1) r += (u - 1) * RN(n + 1, t, u - 2, v);
                    } else {
                        r = Z * RN(n + 1, t, u, v - 1);
                        if (v > 1) r += (v - 1) * RN(n + 1, t, u, v - 2);
                    }
                    RN(n, t, u, v) = r;
                }
            }
    for (int t = 0; t <= L; t++)
        for (int u = 0; u <= L - t; u++)
            for (int v = 0; v <= L - t - u; v++)
                R[(t * d + u) * d + v] = RN(0, t, u, v);
#undef RN
}

#define MAXE 1024

void ora_overlap(const ora_shell_t *A, const ora_shell_t *B, double *out)
{
    const int la = A->am, lb = B->am, na = ora_ncart(la), nb = ora_ncart(lb);
    const int nt = la + lb + 1;
    double Ex[MAXE], Ey[MAXE], Ez[MAXE];
    memset(out, 0, sizeof(double) * na * nb);
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
            md_E(la, lb, a, b, A->xyz[0] - B->xyz[0], Ex);
            md_E(la, lb, a, b, A->xyz[1] - B->xyz[1], Ey);
            md_E(la, lb, a, b, A->xyz[2] - B->xyz[2], Ez);
            double pref = A->coef[pa] * B->coef[pb] * pow(M_PI / p, 1.5);
            for (int ia = 0; ia < na; ia++) {
                int ax, ay, az; ora_cart_powers(la, ia, &ax, &ay, &az);
                for (int ib = 0; ib < nb; ib++) {
                    int bx, by, bz; ora_cart_powers(lb, ib, &bx, &by, &bz);
                    out[ia * nb + ib] += pref * Ex[(ax * (lb + 1) + bx) * nt]
                                              * Ey[(ay * (lb + 1) + by) * nt]
                                              * Ez[(az * (lb + 1) + bz) * nt];
                }
            }
        }
}

void ora_kinetic(const ora_shell_t *A, const ora_shell_t *B, double *out)
{
    const int la = A->am, lb = B->am, na = ora_ncart(la), nb = ora_ncart(lb);
    const int lb2 = lb + 2, nt = la + lb2 + 1;
    double Ex[MAXE], Ey[MAXE], Ez[MAXE];
    memset(out, 0, sizeof(double) * na * nb);
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
            md_E(la, lb2, a, b, A->xyz[0] - B->xyz[0], Ex);
            md_E(la, lb2, a, b, A->xyz[1] - B->xyz[1], Ey);
            md_E(la, lb2, a, b, A->xyz[2] - B->xyz[2], Ez);
            double sq = sqrt(M_PI / p);
            double pref = A->coef[pa] * B->coef[pb];
#define S1(E, i, j) ((j) < 0 ? 0.0 : (E)[((i) * (lb2 + 1) + (j)) * nt] * sq)
#define T1(E, i, j) (-2.0 * b * b * S1(E, i, (j) + 2) + b * (2 * (j) + 1) * S1(E, i, j) \
                     - 0.5 * (j) * ((j) - 1) * S1(E, i, (j) - 2))
            for (int ia = 0; ia < na; ia++) {
                int ax, ay, az; ora_cart_powers(la, ia, &ax, &ay, &az);
                for (int ib = 0; ib < nb; ib++) {
                    int bx, by, bz; ora_cart_powers(lb, ib, &bx, &by, &bz);
                    double sx = S1(Ex, ax, bx), sy = S1(Ey, ay, by), sz = S1(Ez, az, bz);
                    double tx = T1(Ex, ax, bx), ty = T1(Ey, ay, by), tz = T1(Ez, az, bz);
                    out[ia * nb + ib] += pref * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                }
            }
#undef S1
#undef T1
        }
}

void ora_nuclear(const ora_shell_t *A, const ora_shell_t *B, int natoms,
                 const double *Z, const double *axyz, double *out)
{
    const int la = A->am, lb = B->am, na = ora_ncart(la), nb = ora_ncart(lb);
    const int nt = la + lb + 1, L = la + lb, d = L + 1;
    double Ex[MAXE], Ey[MAXE], Ez[MAXE];
    double *R = (double *) malloc(sizeof(double) * d * d * d);
    memset(out, 0, sizeof(double) * na * nb);
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
            double P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A->xyz[x] + b * B->xyz[x]) / p;
            md_E(la, lb, a, b, A->xyz[0] - B->xyz[0], Ex);
            md_E(la, lb, a, b, A->xyz[1] - B->xyz[1], Ey);
            md_E(la, lb, a, b, A->xyz[2] - B->xyz[2], Ez);
            double cab = A->coef[pa] * B->coef[pb] * 2.0 * M_PI / p;
            for (int c = 0; c < natoms; c++) {
                md_R(L, p, P[0] - axyz[3 * c], P[1] - axyz[3 * c + 1], P[2] - axyz[3 * c + 2], R);
                double pref = -Z[c] * cab;
                for (int ia = 0; ia < na; ia++) {
                    int ax, ay, az; ora_cart_powers(la, ia, &ax, &ay, &az);
                    for (int ib = 0; ib < nb; ib++) {
                        int bx, by, bz; ora_cart_powers(lb, ib, &bx, &by, &bz);
                        const double *ex = &Ex[(ax * (lb + 1) + bx) * nt];
                        const double *ey = &Ey[(ay * (lb + 1) + by) * nt];
                        const double *ez = &Ez[(az * (lb + 1) + bz) * nt];
                        double s = 0.0;
                        for (int t = 0; t <= ax + bx; t++)
                            for (int u = 0; u <= ay + by; u++)
                                for (int v = 0; v <= az + bz; v++)
                                    s += ex[t] * ey[u] * ez[v] * R[(t * d + u) * d + v];
                        out[ia * nb + ib] += pref * s;
                    }
                }
            }
        }
    free(R);
}

/* Single-centre Hermite expansion of x^l exp(-g x^2): Ec[l][tau]. */
static void md_E1(int l, double g, double *E)
{
    memset(E, 0, sizeof(double) * (l + 1) * (l + 1));
    E[0] = 1.0;
    for (int i = 1; i <= l; i++)
        for (int t = 0; t <= i; t++) {
            double v = 0.0;
            if (t > 0) v += 0.5 / g * E[(i - 1) * (l + 1) + t - 1];
            if (t + 1 <= i - 1) v += (t + 1) * E[(i - 1) * (l + 1) + t + 1];
            E[i * (l + 1) + t] = v;
        }
}

void ora_eri3(const ora_shell_t *A, const ora_shell_t *B, const ora_shell_t *C, double *out)
{
    const int la = A->am, lb = B->am, lc = C->am;
    const int na = ora_ncart(la), nb = ora_ncart(lb), nc = ora_ncart(lc);
    const int nt = la + lb + 1, L = la + lb + lc, d = L + 1, lab = la + lb;
    double Ex[MAXE], Ey[MAXE], Ez[MAXE], Ecx[64], Ecy[64], Ecz[64];
    double *R = (double *) malloc(sizeof(double) * d * d * d);
    double *W = (double *) malloc(sizeof(double) * nc * (lab + 1) * (lab + 1) * (lab + 1));
    const int dw = lab + 1;
    memset(out, 0, sizeof(double) * na * nb * nc);
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->alpha[pa], b = B->alpha[pb], p = a + b;
            double P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A->xyz[x] + b * B->xyz[x]) / p;
            md_E(la, lb, a, b, A->xyz[0] - B->xyz[0], Ex);
            md_E(la, lb, a, b, A->xyz[1] - B->xyz[1], Ey);
            md_E(la, lb, a, b, A->xyz[2] - B->xyz[2], Ez);
            for (int pc = 0; pc < C->nprim; pc++) {
                double g = C->alpha[pc];
                double alpha = p * g / (p + g);
                md_R(L, alpha, P[0] - C->xyz[0], P[1] - C->xyz[1], P[2] - C->xyz[2], R);
                md_E1(lc, g, Ecx);
                memcpy(Ecy, Ecx, sizeof(double) * (lc + 1) * (lc + 1));
                memcpy(Ecz, Ecx, sizeof(double) * (lc + 1) * (lc + 1));
                double pref = A->coef[pa] * B->coef[pb] * C->coef[pc]
                            * 2.0 * pow(M_PI, 2.5) / (p * g * sqrt(p + g));
                /* W[ic][t][u][v] = sum (-1)^{tau+nu+phi} Ec R[t+tau][u+nu][v+phi] */
                for (int ic = 0; ic < nc; ic++) {
                    int cx, cy, cz; ora_cart_powers(lc, ic, &cx, &cy, &cz);
                    for (int t = 0; t <= lab; t++)
                        for (int u = 0; u <= lab - t; u++)
                            for (int v = 0; v <= lab - t - u; v++) {
                                double s = 0.0;
                                for (int tau = 0; tau <= cx; tau++) {
                                    double ex = Ecx[cx * (lc + 1) + tau];
                                    if (ex == 0.0) continue;
                                    for (int nu = 0; nu <= cy; nu++) {
                                        double ey = Ecy[cy * (lc + 1) + nu];
                                        if (ey == 0.0) continue;
                                        for (int phi = 0; phi <= cz; phi++) {
                                            double ez = Ecz[cz * (lc + 1) + phi];
                                            if (ez == 0.0) continue;
                                            double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                            s += sg * ex * ey * ez * R[((t + tau) * d + u + nu) * d + v + phi];
                                        }
                                    }
                                }
                                W[((ic * dw + t) * dw + u) * dw + v] = s;
                            }
                }
                for (int ia = 0; ia < na; ia++) {
                    int ax, ay, az; ora_cart_powers(la, ia, &ax, &ay, &az);
                    for (int ib = 0; ib < nb; ib++) {
                        int bx, by, bz; ora_cart_powers(lb, ib, &bx, &by, &bz);
                        const double *ex = &Ex[(ax * (lb + 1) + bx) * nt];
                        const double *ey = &Ey[(ay * (lb + 1) + by) * nt];
                        const double *ez = &Ez[(az * (lb + 1) + bz) * nt];
                        for (int ic = 0; ic < nc; ic++) {
                            double s = 0.0;
                            for (int t = 0; t <= ax + bx; t++)
                                for (int u = 0; u <= ay + by; u++)
                                    for (int v = 0; v <= az + bz; v++)
                                        s += ex[t] * ey[u] * ez[v] * W[((ic * dw + t) * dw + u) * dw + v];
                            out[(ia * nb + ib) * nc + ic] += pref * s;
                        }
                    }
                }
            }
        }
    free(R);
    free(W);
}

void ora_eri2(const ora_shell_t *C, const ora_shell_t *D, double *out)
{
    const int lc = C->am, ld = D->am, nc = ora_ncart(lc), nd = ora_ncart(ld);
    const int L = lc + ld, d = L + 1;
    double Ec[64], Ed[64];
    double *R = (double *) malloc(sizeof(double) * d * d * d);
    memset(out, 0, sizeof(double) * nc * nd);
    for (int pc = 0; pc < C->nprim; pc++)
        for (int pd = 0; pd < D->nprim; pd++) {
            double g = C->alpha[pc], h = D->alpha[pd];
            double alpha = g * h / (g + h);
            md_R(L, alpha, C->xyz[0] - D->xyz[0], C->xyz[1] - D->xyz[1], C->xyz[2] - D->xyz[2], R);
            md_E1(lc, g, Ec);
            md_E1(ld, h, Ed);
            double pref = C->coef[pc] * D->coef[pd] * 2.0 * pow(M_PI, 2.5) / (g * h * sqrt(g + h));
            for (int ic = 0; ic < nc; ic++) {
                int cx, cy, cz; ora_cart_powers(lc, ic, &cx, &cy, &cz);
                for (int id = 0; id < nd; id++) {
                    int dx, dy, dz; ora_cart_powers(ld, id, &dx, &dy, &dz);
                    double s = 0.0;
                    for (int t = 0; t <= cx; t++)
                        for (int u = 0; u <= cy; u++)
                            for (int v = 0; v <= cz; v++) {
                                double e1 = Ec[cx * (lc + 1) + t] * Ec[cy * (lc + 1) + u] * Ec[cz * (lc + 1) + v];
                                if (e1 == 0.0) continue;
                                for (int tau = 0; tau <= dx; tau++)
                                    for (int nu = 0; nu <= dy; nu++)
                                        for (int phi = 0; phi <= dz; phi++) {
                                            double e2 = Ed[dx * (ld + 1) + tau] * Ed[dy * (ld + 1) + nu] * Ed[dz * (ld + 1) + phi];
                                            if (e2 == 0.0) continue;
                                            double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                            s += sg * e1 * e2 * R[((t + tau) * d + u + nu) * d + v + phi];
                                        }
                            }
                    out[ic * nd + id] += pref * s;
                }
            }
        }
    free(R);
}

double ora_eri4_elem(const ora_shell_t *A, const ora_shell_t *B,
                     const ora_shell_t *C, const ora_shell_t *D,
                     int ia, int ib, int ic, int id)
{
    const int la = A->am, lb = B->am, lc = C->am, ld = D->am;
    const int L = la + lb + lc + ld, d = L + 1;
    const int nt1 = la + lb + 1, nt2 = lc + ld + 1;
    int ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
    ora_cart_powers(la, ia, &ax, &ay, &az);
    ora_cart_powers(lb, ib, &bx, &by, &bz);
    ora_cart_powers(lc, ic, &cx, &cy, &cz);
    ora_cart_powers(ld, id, &dx, &dy, &dz);
    double E1x[MAXE], E1y[MAXE], E1z[MAXE], E2x[MAXE], E2y[MAXE], E2z[MAXE];
    double *R = (double *) malloc(sizeof(double) * d * d * d);
    double res = 0.0;
    for (int pa = 0; pa < A->nprim; pa++)
        for (int pb = 0; pb < B->nprim; pb++) {
            double a = A->alpha[pa], b = B->alpha[pb], p = a + b, P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A->xyz[x] + b * B->xyz[x]) / p;
            md_E(la, lb, a, b, A->xyz[0] - B->xyz[0], E1x);
            md_E(la, lb, a, b, A->xyz[1] - B->xyz[1], E1y);
            md_E(la, lb, a, b, A->xyz[2] - B->xyz[2], E1z);
            const double *e1x = &E1x[(ax * (lb + 1) + bx) * nt1];
            const double *e1y = &E1y[(ay * (lb + 1) + by) * nt1];
            const double *e1z = &E1z[(az * (lb + 1) + bz) * nt1];
            for (int pc = 0; pc < C->nprim; pc++)
                for (int pd = 0; pd < D->nprim; pd++) {
                    double c = C->alpha[pc], dd = D->alpha[pd], q = c + dd, Q[3];
                    for (int x = 0; x < 3; x++) Q[x] = (c * C->xyz[x] + dd * D->xyz[x]) / q;
                    md_E(lc, ld, c, dd, C->xyz[0] - D->xyz[0], E2x);
                    md_E(lc, ld, c, dd, C->xyz[1] - D->xyz[1], E2y);
                    md_E(lc, ld, c, dd, C->xyz[2] - D->xyz[2], E2z);
                    const double *e2x = &E2x[(cx * (ld + 1) + dx) * nt2];
                    const double *e2y = &E2y[(cy * (ld + 1) + dy) * nt2];
                    const double *e2z = &E2z[(cz * (ld + 1) + dz) * nt2];
                    double alpha = p * q / (p + q);
                    md_R(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
                    double s = 0.0;
                    for (int t = 0; t <= ax + bx; t++)
                        for (int u = 0; u <= ay + by; u++)
                            for (int v = 0; v <= az + bz; v++) {
                                double e1 = e1x[t] * e1y[u] * e1z[v];
                                for (int tau = 0; tau <= cx + dx; tau++)
                                    for (int nu = 0; nu <= cy + dy; nu++)
                                        for (int phi = 0; phi <= cz + dz; phi++) {
                                            double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                            s += sg * e1 * e2x[tau] * e2y[nu] * e2z[phi]
                                               * R[((t + tau) * d + u + nu) * d + v + phi];
                                        }
                            }
                    res += A->coef[pa] * B->coef[pb] * C->coef[pc] * D->coef[pd]
                         * 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * s;
                }
        }
    free(R);
    return res;
}
