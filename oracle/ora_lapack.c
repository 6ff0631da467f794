/*
 * ora_lapack.c -- BLAS/LAPACK provider for the CPU oracle (TEST INFRASTRUCTURE).
 *
 * The reference calls Intel MKL (cblas_dgemm, LAPACKE_dsyev, LAPACKE_dpstrf,
 * LAPACKE_dgesv; call sites listed at build_DF_tensor.c:217,232,250,
 * build_Fock.c:276,321,331, build_density.c:54,56,67, TinySCF.c:65,76,295,
 * DIIS.c:117).  The oracle loads the MKL runtime with dlopen (with the GNU
 * OpenMP threading layer, so it shares libgomp with the oracle's own OpenMP)
 * and uses our own prototypes.  If MKL is absent it falls back to simple
 * built-in routines (blocked OpenMP GEMM, cyclic Jacobi eigensolver, unblocked
 * pivoted Cholesky, Gaussian elimination with partial pivoting); which one ran
 * is reported by ora_lapack_is_mkl().
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>
#include "ora.h"

typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double *, int,
                         const double *, int, double, double *, int);
typedef int (*dsyev_fn)(int, char, char, int, double *, int, double *);
typedef int (*dpstrf_fn)(int, char, int, double *, int, int *, int *, double);
typedef int (*dgesv_fn)(int, int, int, double *, int, int *, double *, int);
typedef void (*setnt_fn)(int);

static dgemm_fn p_dgemm;
static dsyev_fn p_dsyev;
static dpstrf_fn p_dpstrf;
static dgesv_fn p_dgesv;
static setnt_fn p_setnt;
static int lapack_state = -1;   /* -1 not initialised, 0 fallback, 1 MKL */
static int n_threads = 0;

int ora_lapack_init(void)
{
    if (lapack_state >= 0) return lapack_state;
    lapack_state = 0;
    if (getenv("ORA_NO_MKL") == NULL) {
        setenv("MKL_THREADING_LAYER", "GNU", 0);
        setenv("MKL_INTERFACE_LAYER", "LP64", 0);
        const char *names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "libmkl_rt.so.2",
                               "/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", NULL};
        void *h = NULL;
        for (int i = 0; names[i] && !h; i++) h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
        if (h) {
            p_dgemm  = (dgemm_fn) dlsym(h, "cblas_dgemm");
            p_dsyev  = (dsyev_fn) dlsym(h, "LAPACKE_dsyev");
            p_dpstrf = (dpstrf_fn) dlsym(h, "LAPACKE_dpstrf");
            p_dgesv  = (dgesv_fn) dlsym(h, "LAPACKE_dgesv");
            p_setnt  = (setnt_fn) dlsym(h, "MKL_Set_Num_Threads");
            if (p_dgemm && p_dsyev && p_dpstrf && p_dgesv) lapack_state = 1;
        }
    }
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    if (lapack_state == 1 && p_setnt) p_setnt(n_threads);
    return lapack_state;
}

int ora_lapack_is_mkl(void) { return ora_lapack_init() == 1; }

void ora_set_threads(int nt)
{
    if (nt <= 0) nt = omp_get_num_procs();
    n_threads = nt;
    omp_set_num_threads(nt);
    if (ora_lapack_init() == 1 && p_setnt) p_setnt(nt);
}

int ora_get_threads(void) { return n_threads > 0 ? n_threads : omp_get_max_threads(); }

/* ---------------- fallbacks ---------------- */
static void gemm_fallback(int ta, int tb, int m, int n, int k, double alpha,
                          const double *A, int lda, const double *B, int ldb,
                          double beta, double *C, int ldc)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int l = 0; l < k; l++) {
                double a = ta ? A[(size_t) l * lda + i] : A[(size_t) i * lda + l];
                double b = tb ? B[(size_t) j * ldb + l] : B[(size_t) l * ldb + j];
                s += a * b;
            }
            C[(size_t) i * ldc + j] = alpha * s + (beta == 0.0 ? 0.0 : beta * C[(size_t) i * ldc + j]);
        }
    }
}

/* cyclic Jacobi; eigenvalues ascending, eigenvectors as columns (row-major A[i*n+j] = V(i,j)) */
static int syev_fallback(int n, double *A, double *w)
{
    double *V = (double *) calloc((size_t) n * n, sizeof(double));
    for (int i = 0; i < n; i++) V[(size_t) i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < n; i++) for (int j = i + 1; j < n; j++) off += A[(size_t) i * n + j] * A[(size_t) i * n + j];
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(size_t) p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[(size_t) p * n + p], aqq = A[(size_t) q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < n; r++) {
                    double arp = A[(size_t) r * n + p], arq = A[(size_t) r * n + q];
                    A[(size_t) r * n + p] = c * arp - s * arq;
                    A[(size_t) r * n + q] = s * arp + c * arq;
                }
                for (int r = 0; r < n; r++) {
                    double apr = A[(size_t) p * n + r], aqr = A[(size_t) q * n + r];
                    A[(size_t) p * n + r] = c * apr - s * aqr;
                    A[(size_t) q * n + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < n; r++) {
                    double vrp = V[(size_t) r * n + p], vrq = V[(size_t) r * n + q];
                    V[(size_t) r * n + p] = c * vrp - s * vrq;
                    V[(size_t) r * n + q] = s * vrp + c * vrq;
                }
            }
    }
    int *idx = (int *) malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) { w[i] = A[(size_t) i * n + i]; idx[i] = i; }
    for (int i = 1; i < n; i++) {       /* insertion sort, ascending */
        int j = i; int t = idx[i];
        while (j > 0 && w[idx[j - 1]] > w[t]) { idx[j] = idx[j - 1]; j--; }
        idx[j] = t;
    }
    double *w2 = (double *) malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) w2[j] = w[idx[j]];
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) A[(size_t) i * n + j] = V[(size_t) i * n + idx[j]];
    memcpy(w, w2, sizeof(double) * n);
    free(w2); free(idx); free(V);
    return 0;
}

/* unblocked pivoted Cholesky, lower, row-major: P^T A P = L L^T (LAPACK dpstf2 semantics,
 * piv 1-based, first maximal diagonal chosen, stop when pivot <= tol) */
static int pstrf_fallback(int n, double *A, int *piv, int *rank, double tol)
{
    double *work = (double *) calloc(n, sizeof(double));
    for (int i = 0; i < n; i++) piv[i] = i + 1;
    int r = n;
    for (int j = 0; j < n; j++) {
        int pvt = j; double amax = -1e300;
        for (int i = j; i < n; i++) {
            if (j > 0) work[i] += A[(size_t) i * n + j - 1] * A[(size_t) i * n + j - 1];
            double v = A[(size_t) i * n + i] - work[i];
            if (v > amax) { amax = v; pvt = i; }
        }
        if (amax <= tol || isnan(amax)) { A[(size_t) j * n + j] = amax; r = j; break; }
        if (pvt != j) {
            A[(size_t) pvt * n + pvt] = A[(size_t) j * n + j];
            for (int c = 0; c < j; c++) { double t = A[(size_t) j * n + c]; A[(size_t) j * n + c] = A[(size_t) pvt * n + c]; A[(size_t) pvt * n + c] = t; }
            for (int c = pvt + 1; c < n; c++) { double t = A[(size_t) c * n + j]; A[(size_t) c * n + j] = A[(size_t) c * n + pvt]; A[(size_t) c * n + pvt] = t; }
            for (int c = j + 1; c < pvt; c++) { double t = A[(size_t) c * n + j]; A[(size_t) c * n + j] = A[(size_t) pvt * n + c]; A[(size_t) pvt * n + c] = t; }
            double t = work[j]; work[j] = work[pvt]; work[pvt] = t;
            int ip = piv[j]; piv[j] = piv[pvt]; piv[pvt] = ip;
        }
        double ajj = sqrt(amax);
        A[(size_t) j * n + j] = ajj;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t) i * n + j];
            for (int c = 0; c < j; c++) s -= A[(size_t) i * n + c] * A[(size_t) j * n + c];
            A[(size_t) i * n + j] = s / ajj;
        }
    }
    *rank = r;
    free(work);
    return r < n ? 1 : 0;
}

static int gesv_fallback(int n, double *A, int lda, int *ipiv, double *b)
{
    for (int k = 0; k < n; k++) {
        int p = k; double m = fabs(A[(size_t) k * lda + k]);
        for (int i = k + 1; i < n; i++) if (fabs(A[(size_t) i * lda + k]) > m) { m = fabs(A[(size_t) i * lda + k]); p = i; }
        ipiv[k] = p + 1;
        if (m == 0.0) return k + 1;
        if (p != k) {
            for (int c = 0; c < n; c++) { double t = A[(size_t) k * lda + c]; A[(size_t) k * lda + c] = A[(size_t) p * lda + c]; A[(size_t) p * lda + c] = t; }
            double t = b[k]; b[k] = b[p]; b[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            double f = A[(size_t) i * lda + k] / A[(size_t) k * lda + k];
            A[(size_t) i * lda + k] = f;
            for (int c = k + 1; c < n; c++) A[(size_t) i * lda + c] -= f * A[(size_t) k * lda + c];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int c = i + 1; c < n; c++) s -= A[(size_t) i * lda + c] * b[c];
        b[i] = s / A[(size_t) i * lda + i];
    }
    return 0;
}

/* ---------------- dispatch ---------------- */
void ora_dgemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc)
{
    if (m <= 0 || n <= 0) return;
    if (ora_lapack_init() == 1)
        p_dgemm(101, ta ? 112 : 111, tb ? 112 : 111, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        gemm_fallback(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int ora_dsyev(int n, double *A, double *w)
{
    if (ora_lapack_init() == 1) return p_dsyev(101, 'V', 'U', n, A, n, w);
    return syev_fallback(n, A, w);
}

int ora_dpstrf_lower(int n, double *A, int *piv, int *rank, double tol)
{
    if (ora_lapack_init() == 1) return p_dpstrf(101, 'L', n, A, n, piv, rank, tol);
    return pstrf_fallback(n, A, piv, rank, tol);
}

int ora_dgesv(int n, double *A, int lda, int *ipiv, double *b)
{
    if (ora_lapack_init() == 1) return p_dgesv(101, n, 1, A, lda, ipiv, b, 1);
    return gesv_fallback(n, A, lda, ipiv, b);
}
