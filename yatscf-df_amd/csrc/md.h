// md.h -- McMurchie-Davidson integral machinery for gfx950 (device side).
//
// Hermite expansion coefficients E^{ij}_t, Hermite Coulomb integrals
// R_{tuv}, and the Boys function F_n(T) (8-term Taylor expansion from a
// table with spacing 0.05 on [0,40) plus downward recursion; asymptotic form
// and upward recursion above).  The table is built on the host
// (ygpu_api.cpp) from the series e^{-T} sum_k (2T)^k / ((2n+1)...(2n+2k+1)).
//
// Cartesian component order inside a shell: descending x power, then
// descending y power (xx, xy, xz, yy, yz, zz), the order the reference's
// integral provider is assumed to use (CMS_Simint.c:363,527,595).
#pragma once
#include <hip/hip_runtime.h>

#define YG_BOYS_NMAX 24
#define YG_BOYS_TMAX 40.0
#define YG_BOYS_H 0.05
#define YG_BOYS_NPTS 802
#define YG_PI 3.14159265358979323846
#define YG_TWO_PI_25 34.98683665524972   /* 2 pi^{5/2} */

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

__host__ __device__ constexpr int cart_pow(int l, int idx, int dim)
{
    int c = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (c == idx) return dim == 0 ? i : (dim == 1 ? j : l - i - j);
            ++c;
        }
    return 0;
}

// Shell description on the device (structure of arrays is uploaded once).
struct DevShells {
    const int *am, *nprim, *prim_off, *bf_start;
    const double *xyz, *alpha, *coef;
};

// ---------------------------------------------------------------------------
// Boys function, compile-time order
template <int L>
__device__ __forceinline__ void boys_fn(double T, const double *__restrict__ tab, double *F)
{
    if (T < YG_BOYS_TMAX) {
        int k = (int)(T * (1.0 / YG_BOYS_H) + 0.5);
        double d = k * YG_BOYS_H - T;
        const double *f = tab + k * (YG_BOYS_NMAX + 1) + L;
        double s = f[7];
        s = f[6] + s * d * (1.0 / 7.0);
        s = f[5] + s * d * (1.0 / 6.0);
        s = f[4] + s * d * (1.0 / 5.0);
        s = f[3] + s * d * (1.0 / 4.0);
        s = f[2] + s * d * (1.0 / 3.0);
        s = f[1] + s * d * (1.0 / 2.0);
        s = f[0] + s * d;
        F[L] = s;
        if (L > 0) {
            const double eT = exp(-T), t2 = 2.0 * T;
#pragma unroll
            for (int n = L - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + eT) / (double)(2 * n + 1);
        }
    } else {
        const double eT = exp(-T), o2t = 0.5 / T;
        F[0] = 0.5 * sqrt(YG_PI / T);
#pragma unroll
        for (int n = 0; n < L; ++n) F[n + 1] = ((double)(2 * n + 1) * F[n] - eT) * o2t;
    }
}

// Boys function, run-time order (one-time integrals)
__device__ __forceinline__ void boys_rt(int L, double T, const double *__restrict__ tab, double *F)
{
    if (T < YG_BOYS_TMAX) {
        int k = (int)(T * (1.0 / YG_BOYS_H) + 0.5);
        double d = k * YG_BOYS_H - T;
        const double *f = tab + k * (YG_BOYS_NMAX + 1) + L;
        double s = f[7];
        for (int j = 6; j >= 0; --j) s = f[j] + s * d / (double)(j + 1);
        F[L] = s;
        const double eT = exp(-T), t2 = 2.0 * T;
        for (int n = L - 1; n >= 0; --n) F[n] = (t2 * F[n + 1] + eT) / (double)(2 * n + 1);
    } else {
        const double eT = exp(-T), o2t = 0.5 / T;
        F[0] = 0.5 * sqrt(YG_PI / T);
        for (int n = 0; n < L; ++n) F[n + 1] = ((double)(2 * n + 1) * F[n] - eT) * o2t;
    }
}

// ---------------------------------------------------------------------------
// 1D Hermite coefficients, compile-time angular momenta.
// E[(i*(LB+1)+j)*(LA+LB+1)+t] = E^{ij}_t (includes exp(-mu X_AB^2)).
template <int LA, int LB>
__device__ __forceinline__ void herm_E(double a, double b, double XAB, double *E)
{
    constexpr int NT = LA + LB + 1;
    const double p = a + b, oo2p = 0.5 / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
#define EI(i, j, t) E[((i) * (LB + 1) + (j)) * NT + (t)]
#pragma unroll
    for (int q = 0; q < (LA + 1) * (LB + 1) * NT; ++q) E[q] = 0.0;
    EI(0, 0, 0) = exp(-a * b / p * XAB * XAB);
#pragma unroll
    for (int i = 0; i <= LA; ++i) {
        if (i > 0) {
#pragma unroll
            for (int t = 0; t <= i; ++t) {
                double v = 0.0;
                if (t <= i - 1) v = XPA * EI(i - 1, 0, t);
                if (t > 0) v += oo2p * EI(i - 1, 0, t - 1);
                if (t + 1 <= i - 1) v += (double)(t + 1) * EI(i - 1, 0, t + 1);
                EI(i, 0, t) = v;
            }
        }
#pragma unroll
        for (int j = 1; j <= LB; ++j) {
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v = 0.0;
                if (t <= i + j - 1) v = XPB * EI(i, j - 1, t);
                if (t > 0) v += oo2p * EI(i, j - 1, t - 1);
                if (t + 1 <= i + j - 1) v += (double)(t + 1) * EI(i, j - 1, t + 1);
                EI(i, j, t) = v;
            }
        }
    }
#undef EI
}

// run-time version: E sized (la+1)*(lb+1)*(la+lb+1) <= 5*5*9 for la,lb <= 4
__device__ __forceinline__ void herm_E_rt(int la, int lb, double a, double b, double XAB, double *E)
{
    const int NT = la + lb + 1;
    const double p = a + b, oo2p = 0.5 / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
#define EI(i, j, t) E[((i) * (lb + 1) + (j)) * NT + (t)]
    for (int q = 0; q < (la + 1) * (lb + 1) * NT; ++q) E[q] = 0.0;
    EI(0, 0, 0) = exp(-a * b / p * XAB * XAB);
    for (int i = 0; i <= la; ++i) {
        if (i > 0)
            for (int t = 0; t <= i; ++t) {
                double v = 0.0;
                if (t <= i - 1) v = XPA * EI(i - 1, 0, t);
                if (t > 0) v += oo2p * EI(i - 1, 0, t - 1);
                if (t + 1 <= i - 1) v += (double)(t + 1) * EI(i - 1, 0, t + 1);
                EI(i, 0, t) = v;
            }
        for (int j = 1; j <= lb; ++j)
            for (int t = 0; t <= i + j; ++t) {
                double v = 0.0;
                if (t <= i + j - 1) v = XPB * EI(i, j - 1, t);
                if (t > 0) v += oo2p * EI(i, j - 1, t - 1);
                if (t + 1 <= i + j - 1) v += (double)(t + 1) * EI(i, j - 1, t + 1);
                EI(i, j, t) = v;
            }
    }
#undef EI
}

// single-centre Hermite expansion of x^l exp(-g x^2): coefficient of Lambda_tau
template <int L>
__device__ __forceinline__ void herm_E1(double g, double *E /* (L+1)*(L+1) */)
{
#pragma unroll
    for (int q = 0; q < (L + 1) * (L + 1); ++q) E[q] = 0.0;
    E[0] = 1.0;
    const double oo2g = 0.5 / g;
#pragma unroll
    for (int i = 1; i <= L; ++i)
#pragma unroll
        for (int t = 0; t <= i; ++t) {
            double v = 0.0;
            if (t > 0) v += oo2g * E[(i - 1) * (L + 1) + t - 1];
            if (t + 1 <= i - 1) v += (double)(t + 1) * E[(i - 1) * (L + 1) + t + 1];
            E[i * (L + 1) + t] = v;
        }
}

// ---------------------------------------------------------------------------
// Hermite Coulomb integrals, compile-time L.  R[(t*(L+1)+u)*(L+1)+v].
template <int L>
__device__ __forceinline__ void herm_R(double alpha, double X, double Y, double Z,
                                       const double *__restrict__ tab, double *R)
{
    constexpr int D = L + 1;
    double F[L + 1];
    boys_fn<L>(alpha * (X * X + Y * Y + Z * Z), tab, F);
    double Rn[L + 1][D][D][D];
    double f = 1.0;
    const double m2a = -2.0 * alpha;
#pragma unroll
    for (int n = 0; n <= L; ++n) { Rn[n][0][0][0] = f * F[n]; f *= m2a; }
#pragma unroll
    for (int N = 1; N <= L; ++N)
#pragma unroll
        for (int t = 0; t <= N; ++t)
#pragma unroll
            for (int u = 0; u <= N - t; ++u) {
                const int v = N - t - u;
#pragma unroll
                for (int n = 0; n <= L - N; ++n) {
                    double r;
                    if (t > 0) {
                        r = X * Rn[n + 1][t - 1][u][v];
                        if (t > 1) r += (double)(t - 1) * Rn[n + 1][t > 1 ? t - 2 : 0][u][v];
                    } else if (u > 0) {
                        r = Y * Rn[n + 1][t][u - 1][v];
                        if (u > 1) r += (double)(u - 1) * Rn[n + 1][t][u > 1 ? u - 2 : 0][v];
                    } else {
                        r = Z * Rn[n + 1][t][u][v - 1];
                        if (v > 1) r += (double)(v - 1) * Rn[n + 1][t][u][v > 1 ? v - 2 : 0];
                    }
                    Rn[n][t][u][v] = r;
                }
            }
#pragma unroll
    for (int t = 0; t <= L; ++t)
#pragma unroll
        for (int u = 0; u <= L - t; ++u)
#pragma unroll
            for (int v = 0; v <= L - t - u; ++v) R[(t * D + u) * D + v] = Rn[0][t][u][v];
}

// run-time L (<= 8): R and scratch supplied by the caller
#define YG_RT_LMAX 8
__device__ __forceinline__ void herm_R_rt(int L, double alpha, double X, double Y, double Z,
                                          const double *__restrict__ tab, double *R /* (L+1)^3 */)
{
    const int D = L + 1;
    double F[YG_RT_LMAX + 1];
    boys_rt(L, alpha * (X * X + Y * Y + Z * Z), tab, F);
    // two layers of the n recursion suffice when N increases: keep full table by n
    double Rn[(YG_RT_LMAX + 1) * (YG_RT_LMAX + 1) * (YG_RT_LMAX + 1) * (YG_RT_LMAX + 1) / 4 + 1];
    // layout: for each n, a packed (t,u,v) with t+u+v <= L-n; index via offsets
    // simple (L+1)^3 block per n would exceed the budget for L=8, so use the packed index
    auto pidx = [](int t, int u, int v) {
        const int N = t + u + v;
        // packed index of (t,u,v) among all triples with total < N, then within total N
        const int before = N * (N + 1) * (N + 2) / 6;
        const int tt = N - t;                  // number of (u,v) with u+v <= tt ordering
        const int within = tt * (tt + 1) / 2 + (tt - u);
        (void) v;
        return before + within;
    };
    const int npk = (L + 1) * (L + 2) * (L + 3) / 6;   // triples with total <= L
    double f = 1.0;
    const double m2a = -2.0 * alpha;
    for (int n = 0; n <= L; ++n) { Rn[n * npk + pidx(0, 0, 0)] = f * F[n]; f *= m2a; }
    for (int N = 1; N <= L; ++N)
        for (int t = 0; t <= N; ++t)
            for (int u = 0; u <= N - t; ++u) {
                const int v = N - t - u;
                for (int n = 0; n <= L - N; ++n) {
                    double r;
                    const double *Rp = Rn + (n + 1) * npk;
                    if (t > 0) {
                        r = X * Rp[pidx(t - 1, u, v)];
                        if (t > 1) r += (t - 1) * Rp[pidx(t - 2, u, v)];
                    } else if (u > 0) {
                        r = Y * Rp[pidx(t, u - 1, v)];
                        if (u > 1) r += (u - 1) * Rp[pidx(t, u - 2, v)];
                    } else {
                        r = Z * Rp[pidx(t, u, v - 1)];
                        if (v > 1) r += (v - 1) * Rp[pidx(t, u, v - 2)];
                    }
                    Rn[n * npk + pidx(t, u, v)] = r;
                }
            }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[(t * D + u) * D + v] = Rn[pidx(t, u, v)];
}

// ---------------------------------------------------------------------------
// 1D Hermite coefficients E^{ij}_t for one run-time pair (i, j) <= (LA, LB),
// built along the path (0,0) -> (i,0) -> (i,j).  Output E[t], t <= LA+LB (zero
// beyond i+j).  All indexing is compile-time (no dynamically indexed arrays).
template <int LA, int LB>
__device__ __forceinline__ void herm_E1d(double a, double b, double XAB, int i, int j, double *E)
{
    constexpr int NT = LA + LB + 1;
    const double p = a + b, oo2p = 0.5 / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
#pragma unroll
    for (int t = 0; t < NT; ++t) E[t] = 0.0;
    E[0] = exp(-a * b / p * XAB * XAB);
#pragma unroll
    for (int s = 1; s <= LA + LB; ++s) {
        const bool onA = s <= LA;                 // first LA steps raise i, then LB steps raise j
        const bool active = onA ? (s <= i) : (s - LA <= j);
        if (active) {
            const double X = onA ? XPA : XPB;
            double nx[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double v = X * E[t];
                if (t > 0) v += oo2p * E[t - 1];
                if (t + 1 < NT) v += (double)(t + 1) * E[t + 1];
                nx[t] = v;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) E[t] = nx[t];
        }
    }
}

// E^{i,j}_0 for j = 0..LB along (0,0) -> (i,0) -> (i,LB) (overlap factors)
template <int LA, int LB>
__device__ __forceinline__ void herm_E0_row(double a, double b, double XAB, int i, double *E0)
{
    constexpr int NT = LA + LB + 1;
    const double p = a + b, oo2p = 0.5 / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
    double E[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) E[t] = 0.0;
    E[0] = exp(-a * b / p * XAB * XAB);
#pragma unroll
    for (int s = 1; s <= LA; ++s) {
        if (s <= i) {
            double nx[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double v = XPA * E[t];
                if (t > 0) v += oo2p * E[t - 1];
                if (t + 1 < NT) v += (double)(t + 1) * E[t + 1];
                nx[t] = v;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) E[t] = nx[t];
        }
    }
    E0[0] = E[0];
#pragma unroll
    for (int j = 1; j <= LB; ++j) {
        double nx[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            double v = XPB * E[t];
            if (t > 0) v += oo2p * E[t - 1];
            if (t + 1 < NT) v += (double)(t + 1) * E[t + 1];
            nx[t] = v;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) E[t] = nx[t];
        E0[j] = E[0];
    }
}

// ---------------------------------------------------------------------------
// shell access helpers
struct ShellRef {
    int am, nprim, off, bf;
    double x, y, z;
};

__device__ __forceinline__ ShellRef shell_ref(const DevShells &s, int i)
{
    ShellRef r;
    r.am = s.am[i]; r.nprim = s.nprim[i]; r.off = s.prim_off[i]; r.bf = s.bf_start[i];
    r.x = s.xyz[3 * i]; r.y = s.xyz[3 * i + 1]; r.z = s.xyz[3 * i + 2];
    return r;
}

__device__ __forceinline__ void cart_pow_rt(int l, int idx, int &lx, int &ly, int &lz)
{
    int c = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (c == idx) { lx = i; ly = j; lz = l - i - j; return; }
            ++c;
        }
    lx = ly = lz = 0;
}

