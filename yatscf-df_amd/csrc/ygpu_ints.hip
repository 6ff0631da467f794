// ygpu_ints.hip -- one-electron, Schwarz, 2-centre and 3-centre integrals on gfx950.
//
// Replaces the reference's Simint calls:
//   overlap / kinetic / potential  (CMS_Simint.c:402-461, used by TinySCF.c:20-54)
//   (MN|MN) Schwarz elements       (CMS_Simint.c:337-399, used by TinySCF.c:112-177)
//   (P|Q) metric                   (CMS_Simint.c:569-630, build_DF_tensor.c:152-203)
//   (MN|P) 3-centre                (CMS_Simint.c:498-566, build_DF_tensor.c:49-150)
// by the McMurchie-Davidson scheme (md.h).  The 3-centre kernel is the hot
// one: it is instantiated per angular-momentum class (la >= lb <= 2, lc <= 3)
// and per Cartesian component of P so that every Hermite index is a
// compile-time constant (the R recursion stays in registers); one wavefront
// handles one screened shell pair against 64 aux shells of one class, with
// the bra primitive-pair data (P centre, exponent, Hermite coefficients)
// staged once per workgroup in LDS.
#include "ygpu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

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

// ---------------------------------------------------------------------------
// S and Hcore = T + V, one workgroup per shell pair (M,N), one thread per element
__global__ __launch_bounds__(64) void k_ovlp_hcore(DevShells sh, int nshells, int nbf, int natoms,
                                                   const double *__restrict__ Zc, const double *__restrict__ axyz,
                                                   const double *__restrict__ tab, double *__restrict__ S,
                                                   double *__restrict__ H)
{
    const int M = blockIdx.x / nshells, N = blockIdx.x % nshells;
    const ShellRef A = shell_ref(sh, M), B = shell_ref(sh, N);
    const int na = (A.am + 1) * (A.am + 2) / 2, nb = (B.am + 1) * (B.am + 2) / 2;
    const int e = threadIdx.x;
    if (e >= na * nb) return;
    const int ia = e / nb, ib = e % nb;
    int ax, ay, az, bx, by, bz;
    cart_pow_rt(A.am, ia, ax, ay, az);
    cart_pow_rt(B.am, ib, bx, by, bz);
    const int la = A.am, lb2 = B.am + 2, NT = la + lb2 + 1, L = A.am + B.am, D = L + 1;
    double Ex[3 * 5 * 8], Ey[3 * 5 * 8], Ez[3 * 5 * 8], R[5 * 5 * 5];
    double s_acc = 0.0, t_acc = 0.0, v_acc = 0.0;
    for (int pa = 0; pa < A.nprim; ++pa)
        for (int pb = 0; pb < B.nprim; ++pb) {
            const double a = sh.alpha[A.off + pa], b = sh.alpha[B.off + pb], p = a + b;
            const double cab = sh.coef[A.off + pa] * sh.coef[B.off + pb];
            herm_E_rt(la, lb2, a, b, A.x - B.x, Ex);
            herm_E_rt(la, lb2, a, b, A.y - B.y, Ey);
            herm_E_rt(la, lb2, a, b, A.z - B.z, Ez);
            const double sq = sqrt(YG_PI / p);
            auto S1 = [&](const double *E, int i, int j) -> double {
                return j < 0 ? 0.0 : E[(i * (lb2 + 1) + j) * NT] * sq;
            };
            auto T1 = [&](const double *E, int i, int j) -> double {
                return -2.0 * b * b * S1(E, i, j + 2) + b * (2 * j + 1) * S1(E, i, j) - 0.5 * j * (j - 1) * S1(E, i, j - 2);
            };
            const double sx = S1(Ex, ax, bx), sy = S1(Ey, ay, by), sz = S1(Ez, az, bz);
            s_acc += cab * sx * sy * sz;
            t_acc += cab * (T1(Ex, ax, bx) * sy * sz + sx * T1(Ey, ay, by) * sz + sx * sy * T1(Ez, az, bz));
            const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p, Pz = (a * A.z + b * B.z) / p;
            const double cv = cab * 2.0 * YG_PI / p;
            for (int c = 0; c < natoms; ++c) {
                herm_R_rt(L, p, Px - axyz[3 * c], Py - axyz[3 * c + 1], Pz - axyz[3 * c + 2], tab, R);
                double s = 0.0;
                for (int t = 0; t <= ax + bx; ++t)
                    for (int u = 0; u <= ay + by; ++u)
                        for (int v = 0; v <= az + bz; ++v)
                            s += Ex[(ax * (lb2 + 1) + bx) * NT + t] * Ey[(ay * (lb2 + 1) + by) * NT + u] *
                                 Ez[(az * (lb2 + 1) + bz) * NT + v] * R[(t * D + u) * D + v];
                v_acc -= Zc[c] * cv * s;
            }
        }
    const size_t o = (size_t)(A.bf + ia) * nbf + B.bf + ib;
    S[o] = s_acc;
    H[o] = t_acc + v_acc;
}

// ---------------------------------------------------------------------------
// One element (a b | c d) of a 4-centre shell quartet (run-time angular momenta)
__device__ double eri4_elem(const DevShells &sh, const ShellRef &A, const ShellRef &B, const ShellRef &C,
                            const ShellRef &Dd, int ia, int ib, int ic, int id, const double *__restrict__ tab)
{
    int ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
    cart_pow_rt(A.am, ia, ax, ay, az);
    cart_pow_rt(B.am, ib, bx, by, bz);
    cart_pow_rt(C.am, ic, cx, cy, cz);
    cart_pow_rt(Dd.am, id, dx, dy, dz);
    const int L = A.am + B.am + C.am + Dd.am, D = L + 1;
    const int nt1 = A.am + B.am + 1, nt2 = C.am + Dd.am + 1;
    double E1x[3 * 3 * 5], E1y[3 * 3 * 5], E1z[3 * 3 * 5], E2x[3 * 3 * 5], E2y[3 * 3 * 5], E2z[3 * 3 * 5];
    double R[9 * 9 * 9];
    double res = 0.0;
    for (int pa = 0; pa < A.nprim; ++pa)
        for (int pb = 0; pb < B.nprim; ++pb) {
            const double a = sh.alpha[A.off + pa], b = sh.alpha[B.off + pb], p = a + b;
            const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p, Pz = (a * A.z + b * B.z) / p;
            herm_E_rt(A.am, B.am, a, b, A.x - B.x, E1x);
            herm_E_rt(A.am, B.am, a, b, A.y - B.y, E1y);
            herm_E_rt(A.am, B.am, a, b, A.z - B.z, E1z);
            const double cab = sh.coef[A.off + pa] * sh.coef[B.off + pb];
            for (int pc = 0; pc < C.nprim; ++pc)
                for (int pd = 0; pd < Dd.nprim; ++pd) {
                    const double c = sh.alpha[C.off + pc], d = sh.alpha[Dd.off + pd], q = c + d;
                    const double Qx = (c * C.x + d * Dd.x) / q, Qy = (c * C.y + d * Dd.y) / q, Qz = (c * C.z + d * Dd.z) / q;
                    herm_E_rt(C.am, Dd.am, c, d, C.x - Dd.x, E2x);
                    herm_E_rt(C.am, Dd.am, c, d, C.y - Dd.y, E2y);
                    herm_E_rt(C.am, Dd.am, c, d, C.z - Dd.z, E2z);
                    const double alpha = p * q / (p + q);
                    herm_R_rt(L, alpha, Px - Qx, Py - Qy, Pz - Qz, tab, R);
                    const double *e1x = &E1x[(ax * (B.am + 1) + bx) * nt1], *e1y = &E1y[(ay * (B.am + 1) + by) * nt1];
                    const double *e1z = &E1z[(az * (B.am + 1) + bz) * nt1];
                    const double *e2x = &E2x[(cx * (Dd.am + 1) + dx) * nt2], *e2y = &E2y[(cy * (Dd.am + 1) + dy) * nt2];
                    const double *e2z = &E2z[(cz * (Dd.am + 1) + dz) * nt2];
                    double s = 0.0;
                    for (int t = 0; t <= ax + bx; ++t)
                        for (int u = 0; u <= ay + by; ++u)
                            for (int v = 0; v <= az + bz; ++v) {
                                const double e1 = e1x[t] * e1y[u] * e1z[v];
                                for (int tau = 0; tau <= cx + dx; ++tau)
                                    for (int nu = 0; nu <= cy + dy; ++nu)
                                        for (int phi = 0; phi <= cz + dz; ++phi) {
                                            const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                            s += sg * e1 * e2x[tau] * e2y[nu] * e2z[phi] *
                                                 R[((t + tau) * D + u + nu) * D + v + phi];
                                        }
                            }
                    res += cab * sh.coef[C.off + pc] * sh.coef[Dd.off + pd] * YG_TWO_PI_25 / (p * q * sqrt(p + q)) * s;
                }
        }
    return res;
}

// Schwarz values: one workgroup per ordered shell pair (M,N), one thread per (iM,iN)
__global__ __launch_bounds__(64) void k_schwarz(DevShells sh, int nshells, int nbf, int quirk,
                                                const double *__restrict__ tab, double *__restrict__ out)
{
    const int M = blockIdx.x / nshells, N = blockIdx.x % nshells;
    const ShellRef A = shell_ref(sh, M), B = shell_ref(sh, N);
    const int dM = (A.am + 1) * (A.am + 2) / 2, dN = (B.am + 1) * (B.am + 2) / 2;
    const int e = threadIdx.x;
    if (e >= dM * dN) return;
    const int iM = e / dN, iN = e % dN;
    int a, b, c, d;
    if (quirk) {
        // the reference reads index iN*(dM*dN*dM+dM) + iM*(dN*dM+1) of the row-major
        // (MN|MN) block [a][b][c][d] (TinySCF.c:153)
        int idx = iN * (dM * dN * dM + dM) + iM * (dN * dM + 1);
        d = idx % dN; idx /= dN;
        c = idx % dM; idx /= dM;
        b = idx % dN; idx /= dN;
        a = idx;
    } else {
        a = iM; b = iN; c = iM; d = iN;
    }
    out[(size_t)(A.bf + iM) * nbf + B.bf + iN] = fabs(eri4_elem(sh, A, B, A, B, a, b, c, d, tab));
}

// ---------------------------------------------------------------------------
// (P|Q) element between two single-centre shells
__device__ double eri2_elem(const DevShells &sh, const ShellRef &C, const ShellRef &Dd, int ic, int id,
                            const double *__restrict__ tab)
{
    int cx, cy, cz, dx, dy, dz;
    cart_pow_rt(C.am, ic, cx, cy, cz);
    cart_pow_rt(Dd.am, id, dx, dy, dz);
    const int L = C.am + Dd.am, D = L + 1;
    double Ec[3 * 3 * 7], Ed[3 * 3 * 7], R[7 * 7 * 7];
    double res = 0.0;
    for (int pc = 0; pc < C.nprim; ++pc)
        for (int pd = 0; pd < Dd.nprim; ++pd) {
            const double g = sh.alpha[C.off + pc], h = sh.alpha[Dd.off + pd];
            // a single Gaussian is the pair (g, 0) on one centre
            herm_E_rt(C.am, 0, g, 0.0, 0.0, Ec);
            herm_E_rt(Dd.am, 0, h, 0.0, 0.0, Ed);
            herm_R_rt(L, g * h / (g + h), C.x - Dd.x, C.y - Dd.y, C.z - Dd.z, tab, R);
            const int ntc = C.am + 1, ntd = Dd.am + 1;
            double s = 0.0;
            for (int t = 0; t <= cx; ++t)
                for (int u = 0; u <= cy; ++u)
                    for (int v = 0; v <= cz; ++v) {
                        const double e1 = Ec[cx * ntc + t] * Ec[cy * ntc + u] * Ec[cz * ntc + v];
                        if (e1 == 0.0) continue;
                        for (int tau = 0; tau <= dx; ++tau)
                            for (int nu = 0; nu <= dy; ++nu)
                                for (int phi = 0; phi <= dz; ++phi) {
                                    const double e2 = Ed[dx * ntd + tau] * Ed[dy * ntd + nu] * Ed[dz * ntd + phi];
                                    const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                    s += sg * e1 * e2 * R[((t + tau) * D + u + nu) * D + v + phi];
                                }
                    }
            res += sh.coef[C.off + pc] * sh.coef[Dd.off + pd] * YG_TWO_PI_25 / (g * h * sqrt(g + h)) * s;
        }
    return res;
}

// max_k (P_k|P_k) for every aux shell
__global__ void k_df_scr(DevShells sh, int nshells, const double *__restrict__ tab, double *__restrict__ out)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= nshells) return;
    const ShellRef A = shell_ref(sh, P);
    const int n = (A.am + 1) * (A.am + 2) / 2;
    double m = 0.0;
    for (int k = 0; k < n; ++k) m = fmax(m, fabs(eri2_elem(sh, A, A, k, k, tab)));
    out[P] = m;
}

// J_PQ for the screened aux shell pairs P <= Q (both triangles written)
__global__ __launch_bounds__(128) void k_eri2(DevShells sh, const int *__restrict__ plist, const int *__restrict__ qlist,
                                              const double *__restrict__ tab, double *__restrict__ J, int ld)
{
    const int P = plist[blockIdx.x], Q = qlist[blockIdx.x];
    const ShellRef A = shell_ref(sh, P), B = shell_ref(sh, Q);
    const int na = (A.am + 1) * (A.am + 2) / 2, nb = (B.am + 1) * (B.am + 2) / 2;
    const int e = threadIdx.x;
    if (e >= na * nb) return;
    const int ia = e / nb, ib = e % nb;
    const double v = eri2_elem(sh, A, B, ia, ib, tab);
    J[(size_t)(A.bf + ia) * ld + B.bf + ib] = v;
    J[(size_t)(B.bf + ib) * ld + A.bf + ia] = v;
}

// ---------------------------------------------------------------------------
// 3-centre (MN|P) into the packed pair rows
struct Eri3Args {
    DevShells bas, aux;
    const int *pair_lid, *pair_rid;    // shell pairs of this (la,lb) class
    const int *plist;                   // aux shells of this lc class
    int npairs, nP, nshells, nbf;
    const int *upmask;                  // nbf*nbf -> packed row or -1
    const double *sp_scr, *df_scr;
    double scrtol2;
    const double *tab;
    double *out;                        // packed rows, leading dimension ld
    long ld;
    int npp_max;
};

template <int LA, int LB, int LC, int KK>
__device__ __forceinline__ void eri3_body(const Eri3Args &g, const double *__restrict__ lds, int npp,
                                          const ShellRef &A, const ShellRef &B, int P)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), LAB = LA + LB, L = LAB + LC, D = L + 1;
    constexpr int NE1 = (LA + 1) * (LB + 1) * (LAB + 1);
    constexpr int STRIDE = 5 + 3 * NE1;
    constexpr int cx = cart_pow(LC, KK, 0), cy = cart_pow(LC, KK, 1), cz = cart_pow(LC, KK, 2);
    constexpr int DW = LAB + 1;
    const ShellRef C = shell_ref(g.aux, P);
    double acc[NA * NB];
#pragma unroll
    for (int q = 0; q < NA * NB; ++q) acc[q] = 0.0;
    for (int kp = 0; kp < C.nprim; ++kp) {
        const double gam = g.aux.alpha[C.off + kp], cc = g.aux.coef[C.off + kp];
        double Ec[(LC + 1) * (LC + 1)];
        herm_E1<LC>(gam, Ec);
        for (int pp = 0; pp < npp; ++pp) {
            const double *rec = lds + pp * STRIDE;
            const double p = rec[0];
            const double alpha = p * gam / (p + gam);
            double R[D * D * D];
            herm_R<L>(alpha, rec[1] - C.x, rec[2] - C.y, rec[3] - C.z, g.tab, R);
            const double pref = rec[4] * cc * YG_TWO_PI_25 / (p * gam * sqrt(p + gam));
            // W(t,u,v) = sum (-1)^{tau+nu+phi} Ec Ec Ec R[t+tau][u+nu][v+phi]
            double W[DW * DW * DW];
#pragma unroll
            for (int t = 0; t <= LAB; ++t)
#pragma unroll
                for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
                    for (int v = 0; v <= LAB - t - u; ++v) {
                        double s = 0.0;
#pragma unroll
                        for (int tau = 0; tau <= cx; ++tau)
#pragma unroll
                            for (int nu = 0; nu <= cy; ++nu)
#pragma unroll
                                for (int phi = 0; phi <= cz; ++phi) {
                                    constexpr int dummy = 0;
                                    (void) dummy;
                                    const double e = Ec[cx * (LC + 1) + tau] * Ec[cy * (LC + 1) + nu] * Ec[cz * (LC + 1) + phi];
                                    const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                    s += sg * e * R[((t + tau) * D + u + nu) * D + v + phi];
                                }
                        W[(t * DW + u) * DW + v] = s * pref;
                    }
            const double *Ex = rec + 5, *Ey = rec + 5 + NE1, *Ez = rec + 5 + 2 * NE1;
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib) {
                    constexpr int dummy2 = 0;
                    (void) dummy2;
                    const int ax = cart_pow(LA, ia, 0), ay = cart_pow(LA, ia, 1), az = cart_pow(LA, ia, 2);
                    const int bx = cart_pow(LB, ib, 0), by = cart_pow(LB, ib, 1), bz = cart_pow(LB, ib, 2);
                    double s = 0.0;
#pragma unroll
                    for (int t = 0; t <= ax + bx; ++t)
#pragma unroll
                        for (int u = 0; u <= ay + by; ++u) {
                            const double exy = Ex[(ax * (LB + 1) + bx) * (LAB + 1) + t] * Ey[(ay * (LB + 1) + by) * (LAB + 1) + u];
#pragma unroll
                            for (int v = 0; v <= az + bz; ++v)
                                s += exy * Ez[(az * (LB + 1) + bz) * (LAB + 1) + v] * W[(t * DW + u) * DW + v];
                        }
                    acc[ia * NB + ib] += s;
                }
        }
    }
    const int col = C.bf + KK;
    const bool same = (A.bf == B.bf);
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const int m = A.bf + ia, n = B.bf + ib;
            if (same && ia > ib) continue;          // (m,n) and (n,m) share one packed row
            const int lo = m < n ? m : n, hi = m < n ? n : m;
            const int r = g.upmask[(size_t) lo * g.nbf + hi];
            if (r >= 0) g.out[(size_t) r * g.ld + col] = acc[ia * NB + ib];
        }
}

template <int LA, int LB, int LC, int KK>
__device__ __forceinline__ void eri3_dispatch(const Eri3Args &g, const double *lds, int npp, const ShellRef &A,
                                              const ShellRef &B, int P, int kk)
{
    if constexpr (KK < ncart(LC)) {
        if (kk == KK) eri3_body<LA, LB, LC, KK>(g, lds, npp, A, B, P);
        else eri3_dispatch<LA, LB, LC, KK + 1>(g, lds, npp, A, B, P, kk);
    }
}

template <int LA, int LB, int LC>
__global__ __launch_bounds__(64) void k_eri3(Eri3Args g)
{
    extern __shared__ double lds[];
    constexpr int NE1 = (LA + 1) * (LB + 1) * (LA + LB + 1);
    constexpr int STRIDE = 5 + 3 * NE1;
    const int pair = blockIdx.x;
    const int M = g.pair_lid[pair], N = g.pair_rid[pair];
    const ShellRef A = shell_ref(g.bas, M), B = shell_ref(g.bas, N);
    const int npp = A.nprim * B.nprim;
    for (int pp = threadIdx.x; pp < npp; pp += blockDim.x) {
        const int pa = pp / B.nprim, pb = pp % B.nprim;
        const double a = g.bas.alpha[A.off + pa], b = g.bas.alpha[B.off + pb], p = a + b;
        double *rec = lds + pp * STRIDE;
        rec[0] = p;
        rec[1] = (a * A.x + b * B.x) / p;
        rec[2] = (a * A.y + b * B.y) / p;
        rec[3] = (a * A.z + b * B.z) / p;
        rec[4] = g.bas.coef[A.off + pa] * g.bas.coef[B.off + pb];
        double E[NE1];
        herm_E<LA, LB>(a, b, A.x - B.x, E);
#pragma unroll
        for (int q = 0; q < NE1; ++q) rec[5 + q] = E[q];
        herm_E<LA, LB>(a, b, A.y - B.y, E);
#pragma unroll
        for (int q = 0; q < NE1; ++q) rec[5 + NE1 + q] = E[q];
        herm_E<LA, LB>(a, b, A.z - B.z, E);
#pragma unroll
        for (int q = 0; q < NE1; ++q) rec[5 + 2 * NE1 + q] = E[q];
    }
    __syncthreads();
    const int pi = blockIdx.y * blockDim.x + threadIdx.x;
    if (pi >= g.nP) return;
    const int P = g.plist[pi];
    if (g.sp_scr[(size_t) M * g.nshells + N] * g.df_scr[P] < g.scrtol2) return;   // build_DF_tensor.c:99-100
    eri3_dispatch<LA, LB, LC, 0>(g, lds, npp, A, B, P, blockIdx.z);
}

template <int LA, int LB, int LC>
int launch_eri3_class(const Eri3Args &g0, hipStream_t s)
{
    if (g0.npairs == 0 || g0.nP == 0) return YGPU_OK;
    constexpr int NE1 = (LA + 1) * (LB + 1) * (LA + LB + 1);
    const size_t lds = (size_t) g0.npp_max * (5 + 3 * NE1) * sizeof(double);
    if (lds > 64 * 1024) {
        ygpu_set_error(YGPU_ERR_INVALID, "3-centre class (%d%d|%d): %d primitive pairs exceed LDS", LA, LB, LC, g0.npp_max);
        return YGPU_ERR_INVALID;
    }
    dim3 grid(g0.npairs, (g0.nP + 63) / 64, ncart(LC));
    hipLaunchKernelGGL((k_eri3<LA, LB, LC>), grid, dim3(64), lds, s, g0);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

template <int LA, int LB>
int launch_eri3_lc(int lc, const Eri3Args &g, hipStream_t s)
{
    switch (lc) {
    case 0: return launch_eri3_class<LA, LB, 0>(g, s);
    case 1: return launch_eri3_class<LA, LB, 1>(g, s);
    case 2: return launch_eri3_class<LA, LB, 2>(g, s);
    case 3: return launch_eri3_class<LA, LB, 3>(g, s);
    }
    ygpu_set_error(YGPU_ERR_INVALID, "aux angular momentum %d not supported (max f)", lc);
    return YGPU_ERR_INVALID;
}

}  // namespace

// ===========================================================================
int yg_launch_ovlp_hcore(ygpu_ctx *c, double *dS, double *dH)
{
    const BasisDev &b = c->bas;
    if (b.max_am > 2) {
        ygpu_set_error(YGPU_ERR_INVALID, "orbital angular momentum %d not supported (max d)", b.max_am);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_ovlp_hcore, dim3(b.nshells * b.nshells), dim3(64), 0, c->stream, b.dev(), b.nshells, b.nbf,
                       c->natoms, c->d_charge.as<double>(), c->d_atom_xyz.as<double>(), c->boys.as<double>(), dS, dH);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_launch_schwarz(ygpu_ctx *c, int quirk, double *d_bf_scr, double *d_df_scr)
{
    const BasisDev &b = c->bas;
    if (b.max_am > 2 || c->aux.max_am > 3) {
        ygpu_set_error(YGPU_ERR_INVALID, "angular momentum beyond d (orbital) / f (aux) not supported");
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_schwarz, dim3(b.nshells * b.nshells), dim3(64), 0, c->stream, b.dev(), b.nshells, b.nbf,
                       quirk, c->boys.as<double>(), d_bf_scr);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_df_scr, dim3((c->aux.nshells + 63) / 64), dim3(64), 0, c->stream, c->aux.dev(),
                       c->aux.nshells, c->boys.as<double>(), d_df_scr);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_launch_eri2(ygpu_ctx *c, double *dJpq, int ld)
{
    const BasisDev &a = c->aux;
    std::vector<int> pl, ql;
    for (int P = 0; P < a.nshells; ++P)
        for (int Q = P; Q < a.nshells; ++Q) {
            if (c->h_df_sp_scrval[P] * c->h_df_sp_scrval[Q] < c->scrtol2) continue;   // build_DF_tensor.c:176
            pl.push_back(P);
            ql.push_back(Q);
        }
    if (pl.empty()) return YGPU_OK;
    YG_TRY(c->d_gA.ensure(pl.size() * sizeof(int)));
    YG_TRY(c->d_gB.ensure(ql.size() * sizeof(int)));
    YG_HIP(hipMemcpyAsync(c->d_gA.p, pl.data(), pl.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(c->d_gB.p, ql.data(), ql.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_eri2, dim3(pl.size()), dim3(128), 0, c->stream, a.dev(), c->d_gA.as<int>(), c->d_gB.as<int>(),
                       c->boys.as<double>(), dJpq, ld);
    YG_HIP(hipGetLastError());
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
}

int yg_launch_eri3(ygpu_ctx *c, double *dpqA)
{
    const BasisDev &b = c->bas, &a = c->aux;
    if (b.max_am > 2 || a.max_am > 3) {
        ygpu_set_error(YGPU_ERR_INVALID, "angular momentum beyond d (orbital) / f (aux) not supported");
        return YGPU_ERR_INVALID;
    }
    // pair lists per (la,lb) class, aux lists per lc
    std::vector<int> lid[3][3], rid[3][3];
    int ppmax[3][3] = {};
    for (int i = 0; i < c->num_uniq_sp; ++i) {
        const int M = c->h_uniq_lid[i], N = c->h_uniq_rid[i];
        const int la = b.am[M], lb = b.am[N];
        if (la < lb) {
            ygpu_set_error(YGPU_ERR_INVALID, "unique shell pair %d has AM(lid) < AM(rid)", i);
            return YGPU_ERR_INVALID;
        }
        lid[la][lb].push_back(M);
        rid[la][lb].push_back(N);
        ppmax[la][lb] = std::max(ppmax[la][lb], b.nprim[M] * b.nprim[N]);
    }
    std::vector<int> plist[4];
    for (int P = 0; P < a.nshells; ++P) plist[a.am[P]].push_back(P);
    // one device buffer for all lists
    std::vector<int> all;
    size_t off_l[3][3], off_r[3][3], off_p[4];
    for (int la = 0; la < 3; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            off_l[la][lb] = all.size(); all.insert(all.end(), lid[la][lb].begin(), lid[la][lb].end());
            off_r[la][lb] = all.size(); all.insert(all.end(), rid[la][lb].begin(), rid[la][lb].end());
        }
    for (int lc = 0; lc < 4; ++lc) { off_p[lc] = all.size(); all.insert(all.end(), plist[lc].begin(), plist[lc].end()); }
    YG_TRY(c->d_gC.ensure(std::max<size_t>(1, all.size()) * sizeof(int)));
    YG_HIP(hipMemcpyAsync(c->d_gC.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int *base = c->d_gC.as<int>();
    Eri3Args g{};
    g.bas = b.dev(); g.aux = a.dev();
    g.nshells = b.nshells; g.nbf = b.nbf;
    g.upmask = c->d_upmask.as<int>();
    g.sp_scr = c->d_sp_scr.as<double>(); g.df_scr = c->d_df_scr.as<double>();
    g.scrtol2 = c->scrtol2;
    g.tab = c->boys.as<double>();
    g.out = dpqA; g.ld = c->naux_pad;
    for (int la = 0; la < 3; ++la)
        for (int lb = 0; lb <= la; ++lb)
            for (int lc = 0; lc < 4; ++lc) {
                g.pair_lid = base + off_l[la][lb]; g.pair_rid = base + off_r[la][lb];
                g.npairs = (int) lid[la][lb].size();
                g.plist = base + off_p[lc]; g.nP = (int) plist[lc].size();
                g.npp_max = ppmax[la][lb];
                int st = YGPU_OK;
                if (la == 0) st = launch_eri3_lc<0, 0>(lc, g, c->stream);
                else if (la == 1 && lb == 0) st = launch_eri3_lc<1, 0>(lc, g, c->stream);
                else if (la == 1) st = launch_eri3_lc<1, 1>(lc, g, c->stream);
                else if (lb == 0) st = launch_eri3_lc<2, 0>(lc, g, c->stream);
                else if (lb == 1) st = launch_eri3_lc<2, 1>(lc, g, c->stream);
                else st = launch_eri3_lc<2, 2>(lc, g, c->stream);
                if (st != YGPU_OK) return st;
            }
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
}
