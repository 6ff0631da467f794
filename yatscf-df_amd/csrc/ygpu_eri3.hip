// ygpu_eri3.hip -- the 3-centre integrals (mn|P) on gfx950 (hot, one-time per SCF).
//
// Replaces the reference's calc_DF_3center_integrals (build_DF_tensor.c:49-150)
// and its Simint batch calls (CMS_Simint.c:474-566).  McMurchie-Davidson with
// the aux shell as a single-centre Hermite expansion (the reference pairs each
// aux shell with an exponent-0 unit shell, CMS_Simint.c:131-141).  Instantiated
// per angular-momentum class (la >= lb <= 2, lc <= 3) and per group of Cartesian
// components of P so that every Hermite index is a compile-time constant (the R
// recursion stays in registers, and is formed once for the group); one
// wavefront handles one screened shell pair against 64 aux shells of one class,
// with the bra primitive-pair data (P centre, exponent, coefficient product,
// Hermite coefficients) staged once per workgroup in LDS.  Results go straight
// into the packed pair rows (i <= j).
#include "ygpu_internal.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace {

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

// Cartesian components of P computed together by one lane: as many as keep the accumulators of
// the group within 48 doubles (measured: two for the (d p|*) classes, one for (d d|*)), the
// components dealt evenly over the groups.  The Hermite integrals R of a primitive quadruple are
// formed once per group instead of once per component.
__host__ __device__ constexpr int eri3_groups(int la, int lb, int lc)
{
    const int nab = ncart(la) * ncart(lb), nc = ncart(lc);
    const int cap = 48 / nab;
    const int g = cap < 1 ? 1 : (cap > nc ? nc : cap);
    return (nc + g - 1) / g;
}
__host__ __device__ constexpr int eri3_group_size(int la, int lb, int lc)
{
    const int ng = eri3_groups(la, lb, lc), nc = ncart(lc);
    return (nc + ng - 1) / ng;
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// components K0 .. K0 + G - 1 of P for one lane
template <int LA, int LB, int LC, int K0, int G>
__device__ __forceinline__ void eri3_body(const Eri3Args &g, const double *__restrict__ lds, int npp,
                                          const ShellRef &A, const ShellRef &B, int P)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), LAB = LA + LB, L = LAB + LC, D = L + 1;
    constexpr int NE1 = (LA + 1) * (LB + 1) * (LAB + 1);
    constexpr int STRIDE = 5 + 3 * NE1;
    constexpr int DW = LAB + 1;
    const ShellRef C = shell_ref(g.aux, P);
    double acc[G][NA * NB];
#pragma unroll
    for (int k = 0; k < G; ++k)
#pragma unroll
        for (int q = 0; q < NA * NB; ++q) acc[k][q] = 0.0;
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
            const double *Ex = rec + 5, *Ey = rec + 5 + NE1, *Ez = rec + 5 + 2 * NE1;
            static_for<0, G>([&](auto gi) {
                constexpr int KK = K0 + decltype(gi)::value;
                constexpr int cx = cart_pow(LC, KK, 0), cy = cart_pow(LC, KK, 1), cz = cart_pow(LC, KK, 2);
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
                                        const double e = Ec[cx * (LC + 1) + tau] * Ec[cy * (LC + 1) + nu] *
                                                         Ec[cz * (LC + 1) + phi];
                                        const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                        s += sg * e * R[((t + tau) * D + u + nu) * D + v + phi];
                                    }
                            W[(t * DW + u) * DW + v] = s * pref;
                        }
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ib = 0; ib < NB; ++ib) {
                        const int ax = cart_pow(LA, ia, 0), ay = cart_pow(LA, ia, 1), az = cart_pow(LA, ia, 2);
                        const int bx = cart_pow(LB, ib, 0), by = cart_pow(LB, ib, 1), bz = cart_pow(LB, ib, 2);
                        double s = 0.0;
#pragma unroll
                        for (int t = 0; t <= ax + bx; ++t)
#pragma unroll
                            for (int u = 0; u <= ay + by; ++u) {
                                const double exy = Ex[(ax * (LB + 1) + bx) * (LAB + 1) + t] *
                                                   Ey[(ay * (LB + 1) + by) * (LAB + 1) + u];
#pragma unroll
                                for (int v = 0; v <= az + bz; ++v)
                                    s += exy * Ez[(az * (LB + 1) + bz) * (LAB + 1) + v] * W[(t * DW + u) * DW + v];
                            }
                        acc[decltype(gi)::value][ia * NB + ib] += s;
                    }
            });
        }
    }
    const bool same = (A.bf == B.bf);
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int col = C.bf + K0 + k;
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                const int m = A.bf + ia, n = B.bf + ib;
                if (same && ia > ib) continue;          // (m,n) and (n,m) share one packed row
                const int lo = m < n ? m : n, hi = m < n ? n : m;
                const int r = g.upmask[(size_t) lo * g.nbf + hi];
                if (r >= 0) g.out[(size_t) r * g.ld + col] = acc[k][ia * NB + ib];
            }
    }
}

// group z of the components: K0 = z GS, the last group possibly shorter
template <int LA, int LB, int LC, int Z>
__device__ __forceinline__ void eri3_dispatch(const Eri3Args &g, const double *lds, int npp, const ShellRef &A,
                                              const ShellRef &B, int P, int z)
{
    constexpr int GS = eri3_group_size(LA, LB, LC), NC = ncart(LC);
    if constexpr (Z * GS < NC) {
        constexpr int K0 = Z * GS, G = (NC - K0 < GS) ? NC - K0 : GS;
        if (z == Z) eri3_body<LA, LB, LC, K0, G>(g, lds, npp, A, B, P);
        else eri3_dispatch<LA, LB, LC, Z + 1>(g, lds, npp, A, B, P, z);
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
    dim3 grid(g0.npairs, (g0.nP + 63) / 64, eri3_groups(LA, LB, LC));
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

int yg_launch_eri3(ygpu_ctx *c, double *dpqA, hipStream_t s)
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
    YG_HIP(hipMemcpyAsync(c->d_gC.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, s));
    YG_HIP(hipStreamSynchronize(s));            // the host lists go out of scope on return
    const int *base = c->d_gC.as<int>();
    Eri3Args g{};
    g.bas = b.dev(); g.aux = a.dev();
    g.nshells = b.nshells; g.nbf = b.nbf;
    g.upmask = c->d_upmask.as<int>();
    g.sp_scr = c->d_sp_scr.as<double>(); g.df_scr = c->d_df_scr.as<double>();
    g.scrtol2 = c->scrtol2;
    g.tab = c->boys.as<double>();
    g.out = dpqA; g.ld = c->naux_cart_pad;   // the Cartesian auxiliary functions
    for (int la = 0; la < 3; ++la)
        for (int lb = 0; lb <= la; ++lb)
            for (int lc = 0; lc < 4; ++lc) {
                g.pair_lid = base + off_l[la][lb]; g.pair_rid = base + off_r[la][lb];
                g.npairs = (int) lid[la][lb].size();
                g.plist = base + off_p[lc]; g.nP = (int) plist[lc].size();
                g.npp_max = ppmax[la][lb];
                int st = YGPU_OK;
                if (la == 0) st = launch_eri3_lc<0, 0>(lc, g, s);
                else if (la == 1 && lb == 0) st = launch_eri3_lc<1, 0>(lc, g, s);
                else if (la == 1) st = launch_eri3_lc<1, 1>(lc, g, s);
                else if (lb == 0) st = launch_eri3_lc<2, 0>(lc, g, s);
                else if (lb == 1) st = launch_eri3_lc<2, 1>(lc, g, s);
                else st = launch_eri3_lc<2, 2>(lc, g, s);
                if (st != YGPU_OK) return st;
            }
    return YGPU_OK;
}
