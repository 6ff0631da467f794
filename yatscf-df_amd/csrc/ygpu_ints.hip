// ygpu_ints.hip -- one-electron, Schwarz and 2-centre integrals on gfx950.
//
// Replaces the reference's Simint calls:
//   overlap / kinetic / potential  (CMS_Simint.c:402-461, used by TinySCF.c:20-54)
//   (MN|MN) Schwarz elements       (CMS_Simint.c:337-399, used by TinySCF.c:112-177)
//   (P|Q) metric                   (CMS_Simint.c:569-630, build_DF_tensor.c:152-203)
// by the McMurchie-Davidson scheme (md.h).  One-time work: one workgroup per
// shell pair of an angular-momentum class; each Cartesian element gets
// 64 / next_pow2(elements) lanes that split its primitive-pair loop.
#include "ygpu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

// ---------------------------------------------------------------------------
// One-time integrals, templated on the shell-pair angular momenta so that all
// Hermite arrays have compile-time shapes (no scratch); Cartesian components
// are run-time per thread and enter only through path-wise recursions and
// select chains.

// lanes sharing one output element: 64 / (smallest power of two >= n)
__host__ __device__ constexpr int lanes_per_elem(int n)
{
    int p = 1;
    while (p < n) p *= 2;
    return p >= 64 ? 1 : 64 / p;
}

// sum over the LPE adjacent lanes that share an element (LPE a power of two)
template <int LPE>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = LPE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// value E0[j] for a run-time j in [lo, hi], zero outside
template <int N>
__device__ __forceinline__ double pick(const double *E0, int j)
{
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q)
        if (q == j) v = E0[q];
    return v;
}

// sum_{t+u+v <= LAB} Ex[t] Ey[u] Ez[v] sum_{tau+nu+phi <= LCD} (-1)^{tau+nu+phi} Fx Fy Fz R[t+tau][u+nu][v+phi]
template <int LAB, int LCD>
__device__ __forceinline__ double herm_contract(const double *Ex, const double *Ey, const double *Ez,
                                                const double *Fx, const double *Fy, const double *Fz, const double *R)
{
    constexpr int D = LAB + LCD + 1;
    double s = 0.0;
#pragma unroll
    for (int t = 0; t <= LAB; ++t)
#pragma unroll
        for (int u = 0; u <= LAB - t; ++u)
#pragma unroll
            for (int v = 0; v <= LAB - t - u; ++v) {
                double w = 0.0;
#pragma unroll
                for (int tau = 0; tau <= LCD; ++tau)
#pragma unroll
                    for (int nu = 0; nu <= LCD - tau; ++nu)
#pragma unroll
                        for (int phi = 0; phi <= LCD - tau - nu; ++phi) {
                            const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                            w += sg * Fx[tau] * Fy[nu] * Fz[phi] * R[((t + tau) * D + u + nu) * D + v + phi];
                        }
                s += Ex[t] * Ey[u] * Ez[v] * w;
            }
    return s;
}

// S and Hcore = T + V: one workgroup per shell pair of the (LA,LB) class, LPE lanes per element
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_ovlp_hcore(DevShells sh, const int *__restrict__ pM, const int *__restrict__ pN,
                                                   int nbf, int natoms, const double *__restrict__ Zc,
                                                   const double *__restrict__ axyz, const double *__restrict__ tab,
                                                   double *__restrict__ S, double *__restrict__ H)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), LAB = LA + LB;
    constexpr int LPE = lanes_per_elem(NA * NB);
    const ShellRef A = shell_ref(sh, pM[blockIdx.x]), B = shell_ref(sh, pN[blockIdx.x]);
    const int e = threadIdx.x / LPE, sub = threadIdx.x % LPE;
    const bool active = e < NA * NB;
    const int ia = active ? e / NB : 0, ib = active ? e % NB : 0;
    int ax, ay, az, bx, by, bz;
    cart_pow_rt(LA, ia, ax, ay, az);
    cart_pow_rt(LB, ib, bx, by, bz);
    const int bp[3] = {bx, by, bz}, ap[3] = {ax, ay, az};
    double s_acc = 0.0, t_acc = 0.0, v_acc = 0.0;
    const int npp = A.nprim * B.nprim;
    for (int q = sub; q < npp; q += LPE) {
        const int pa = q / B.nprim, pb = q % B.nprim;
        const double a = sh.alpha[A.off + pa], b = sh.alpha[B.off + pb], p = a + b;
        const double cab = sh.coef[A.off + pa] * sh.coef[B.off + pb];
        const double d3[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
        const double sq = sqrt(YG_PI / p);
        double S1[3], T1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double E0[LB + 3];
            herm_E0_row<LA, LB + 2>(a, b, d3[k], ap[k], E0);
            const int j = bp[k];
            const double s0 = pick<LB + 3>(E0, j) * sq;
            const double sp2 = pick<LB + 3>(E0, j + 2) * sq;
            const double sm2 = pick<LB + 3>(E0, j - 2) * sq;
            S1[k] = s0;
            T1[k] = -2.0 * b * b * sp2 + b * (2 * j + 1) * s0 - 0.5 * j * (j - 1) * sm2;
        }
        s_acc += cab * S1[0] * S1[1] * S1[2];
        t_acc += cab * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2]);
        double Ex[LAB + 1], Ey[LAB + 1], Ez[LAB + 1];
        herm_E1d<LA, LB>(a, b, d3[0], ax, bx, Ex);
        herm_E1d<LA, LB>(a, b, d3[1], ay, by, Ey);
        herm_E1d<LA, LB>(a, b, d3[2], az, bz, Ez);
        const double Px = (a * A.x + b * B.x) / p, Py = (a * A.y + b * B.y) / p, Pz = (a * A.z + b * B.z) / p;
        const double cv = cab * 2.0 * YG_PI / p;
        const double one[1] = {1.0};
        for (int c = 0; c < natoms; ++c) {
            double R[(LAB + 1) * (LAB + 1) * (LAB + 1)];
            herm_R<LAB>(p, Px - axyz[3 * c], Py - axyz[3 * c + 1], Pz - axyz[3 * c + 2], tab, R);
            v_acc -= Zc[c] * cv * herm_contract<LAB, 0>(Ex, Ey, Ez, one, one, one, R);
        }
    }
    s_acc = group_sum<LPE>(s_acc);
    t_acc = group_sum<LPE>(t_acc);
    v_acc = group_sum<LPE>(v_acc);
    if (!active || sub != 0) return;
    const size_t o = (size_t)(A.bf + ia) * nbf + B.bf + ib;
    S[o] = s_acc;
    H[o] = t_acc + v_acc;
}

// Schwarz values: one workgroup per ordered shell pair (M,N) of the (LA,LB) class,
// LPE lanes per (iM,iN); the element read is the reference's (TinySCF.c:153) when quirk != 0
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz(DevShells sh, const int *__restrict__ pM, const int *__restrict__ pN,
                                                int nbf, int quirk, const double *__restrict__ tab,
                                                double *__restrict__ out)
{
    constexpr int dM = ncart(LA), dN = ncart(LB), LAB = LA + LB, L = 2 * LAB;
    constexpr int LPE = lanes_per_elem(dM * dN);
    const ShellRef A = shell_ref(sh, pM[blockIdx.x]), B = shell_ref(sh, pN[blockIdx.x]);
    const int e = threadIdx.x / LPE, sub = threadIdx.x % LPE;
    const bool active = e < dM * dN;
    const int iM = active ? e / dN : 0, iN = active ? e % dN : 0;
    int a, b, c, d;
    if (quirk) {
        int idx = iN * (dM * dN * dM + dM) + iM * (dN * dM + 1);
        d = idx % dN; idx /= dN;
        c = idx % dM; idx /= dM;
        b = idx % dN; idx /= dN;
        a = idx;
    } else {
        a = iM; b = iN; c = iM; d = iN;
    }
    int ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
    cart_pow_rt(LA, a, ax, ay, az);
    cart_pow_rt(LB, b, bx, by, bz);
    cart_pow_rt(LA, c, cx, cy, cz);
    cart_pow_rt(LB, d, dx, dy, dz);
    const double XAB = A.x - B.x, YAB = A.y - B.y, ZAB = A.z - B.z;
    double res = 0.0;
    const int npp = A.nprim * B.nprim;
    // bra primitive pairs spread over the LPE lanes of this element
    for (int qb = sub; qb < npp; qb += LPE) {
        const int pa = qb / B.nprim, pb = qb % B.nprim;
        const double al = sh.alpha[A.off + pa], be = sh.alpha[B.off + pb], p = al + be;
        const double Px = (al * A.x + be * B.x) / p, Py = (al * A.y + be * B.y) / p, Pz = (al * A.z + be * B.z) / p;
        double Ex[LAB + 1], Ey[LAB + 1], Ez[LAB + 1];
        herm_E1d<LA, LB>(al, be, XAB, ax, bx, Ex);
        herm_E1d<LA, LB>(al, be, YAB, ay, by, Ey);
        herm_E1d<LA, LB>(al, be, ZAB, az, bz, Ez);
        const double cab = sh.coef[A.off + pa] * sh.coef[B.off + pb];
        for (int pc = 0; pc < A.nprim; ++pc)
            for (int pd = 0; pd < B.nprim; ++pd) {
                const double ga = sh.alpha[A.off + pc], de = sh.alpha[B.off + pd], q = ga + de;
                const double Qx = (ga * A.x + de * B.x) / q, Qy = (ga * A.y + de * B.y) / q, Qz = (ga * A.z + de * B.z) / q;
                double Fx[LAB + 1], Fy[LAB + 1], Fz[LAB + 1];
                herm_E1d<LA, LB>(ga, de, XAB, cx, dx, Fx);
                herm_E1d<LA, LB>(ga, de, YAB, cy, dy, Fy);
                herm_E1d<LA, LB>(ga, de, ZAB, cz, dz, Fz);
                double R[(L + 1) * (L + 1) * (L + 1)];
                herm_R<L>(p * q / (p + q), Px - Qx, Py - Qy, Pz - Qz, tab, R);
                res += cab * sh.coef[A.off + pc] * sh.coef[B.off + pd] * YG_TWO_PI_25 / (p * q * sqrt(p + q)) *
                       herm_contract<LAB, LAB>(Ex, Ey, Ez, Fx, Fy, Fz, R);
            }
    }
    res = group_sum<LPE>(res);
    if (!active || sub != 0) return;
    out[(size_t)(A.bf + iM) * nbf + B.bf + iN] = fabs(res);
}

// (P_i|Q_j) between two single-centre shells of angular momenta (LP, LQ)
template <int LP, int LQ>
__device__ __forceinline__ double eri2_elem(const DevShells &sh, const ShellRef &C, const ShellRef &Dd, int ic, int id,
                                            const double *__restrict__ tab)
{
    constexpr int L = LP + LQ, D = L + 1;
    int cx, cy, cz, dx, dy, dz;
    cart_pow_rt(LP, ic, cx, cy, cz);
    cart_pow_rt(LQ, id, dx, dy, dz);
    double res = 0.0;
    for (int pc = 0; pc < C.nprim; ++pc)
        for (int pd = 0; pd < Dd.nprim; ++pd) {
            const double g = sh.alpha[C.off + pc], h = sh.alpha[Dd.off + pd];
            // a single Gaussian is the pair (g, 0) on one centre
            double Ex[LP + 1], Ey[LP + 1], Ez[LP + 1], Fx[LQ + 1], Fy[LQ + 1], Fz[LQ + 1];
            herm_E1d<LP, 0>(g, 0.0, 0.0, cx, 0, Ex);
            herm_E1d<LP, 0>(g, 0.0, 0.0, cy, 0, Ey);
            herm_E1d<LP, 0>(g, 0.0, 0.0, cz, 0, Ez);
            herm_E1d<LQ, 0>(h, 0.0, 0.0, dx, 0, Fx);
            herm_E1d<LQ, 0>(h, 0.0, 0.0, dy, 0, Fy);
            herm_E1d<LQ, 0>(h, 0.0, 0.0, dz, 0, Fz);
            double R[D * D * D];
            herm_R<L>(g * h / (g + h), C.x - Dd.x, C.y - Dd.y, C.z - Dd.z, tab, R);
            res += sh.coef[C.off + pc] * sh.coef[Dd.off + pd] * YG_TWO_PI_25 / (g * h * sqrt(g + h)) *
                   herm_contract<LP, LQ>(Ex, Ey, Ez, Fx, Fy, Fz, R);
        }
    (void) D;
    return res;
}

// max_k (P_k|P_k) for the aux shells of one angular momentum
template <int L>
__global__ __launch_bounds__(64) void k_df_scr(DevShells sh, const int *__restrict__ plist, int n, const double *__restrict__ tab,
                         double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int P = plist[i];
    const ShellRef A = shell_ref(sh, P);
    double m = 0.0;
    for (int k = 0; k < ncart(L); ++k) m = fmax(m, fabs(eri2_elem<L, L>(sh, A, A, k, k, tab)));
    out[P] = m;
}

// J_PQ for the screened aux shell pairs of one (LP,LQ) class (both triangles written)
template <int LP, int LQ>
__global__ __launch_bounds__(128) void k_eri2(DevShells sh, const int *__restrict__ plist, const int *__restrict__ qlist,
                                              const double *__restrict__ tab, double *__restrict__ J, int ld)
{
    constexpr int NP = ncart(LP), NQ = ncart(LQ);
    const ShellRef A = shell_ref(sh, plist[blockIdx.x]), B = shell_ref(sh, qlist[blockIdx.x]);
    const int e = threadIdx.x;
    if (e >= NP * NQ) return;
    const int ia = e / NQ, ib = e % NQ;
    const double v = eri2_elem<LP, LQ>(sh, A, B, ia, ib, tab);
    J[(size_t)(A.bf + ia) * ld + B.bf + ib] = v;
    J[(size_t)(B.bf + ib) * ld + A.bf + ia] = v;
}

}  // namespace

// ===========================================================================
namespace {

// per-class work lists uploaded in one buffer: returns device offsets
struct ClassLists {
    std::vector<int> all;
    size_t off[5][5][2] = {};
    int count[5][5] = {};
};

int upload_lists(ygpu_ctx *c, ClassLists &cl, DBuf &buf)
{
    YG_TRY(buf.ensure(std::max<size_t>(1, cl.all.size()) * sizeof(int)));
    YG_HIP(hipMemcpyAsync(buf.p, cl.all.data(), cl.all.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    return YGPU_OK;
}

// ordered pairs (M,N) of the orbital basis grouped by (am(M), am(N))
void pair_lists(const BasisDev &b, ClassLists &cl, bool upper_only, const std::vector<double> *scr = nullptr,
                double tol2 = 0.0)
{
    std::vector<int> L[5][5][2];
    for (int M = 0; M < b.nshells; ++M)
        for (int N = upper_only ? M : 0; N < b.nshells; ++N) {
            if (scr && (*scr)[M] * (*scr)[N] < tol2) continue;
            L[b.am[M]][b.am[N]][0].push_back(M);
            L[b.am[M]][b.am[N]][1].push_back(N);
        }
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            cl.count[i][j] = (int) L[i][j][0].size();
            for (int k = 0; k < 2; ++k) {
                cl.off[i][j][k] = cl.all.size();
                cl.all.insert(cl.all.end(), L[i][j][k].begin(), L[i][j][k].end());
            }
        }
}

}  // namespace

#define YG_CLASS_SWITCH2(LMAXA, LMAXB, la, lb, CALL)                                         \
    do {                                                                                     \
        bool done_ = false;                                                                  \
        if (la == 0 && lb == 0) { CALL(0, 0); done_ = true; }                                \
        if (la == 0 && lb == 1) { CALL(0, 1); done_ = true; }                                \
        if (la == 0 && lb == 2) { CALL(0, 2); done_ = true; }                                \
        if (la == 1 && lb == 0) { CALL(1, 0); done_ = true; }                                \
        if (la == 1 && lb == 1) { CALL(1, 1); done_ = true; }                                \
        if (la == 1 && lb == 2) { CALL(1, 2); done_ = true; }                                \
        if (la == 2 && lb == 0) { CALL(2, 0); done_ = true; }                                \
        if (la == 2 && lb == 1) { CALL(2, 1); done_ = true; }                                \
        if (la == 2 && lb == 2) { CALL(2, 2); done_ = true; }                                \
        if (!done_) {                                                                        \
            ygpu_set_error(YGPU_ERR_INVALID, "angular momentum pair (%d,%d) not supported", la, lb); \
            return YGPU_ERR_INVALID;                                                         \
        }                                                                                    \
    } while (0)

int yg_launch_ovlp_hcore(ygpu_ctx *c, double *dS, double *dH)
{
    const BasisDev &b = c->bas;
    if (b.max_am > 2) {
        ygpu_set_error(YGPU_ERR_INVALID, "orbital angular momentum %d not supported (max d)", b.max_am);
        return YGPU_ERR_INVALID;
    }
    ClassLists cl;
    pair_lists(b, cl, false);
    YG_TRY(upload_lists(c, cl, c->d_gA));
    const int *base = c->d_gA.as<int>();
    for (int la = 0; la <= 2; ++la)
        for (int lb = 0; lb <= 2; ++lb) {
            const int n = cl.count[la][lb];
            if (n == 0) continue;
            const int *pM = base + cl.off[la][lb][0], *pN = base + cl.off[la][lb][1];
#define CALL(A_, B_)                                                                                         \
    hipLaunchKernelGGL((k_ovlp_hcore<A_, B_>), dim3(n), dim3(64), 0, c->stream, b.dev(), pM, pN, b.nbf,    \
                       c->natoms, c->d_charge.as<double>(), c->d_atom_xyz.as<double>(), c->boys.as<double>(), dS, dH)
            YG_CLASS_SWITCH2(2, 2, la, lb, CALL);
#undef CALL
            YG_HIP(hipGetLastError());
        }
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
}

int yg_launch_schwarz(ygpu_ctx *c, int quirk, double *d_bf_scr, double *d_df_scr)
{
    const BasisDev &b = c->bas, &a = c->aux;
    if (b.max_am > 2 || a.max_am > 3) {
        ygpu_set_error(YGPU_ERR_INVALID, "angular momentum beyond d (orbital) / f (aux) not supported");
        return YGPU_ERR_INVALID;
    }
    ClassLists cl;
    pair_lists(b, cl, false);
    // aux shells by angular momentum, appended to the same buffer
    size_t aoff[4];
    int acnt[4];
    for (int l = 0; l < 4; ++l) {
        aoff[l] = cl.all.size();
        acnt[l] = 0;
        for (int P = 0; P < a.nshells; ++P)
            if (a.am[P] == l) { cl.all.push_back(P); acnt[l]++; }
    }
    YG_TRY(upload_lists(c, cl, c->d_gA));
    const int *base = c->d_gA.as<int>();
    for (int la = 0; la <= 2; ++la)
        for (int lb = 0; lb <= 2; ++lb) {
            const int n = cl.count[la][lb];
            if (n == 0) continue;
            const int *pM = base + cl.off[la][lb][0], *pN = base + cl.off[la][lb][1];
#define CALL(A_, B_)                                                                                     \
    hipLaunchKernelGGL((k_schwarz<A_, B_>), dim3(n), dim3(64), 0, c->stream, b.dev(), pM, pN, b.nbf, quirk, \
                       c->boys.as<double>(), d_bf_scr)
            YG_CLASS_SWITCH2(2, 2, la, lb, CALL);
#undef CALL
            YG_HIP(hipGetLastError());
        }
    for (int l = 0; l < 4; ++l) {
        if (acnt[l] == 0) continue;
        const dim3 g((acnt[l] + 63) / 64);
        const int *pl = base + aoff[l];
        if (l == 0) hipLaunchKernelGGL((k_df_scr<0>), g, dim3(64), 0, c->stream, a.dev(), pl, acnt[l], c->boys.as<double>(), d_df_scr);
        if (l == 1) hipLaunchKernelGGL((k_df_scr<1>), g, dim3(64), 0, c->stream, a.dev(), pl, acnt[l], c->boys.as<double>(), d_df_scr);
        if (l == 2) hipLaunchKernelGGL((k_df_scr<2>), g, dim3(64), 0, c->stream, a.dev(), pl, acnt[l], c->boys.as<double>(), d_df_scr);
        if (l == 3) hipLaunchKernelGGL((k_df_scr<3>), g, dim3(64), 0, c->stream, a.dev(), pl, acnt[l], c->boys.as<double>(), d_df_scr);
        YG_HIP(hipGetLastError());
    }
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
}

int yg_launch_eri2(ygpu_ctx *c, double *dJpq, int ld)
{
    const BasisDev &a = c->aux;
    if (a.max_am > 3) {
        ygpu_set_error(YGPU_ERR_INVALID, "aux angular momentum %d not supported (max f)", a.max_am);
        return YGPU_ERR_INVALID;
    }
    ClassLists cl;
    pair_lists(a, cl, true, &c->h_df_sp_scrval, c->scrtol2);   // build_DF_tensor.c:176
    YG_TRY(upload_lists(c, cl, c->d_gB));
    const int *base = c->d_gB.as<int>();
    for (int lp = 0; lp <= 3; ++lp)
        for (int lq = 0; lq <= 3; ++lq) {
            const int n = cl.count[lp][lq];
            if (n == 0) continue;
            const int *pl = base + cl.off[lp][lq][0], *ql = base + cl.off[lp][lq][1];
#define CALL(A_, B_) \
    hipLaunchKernelGGL((k_eri2<A_, B_>), dim3(n), dim3(128), 0, c->stream, a.dev(), pl, ql, c->boys.as<double>(), dJpq, ld)
            switch (lp * 4 + lq) {
            case 0: CALL(0, 0); break;  case 1: CALL(0, 1); break;  case 2: CALL(0, 2); break;  case 3: CALL(0, 3); break;
            case 4: CALL(1, 0); break;  case 5: CALL(1, 1); break;  case 6: CALL(1, 2); break;  case 7: CALL(1, 3); break;
            case 8: CALL(2, 0); break;  case 9: CALL(2, 1); break;  case 10: CALL(2, 2); break; case 11: CALL(2, 3); break;
            case 12: CALL(3, 0); break; case 13: CALL(3, 1); break; case 14: CALL(3, 2); break; case 15: CALL(3, 3); break;
            }
#undef CALL
            YG_HIP(hipGetLastError());
        }
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
}

