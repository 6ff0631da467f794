// ygpu_sph.hip -- spherical-harmonic basis option.
//
// The reference reads a "spherical" flag from the first line of a .gbs file
// (CMS_BasisSet.c:464-483) and then counts 2l+1 functions per shell (:277-281)
// but has no integral code for such shells (its Simint path is Cartesian).
// Here a spherical basis is the span of the real solid harmonics inside the
// Cartesian one: every integral stays Cartesian, and
//   - for the orbital basis the host orthogonaliser X = T (T^T S T)^{-1/2}
//     (zero columns beyond the spherical count) confines the orbitals to that
//     span, so the SCF is the spherical-basis SCF written in Cartesian
//     components;
//   - for the auxiliary basis the columns of (mn|P) and both sides of J_PQ are
//     transformed by T before the metric factor, so the fit is the
//     spherical-basis fit.
// T per shell: real solid harmonics of degree l, m = -l..l (Helgaker,
// Jorgensen and Olsen, Molecular Electronic-Structure Theory, eqs. 6.4.47-49),
// over the Cartesian components x^a y^b z^c (a = l..0, b = l-a..0) normalised
// like x^l; each column scaled to unit norm in the metric of those components.
// s and p shells: the identity (the two bases coincide).
#include "ygpu_internal.h"

#include <cmath>
#include <vector>

namespace {

double binom(int n, int k)
{
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

double dfact(int n)      // n!! with (-1)!! = 1
{
    double r = 1.0;
    for (int i = n; i > 1; i -= 2) r *= i;
    return r;
}

int cart_index(int l, int a, int b)
{
    // position of x^a y^b z^(l-a-b) in the order a = l..0, b = l-a..0
    int idx = 0;
    for (int i = l; i > a; --i) idx += l - i + 1;
    return idx + (l - a - b);
}

// y[r][q] = sum_c t_q[c] x[r][cs_q + c] for q < ncols, 0 for ncols <= q < ldout
__global__ void k_aux_sph_cols(long rows, const double *__restrict__ in, long ldin, double *__restrict__ out,
                               long ldout, int ncols, const int *__restrict__ cs, const int *__restrict__ nc,
                               const double *__restrict__ coef, int cmax)
{
    const long total = rows * ldout;
    for (long idx = (long) blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long) gridDim.x * blockDim.x) {
        const long r = idx / ldout;
        const int q = (int) (idx - r * ldout);
        double s = 0.0;
        if (q < ncols) {
            const double *x = in + r * ldin + cs[q], *t = coef + (size_t) q * cmax;
            for (int c = 0; c < nc[q]; ++c) s += t[c] * x[c];
        }
        out[idx] = s;
    }
}

// A[k][k] += shift for n_orth <= k < n
__global__ void k_shift_tail_diag(int n, int n_orth, double *__restrict__ A, double shift)
{
    const int k = n_orth + (int) (blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) A[(size_t) k * n + k] += shift;
}

}  // namespace

int yg_sph_coefficients(int l, double *T)
{
    if (l < 0 || l > 4 || !T) return YGPU_ERR_INVALID;
    const int nc = (l + 1) * (l + 2) / 2, ns = 2 * l + 1;
    for (int i = 0; i < nc * ns; ++i) T[i] = 0.0;
    if (l <= 1) {
        for (int i = 0; i < nc; ++i) T[i * ns + i] = 1.0;
        return YGPU_OK;
    }
    for (int m = -l; m <= l; ++m) {
        const int col = m + l, am = m < 0 ? -m : m, vm2 = m < 0 ? 1 : 0;   // vm2 = 2 v_m
        for (int t = 0; t <= (l - am) / 2; ++t)
            for (int u = 0; u <= t; ++u)
                for (int vv = vm2; vv <= am; vv += 2) {                      // vv = 2 v
                    const double sign = ((t + (vv - vm2) / 2) & 1) ? -1.0 : 1.0;
                    const double cf = sign * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                                      binom(am, vv);
                    const int ax = 2 * t + am - 2 * u - vv, ay = 2 * u + vv;
                    T[cart_index(l, ax, ay) * ns + col] += cf;
                }
    }
    // unit norm in the metric of the x^l-normalised components:
    // <a b c | a' b' c'> = (a+a'-1)!! (b+b'-1)!! (c+c'-1)!! / (2l-1)!!, zero for an odd sum
    std::vector<int> pa(nc), pb(nc), pc(nc);
    for (int a = l, i = 0; a >= 0; --a)
        for (int b = l - a; b >= 0; --b, ++i) {
            pa[i] = a;
            pb[i] = b;
            pc[i] = l - a - b;
        }
    const double dl = dfact(2 * l - 1);
    for (int s = 0; s < ns; ++s) {
        double n2 = 0.0;
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) {
                const int x = pa[i] + pa[j], y = pb[i] + pb[j], z = pc[i] + pc[j];
                if ((x | y | z) & 1) continue;
                n2 += T[i * ns + s] * T[j * ns + s] * dfact(x - 1) * dfact(y - 1) * dfact(z - 1) / dl;
            }
        const double f = 1.0 / std::sqrt(n2);
        for (int i = 0; i < nc; ++i) T[i * ns + s] *= f;
    }
    return YGPU_OK;
}

int yg_aux_sph_setup(ygpu_ctx *c)
{
    const BasisDev &a = c->aux;
    int nsph = 0, cmax = 1;
    for (int P = 0; P < a.nshells; ++P) {
        nsph += 2 * a.am[P] + 1;
        cmax = std::max(cmax, (a.am[P] + 1) * (a.am[P] + 2) / 2);
    }
    std::vector<int> cs(nsph), nc(nsph);
    std::vector<double> coef((size_t) nsph * cmax, 0.0), T(15 * 9);
    for (int P = 0, q = 0; P < a.nshells; ++P) {
        const int l = a.am[P], ncl = (l + 1) * (l + 2) / 2, ns = 2 * l + 1;
        YG_TRY(yg_sph_coefficients(l, T.data()));
        for (int s = 0; s < ns; ++s, ++q) {
            cs[q] = a.bf_start[P];
            nc[q] = ncl;
            for (int i = 0; i < ncl; ++i) coef[(size_t) q * cmax + i] = T[i * ns + s];
        }
    }
    YG_TRY(c->d_sph_cs.ensure(std::max(1, nsph) * sizeof(int)));
    YG_TRY(c->d_sph_nc.ensure(std::max(1, nsph) * sizeof(int)));
    YG_TRY(c->d_sph_coef.ensure(std::max<size_t>(1, coef.size()) * sizeof(double)));
    YG_HIP(hipMemcpyAsync(c->d_sph_cs.p, cs.data(), nsph * sizeof(int), hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(c->d_sph_nc.p, nc.data(), nsph * sizeof(int), hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(c->d_sph_coef.p, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->sph_cmax = cmax;
    c->naux = nsph;
    return YGPU_OK;
}

int yg_aux_sph_cols(ygpu_ctx *c, hipStream_t s, long rows, const double *in, long ldin, double *out, long ldout)
{
    if (rows <= 0 || ldout <= 0) return YGPU_OK;
    const long total = rows * ldout;
    const long nb = std::min<long>((total + 255) / 256, 1L << 16);
    hipLaunchKernelGGL(k_aux_sph_cols, dim3((unsigned) nb), dim3(256), 0, s, rows, in, ldin, out, ldout, c->naux,
                       c->d_sph_cs.as<int>(), c->d_sph_nc.as<int>(), c->d_sph_coef.as<double>(), c->sph_cmax);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_shift_tail_diag(hipStream_t s, int n, int n_orth, double *A, double shift)
{
    if (n_orth <= 0 || n_orth >= n) return YGPU_OK;
    const int cnt = n - n_orth;
    hipLaunchKernelGGL(k_shift_tail_diag, dim3((cnt + 255) / 256), dim3(256), 0, s, n, n_orth, A, shift);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}
