// ygpu_purify.hip -- density matrix by purification of the orthogonal-basis Fock matrix.
//
// The reference forms D from the n_occ lowest eigenvectors of F' = X^T F X (dsyev,
// build_density.c:37-69).  D only depends on the spectral projector P onto those
// eigenvectors, which the second-order spectral projection (SP2, Niklasson, Phys. Rev. B 66,
// 155115 (2002)) reaches with matrix squares alone:
//   X_0 = (e_max I - F') / (e_max - e_min)       (e_min, e_max: Gershgorin bounds of F')
//   X_k = X_{k-1}^2        if |tr(X^2) - N| <= |2 tr(X) - tr(X^2) - N|
//         2 X_{k-1} - X_{k-1}^2  otherwise,         N = n_occ,
// every eigenvalue of X_k moving to 1 (occupied) or 0 (virtual) quadratically once the
// iteration has separated them.  tr(X^2) = |X|_F^2 for the symmetric X, so both traces come
// from the elementwise pass that writes X; e_k = tr(X_k) - tr(X_k^2) = sum lambda (1 - lambda)
// measures the distance from idempotency.  Each step is one FP64 MFMA SYRK (the upper tiles of
// X X^T, split over k into partials) and one pass that adds the partials in a fixed order and
// applies the choice, so the result does not depend on scheduling.
//
// Rows and columns at or beyond n_orth (the zero directions of a spherical basis, whose
// diagonal the density build shifts far above the spectrum) are set to zero in X_0 and stay
// zero; X is stored n x ldx with zero padding columns, ldx a multiple of 16 (the SYRK's k step).
#include "ygpu_internal.h"

#include <algorithm>
#include <cfloat>

namespace {

constexpr int SP2_RPB = 1;         // row pairs per workgroup of the initial pass
constexpr int UT = 64;             // tile of the update pass

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// sum of the nb block partials in a fixed order (every workgroup forms the same value);
// the 256 threads each add a strided slice, then the wavefront and block sums
__device__ double block_total(const double *__restrict__ part, int nb, double *sred)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[b];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sred[w] = v;
    __syncthreads();
    return (sred[0] + sred[1]) + (sred[2] + sred[3]);
}

// Gershgorin interval of rows < n_orth: one wavefront per row, block partials of
// min(F_ii - r_i) and max(F_ii + r_i), r_i = sum over j != i, j < n_orth of |F_ij|
__global__ __launch_bounds__(256) void k_sp2_bounds(int n_orth, long ldf, const double *__restrict__ F,
                                                    double *__restrict__ plo, double *__restrict__ phi)
{
    __shared__ double slo[4], shi[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + w;
    double lo = DBL_MAX, hi = -DBL_MAX;
    if (i < n_orth) {
        const double *row = F + (size_t) i * ldf;
        double r = 0.0;
        for (int j = lane; j < n_orth; j += 64) r += (j == i) ? 0.0 : fabs(row[j]);
        r = wave_sum(r);
        lo = row[i] - r;
        hi = row[i] + r;
    }
    if (lane == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        plo[blockIdx.x] = fmin(fmin(slo[0], slo[1]), fmin(slo[2], slo[3]));
        phi[blockIdx.x] = fmax(fmax(shi[0], shi[1]), fmax(shi[2], shi[3]));
    }
}

// rows of workgroup b: the pairs (p, n - 1 - p) for p in [b R, b R + R), so that every
// workgroup covers about R (n + 1) elements of the upper triangle
__device__ __forceinline__ int pair_row(int n, int b, int q)
{
    const int p = b * SP2_RPB + (q >> 1);
    if (p >= (n + 1) / 2) return -1;
    const int r = (q & 1) ? n - 1 - p : p;
    return ((q & 1) && r == p) ? -1 : r;
}

// X_0 = (e_max I - F) / (e_max - e_min) on rows / columns < n_orth, zero elsewhere; block
// partials of tr(X_0) and |X_0|_F^2; bounds[0..1] = e_min, e_max
__global__ __launch_bounds__(256) void k_sp2_init(int n, int n_orth, long ldf, const double *__restrict__ F,
                                                  int nbb, const double *__restrict__ plo,
                                                  const double *__restrict__ phi, long ldx, double *__restrict__ X,
                                                  double *__restrict__ ptr, double *__restrict__ psq,
                                                  double *__restrict__ bounds)
{
    __shared__ double sred[4];
    __shared__ double sb[2];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
        double lo = DBL_MAX, hi = -DBL_MAX;
        for (int b = lane; b < nbb; b += 64) {
            lo = fmin(lo, plo[b]);
            hi = fmax(hi, phi[b]);
        }
        lo = wave_min(lo);
        hi = wave_max(hi);
        if (lane == 0) {
            sb[0] = lo;
            sb[1] = hi;
        }
    }
    __syncthreads();
    const double emin = sb[0], emax = sb[1];
    const double scale = 1.0 / fmax(emax - emin, 1e-300);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bounds[0] = emin;
        bounds[1] = emax;
    }
    double tr = 0.0, sq = 0.0;
    for (int q = 0; q < 2 * SP2_RPB; ++q) {
        const int i = pair_row(n, blockIdx.x, q);
        if (i < 0) continue;
        const double *frow = F + (size_t) i * ldf;
        double *xrow = X + (size_t) i * ldx;
        for (int j = threadIdx.x; j < ldx; j += 256) {
            double v = 0.0;
            if (i < n_orth && j < n_orth) v = ((i == j ? emax : 0.0) - frow[j]) * scale;
            xrow[j] = v;
            if (j < n) {
                sq += v * v;
                if (j == i) tr += v;
            }
        }
    }
    tr = wave_sum(tr);
    sq = wave_sum(sq);
    __syncthreads();
    if (lane == 0) sred[w] = tr;
    __syncthreads();
    const double btr = (sred[0] + sred[1]) + (sred[2] + sred[3]);
    __syncthreads();
    if (lane == 0) sred[w] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        ptr[blockIdx.x] = btr;
        psq[blockIdx.x] = (sred[0] + sred[1]) + (sred[2] + sred[3]);
    }
}

// One SP2 step after the SYRK of X, on the upper 64 x 64 tiles (ti <= tj): S = sum_z part[z] (the
// SYRK's upper tiles hold every element of these), X <- S or 2X - S in place (each element of an
// upper tile is read and written by one thread; an off-diagonal tile's transpose, staged in LDS,
// is written row by row into the lower tile, which no workgroup reads); block partials of the new
// tr(X) and |X|_F^2.  elog[step] = tr(X) - tr(X^2) of the X before the step, slog[step] = 1 for a
// square, 0 for 2X - X^2
__global__ __launch_bounds__(256) void k_sp2_update(int n, long ldx, double *__restrict__ X,
                                                    const double *__restrict__ part, int nsplit, int N, int nt,
                                                    int nb_in, const double *__restrict__ ptr_in,
                                                    const double *__restrict__ psq_in, double *__restrict__ ptr_out,
                                                    double *__restrict__ psq_out, double *__restrict__ elog,
                                                    int *__restrict__ slog, int step)
{
    __shared__ double T[UT][UT + 1];
    __shared__ double sred[4];
    const double tX = block_total(ptr_in, nb_in, sred);
    __syncthreads();
    const double tX2 = block_total(psq_in, nb_in, sred);
    const bool square = fabs(tX2 - N) <= fabs(2.0 * tX - tX2 - N);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        elog[step] = tX - tX2;
        slog[step] = square ? 1 : 0;
    }
    int t = blockIdx.x, ti = 0;
    while (t >= nt - ti) {
        t -= nt - ti;
        ++ti;
    }
    const int tj = ti + t;
    const int i0 = ti * UT, j0 = tj * UT;
    const int c = threadIdx.x & 63, rq = threadIdx.x >> 6;
    const size_t nn = (size_t) n * n;
    double tr = 0.0, sq = 0.0;
    for (int r = rq; r < UT; r += 4) {
        const int i = i0 + r, j = j0 + c;
        double v = 0.0;
        if (i < n && j < n) {
            const double *pz = part + (size_t) i * n + j;
            double sz = 0.0;
            for (int z = 0; z < nsplit; ++z) sz += pz[(size_t) z * nn];
            double *x = X + (size_t) i * ldx + j;
            v = square ? sz : 2.0 * *x - sz;
            *x = v;
            if (ti == tj) {
                sq += v * v;
                if (i == j) tr += v;
            } else {
                sq += 2.0 * v * v;
            }
        }
        T[c][r] = v;
    }
    if (ti != tj) {
        __syncthreads();
        for (int r = rq; r < UT; r += 4) {
            const int i = j0 + r, j = i0 + c;
            if (i < n && j < n) X[(size_t) i * ldx + j] = T[r][c];
        }
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    tr = wave_sum(tr);
    sq = wave_sum(sq);
    __syncthreads();
    if (lane == 0) sred[w] = tr;
    __syncthreads();
    const double btr = (sred[0] + sred[1]) + (sred[2] + sred[3]);
    __syncthreads();
    if (lane == 0) sred[w] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        ptr_out[blockIdx.x] = btr;
        psq_out[blockIdx.x] = (sred[0] + sred[1]) + (sred[2] + sred[3]);
    }
}

// out[0] = min over i < n of L[i][i]^2 (one workgroup)
__global__ __launch_bounds__(256) void k_min_diag_sq(int n, long ld, const double *__restrict__ L,
                                                     double *__restrict__ out)
{
    __shared__ double sred[4];
    double v = DBL_MAX;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double d = L[(size_t) i * ld + i];
        v = fmin(v, d * d);
    }
    v = wave_min(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sred[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmin(fmin(sred[0], sred[1]), fmin(sred[2], sred[3]));
}

}  // namespace

// workgroups of the initial pass (row pairs) and of the update pass (upper 64 x 64 tiles)
static int sp2_init_blocks(int n) { return std::max(1, ((n + 1) / 2 + SP2_RPB - 1) / SP2_RPB); }
static int sp2_tiles(int n)
{
    const int nt = (n + UT - 1) / UT;
    return nt * (nt + 1) / 2;
}
static int sp2_part_len(int n) { return std::max(sp2_init_blocks(n), sp2_tiles(n)); }

// layout of the summation workspace: Gershgorin partials (2 x (n + 3) / 4), then two buffers of
// trace and norm partials (2 x 2 x sp2_part_len(n))
size_t yg_sp2_wsum_doubles(int n) { return 2 * (size_t) ((n + 3) / 4) + 4 * (size_t) sp2_part_len(n); }

static double *sp2_tr_part(double *wsum, int n, int buf)
{
    return wsum + 2 * ((n + 3) / 4) + (size_t) buf * 2 * sp2_part_len(n);
}

int yg_sp2_begin(hipStream_t s, int n, int n_orth, const double *F, long ldf, double *X, long ldx, double *wsum,
                 double *bounds)
{
    if (n <= 0 || ldx < n || ldx % 16) {
        ygpu_set_error(YGPU_ERR_INVALID, "SP2: order %d with leading dimension %ld (a multiple of 16 >= n)", n, ldx);
        return YGPU_ERR_INVALID;
    }
    const int no = (n_orth > 0 && n_orth < n) ? n_orth : n;
    const int nbb = (no + 3) / 4, nb = sp2_init_blocks(n), len = sp2_part_len(n);
    double *plo = wsum, *phi = wsum + nbb;
    double *ptr = sp2_tr_part(wsum, n, 0), *psq = ptr + len;
    hipLaunchKernelGGL(k_sp2_bounds, dim3(nbb), dim3(256), 0, s, no, ldf, F, plo, phi);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sp2_init, dim3(nb), dim3(256), 0, s, n, no, ldf, F, nbb, plo, phi, ldx, X, ptr, psq, bounds);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_sp2_steps(hipStream_t s, int n, long ldx, double *X, int N, double *part, long kchunk, double *wsum,
                 int first, int count, double *elog, int *slog)
{
    const int nt = (n + UT - 1) / UT, ntiles = sp2_tiles(n), len = sp2_part_len(n);
    double *ptr[2] = {sp2_tr_part(wsum, n, 0), sp2_tr_part(wsum, n, 1)};
    double *psq[2] = {ptr[0] + len, ptr[1] + len};
    const int nsplit = (int) ((ldx + kchunk - 1) / kchunk);
    for (int k = first; k < first + count; ++k) {
        YG_TRY(yg_syrk_upper_parts(s, n, ldx, X, part, kchunk));
        // the partials of X_k: buffer k % 2, written by the initial pass (k = 0) or step k - 1
        const int a = k & 1, b = a ^ 1;
        const int nb_in = k == 0 ? sp2_init_blocks(n) : ntiles;
        hipLaunchKernelGGL(k_sp2_update, dim3(ntiles), dim3(256), 0, s, n, ldx, X, part, nsplit, N, nt, nb_in, ptr[a],
                           psq[a], ptr[b], psq[b], elog, slog, k);
        YG_HIP(hipGetLastError());
    }
    return YGPU_OK;
}

int yg_min_diag_sq(hipStream_t s, int n, long ld, const double *L, double *out)
{
    hipLaunchKernelGGL(k_min_diag_sq, dim3(1), dim3(256), 0, s, n, ld, L, out);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}
