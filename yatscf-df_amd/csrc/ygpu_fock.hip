// ygpu_fock.hip -- bandwidth-bound J build and the small elementwise kernels.
//
// J (build_Fock.c:37-130):
//   pass 1 "aux. J tensor": temp_J[p] = sum_r w_r B[r][p] over the packed rows
//     r = (i <= j), w_r = D_ii on the diagonal and 2 D_ij off it; a grid of
//     (column block) x (contiguous row range) workgroups streams each row once,
//     four wavefronts on adjacent column tiles of the same rows, and leaves one
//     partial per row range, summed in a fixed order (deterministic; replaces
//     the per-thread buffers and tree reduction of :15-33, :61-98);
//   pass 2 "J matrix": J_ij = B[r] . temp_J with temp_J staged in LDS, four
//     rows in flight per wavefront (each temp_J value read once for the four),
//     cross-lane reductions with shuffles.
// Every lane moves 16 bytes of a row per load in both passes.
// Screened pairs keep J = 0 (quirk A7 of the reference made explicit).
// Both passes read the tensor as FP64 or, in the mixed-precision phase, as its
// FP32 copy (half the bytes); the arithmetic is FP64 either way.
#include "ygpu_internal.h"

#include <algorithm>

namespace {

constexpr int JWAVES = 4;      // wavefronts per pass-1 workgroup, each on its own rows

// columns per lane: every lane moves 16 bytes per row (two FP64 or four FP32 values); rows are
// ld values apart with ld a multiple of 64, so these loads are aligned
template <class TB> constexpr int cpl() { return (int) (16 / sizeof(TB)); }

template <class TB>
__device__ __forceinline__ void load_cols(const TB *p, double (&v)[cpl<TB>()])
{
    if constexpr (sizeof(TB) == 8) {
        const double2 f = *(const double2 *) p;
        v[0] = f.x;
        v[1] = f.y;
    } else {
        const float4 f = *(const float4 *) p;
        v[0] = f.x;
        v[1] = f.y;
        v[2] = f.z;
        v[3] = f.w;
    }
}

// w_r = D_ii on the diagonal pairs and 2 D_ij off it, for the packed rows r = (i <= j)
__global__ void k_pair_weights(long n_up, const int *__restrict__ up_i, const int *__restrict__ up_j,
                               const double *__restrict__ D, int nbf, double *__restrict__ wts)
{
    const long r = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_up) return;
    const int i = up_i[r], j = up_j[r];
    wts[r] = (i == j ? 1.0 : 2.0) * D[(size_t) i * nbf + j];
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int) b, l), hi = __builtin_amdgcn_readlane((int) (b >> 32), l);
    return __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}

// pass 1: partial[g][c] = sum of w_r B[r][c] over the contiguous rows r of row group g,
// [g rpg, min(n_up, (g + 1) rpg)); grid (column blocks, row groups).  The JWAVES wavefronts of a
// workgroup take adjacent column tiles of 64 lanes x cpl columns of the same rows, so a
// workgroup reads JWAVES KiB of each row side by side, and the column blocks of one row group,
// consecutive in the grid, cover its rows whole and walk them in order.  The weights of 64 rows
// arrive with one load per wavefront and each row takes its own from the lane that holds it, so
// the row loop carries no dependent load; four rows in flight per wavefront.
// ld is a multiple of 64, so a lane's cpl columns are all inside or all outside.
template <class TB>
__global__ __launch_bounds__(64 * JWAVES) void k_tempj_part(const TB *__restrict__ B, long ld, long n_up, long rpg,
                                                            const double *__restrict__ wts, double *__restrict__ part)
{
    constexpr int C = cpl<TB>();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long col0 = ((long) blockIdx.x * JWAVES + w) * 64 * C;
    if (col0 >= ld) return;                     // a whole wavefront beyond the last column
    const long col = col0 + (long) lane * C;
    const bool cin = col < ld;
    const long r0 = (long) blockIdx.y * rpg, r1 = std::min(n_up, r0 + rpg);
    double a[4][C];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) a[q][c] = 0.0;
    for (long rb = r0; rb < r1; rb += 64) {
        const double wv = rb + lane < r1 ? wts[rb + lane] : 0.0;
        const int nr = (int) std::min<long>(64, r1 - rb);
        const TB *Bb = B + (size_t) rb * ld + col;
        int q0 = 0;
        for (; q0 + 3 < nr; q0 += 4) {
            double bv[4][C];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (cin) load_cols<TB>(Bb + (size_t) (q0 + q) * ld, bv[q]);
                else
#pragma unroll
                    for (int c = 0; c < C; ++c) bv[q][c] = 0.0;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double wt = readlane_f64(wv, q0 + q);
#pragma unroll
                for (int c = 0; c < C; ++c) a[q][c] += wt * bv[q][c];
            }
        }
        for (; q0 < nr; ++q0) {
            double bv[C];
            if (cin) load_cols<TB>(Bb + (size_t) q0 * ld, bv);
            else
#pragma unroll
                for (int c = 0; c < C; ++c) bv[c] = 0.0;
            const double wt = readlane_f64(wv, q0);
#pragma unroll
            for (int c = 0; c < C; ++c) a[0][c] += wt * bv[c];
        }
    }
    if (cin)
#pragma unroll
        for (int c = 0; c < C; ++c) part[(size_t) blockIdx.y * ld + col + c] = (a[0][c] + a[1][c]) + (a[2][c] + a[3][c]);
}

// temp_J[c] = sum over the row groups of partial[g][c] in a fixed order (deterministic): a
// workgroup takes TJR_COLS columns, TJR_SEG threads per column each add a contiguous range of
// the groups in order, and the TJR_SEG sums are added in order
constexpr int TJR_COLS = 32, TJR_SEG = 8;
__global__ __launch_bounds__(TJR_COLS * TJR_SEG) void k_tempj_reduce(const double *__restrict__ part, int ngroup,
                                                                     long ld, double *__restrict__ tempJ)
{
    __shared__ double red[TJR_SEG][TJR_COLS];
    const int cl = threadIdx.x % TJR_COLS, sg = threadIdx.x / TJR_COLS;
    const long col = (long) blockIdx.x * TJR_COLS + cl;
    const int per = (ngroup + TJR_SEG - 1) / TJR_SEG, g0 = sg * per, g1 = min(ngroup, g0 + per);
    double s = 0.0;
    if (col < ld)
        for (int g = g0; g < g1; ++g) s += part[(size_t) g * ld + col];
    red[sg][cl] = s;
    __syncthreads();
    if (sg == 0 && col < ld) {
        double t = red[0][cl];
#pragma unroll
        for (int q = 1; q < TJR_SEG; ++q) t += red[q][cl];
        tempJ[col] = t;
    }
}

constexpr int JM_ROWS = 4;      // pass-2 rows in flight per wavefront
constexpr int JM_THREADS = 1024;

// pass 2: wavefront g of the grid-stride sequence takes rows JM_ROWS g .. JM_ROWS g + 3 (the
// last group's missing rows read row n_up - 1 again and are not stored)
template <class TB>
__global__ __launch_bounds__(JM_THREADS) void k_jmat(const TB *__restrict__ B, long ld, long n_up,
                                                     const int *__restrict__ up_i, const int *__restrict__ up_j,
                                                     const double *__restrict__ tempJ, int nbf,
                                                     double *__restrict__ J)
{
    extern __shared__ double tj[];
    for (long c = threadIdx.x; c < ld; c += blockDim.x) tj[c] = tempJ[c];
    __syncthreads();
    constexpr int C = cpl<TB>();
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (long g = (long) blockIdx.x * nw + (threadIdx.x >> 6); g * JM_ROWS < n_up; g += (long) gridDim.x * nw) {
        const TB *row[JM_ROWS];
#pragma unroll
        for (int q = 0; q < JM_ROWS; ++q) row[q] = B + (size_t) std::min<long>(g * JM_ROWS + q, n_up - 1) * ld;
        double s[JM_ROWS];
#pragma unroll
        for (int q = 0; q < JM_ROWS; ++q) s[q] = 0.0;
        for (long c = (long) lane * C; c < ld; c += 64 * C) {
            double v[JM_ROWS][C], t[C];
#pragma unroll
            for (int q = 0; q < JM_ROWS; ++q) load_cols<TB>(row[q] + c, v[q]);
#pragma unroll
            for (int e = 0; e < C; ++e) t[e] = tj[c + e];
#pragma unroll
            for (int q = 0; q < JM_ROWS; ++q)
#pragma unroll
                for (int e = 0; e < C; ++e) s[q] += v[q][e] * t[e];
        }
#pragma unroll
        for (int q = 0; q < JM_ROWS; ++q) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off);
            const long r = g * JM_ROWS + q;
            if (lane == 0 && r < n_up) J[(size_t) up_i[r] * nbf + up_j[r]] = s[q];
        }
    }
}

// F = H + 2 J - K from the upper triangles of J and K; J is mirrored in place
__global__ void k_fock(int nbf, const double *__restrict__ H, double *__restrict__ J, const double *__restrict__ K,
                       double *__restrict__ F)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) nbf * nbf) return;
    const int i = (int)(idx / nbf), j = (int)(idx % nbf);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const double jv = J[(size_t) lo * nbf + hi];
    const double kv = K[(size_t) lo * nbf + hi];
    F[idx] = H[idx] + 2.0 * jv - kv;
    if (i > j) J[idx] = jv;
}

// Energy terms of a Fock build, block partials (fixed grid, fixed order): part[2b] = sum D (H + F),
// part[2b + 1] = sum D (F_r - H) - D_r (F - H) when a reference (F_r, D_r) is given, each product
// pair formed per element so that the near-cancelling terms cancel before they are summed
constexpr int EN_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_energy_terms(long nn, const double *__restrict__ D,
                                                      const double *__restrict__ H, const double *__restrict__ F,
                                                      const double *__restrict__ Dr, const double *__restrict__ Fr,
                                                      double *__restrict__ part)
{
    __shared__ double s1[4], s2[4];
    double a = 0.0, b = 0.0;
    for (long e = (long) blockIdx.x * 256 + threadIdx.x; e < nn; e += (long) EN_BLOCKS * 256) {
        const double d = D[e], h = H[e], f = F[e];
        a += d * (h + f);
        if (Dr) b += d * (Fr[e] - h) - Dr[e] * (f - h);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s1[w] = a;
        s2[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
        part[2 * blockIdx.x + 1] = (s2[0] + s2[1]) + (s2[2] + s2[3]);
    }
}

// mode 0: A[r][c] *= w[r]; mode 1: A[r][c] *= 1/sqrt(w[r])
__global__ void k_scale_rows(int n, int ncol, double *__restrict__ A, int lda, const double *__restrict__ w, int mode)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) n * ncol) return;
    const int r = (int)(idx / ncol), c = (int)(idx % ncol);
    const double s = mode == 0 ? w[r] : 1.0 / sqrt(w[r]);
    A[(size_t) r * lda + c] *= s;
}

__global__ void k_gather_cols(int nbf, int n_occ, const double *__restrict__ C, const int *__restrict__ idx,
                              double *__restrict__ Cocc)
{
    const long e = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long) nbf * n_occ) return;
    const int i = (int)(e / n_occ), s = (int)(e % n_occ);
    Cocc[e] = C[(size_t) i * nbf + (idx ? idx[s] : s)];
}

__global__ void k_transpose(int rows, int cols, const double *__restrict__ A, double *__restrict__ AT)
{
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = by + k, c = bx + tx;
        if (r < rows && c < cols) tile[k][tx] = A[(size_t) r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int r = bx + k, c = by + tx;   // AT is cols x rows
        if (r < cols && c < rows) AT[(size_t) r * rows + c] = tile[tx][k];
    }
}

__global__ void k_antisym(int n, const double *__restrict__ A, double *__restrict__ out)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    out[idx] = A[idx] - A[(size_t) j * n + i];
}

// A = (A + A^T) / 2 on the n x n block (one thread per upper element)
__global__ void k_symmetrize(int n, double *__restrict__ A, long lda)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) n * n) return;
    const int i = (int) (idx / n), j = (int) (idx % n);
    if (i >= j) return;
    const double v = 0.5 * (A[(size_t) i * lda + j] + A[(size_t) j * lda + i]);
    A[(size_t) i * lda + j] = v;
    A[(size_t) j * lda + i] = v;
}

// difference density of the incremental J: dD = D - Dprev, then Dprev = D
__global__ void k_delta(long n, const double *__restrict__ D, double *__restrict__ Dprev, double *__restrict__ dD)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const double d = D[idx];
    dD[idx] = d - Dprev[idx];
    Dprev[idx] = d;
}

// J += Jprev, then Jprev = J
__global__ void k_accum(long n, double *__restrict__ J, double *__restrict__ Jprev)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const double v = J[idx] + Jprev[idx];
    J[idx] = v;
    Jprev[idx] = v;
}

// out[0] = max |x|, one workgroup
__global__ __launch_bounds__(1024) void k_maxabs(long n, const double *__restrict__ x, double *__restrict__ out)
{
    __shared__ double red[16];
    double m = 0.0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(x[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int) (blockDim.x >> 6); ++w) m = fmax(m, red[w]);
        out[0] = m;
    }
}

// part[j][chunk] = sum over the chunk of x[k] * y_j[k]; with pmax, also pmax[chunk] = max over
// the chunk of |x[k]| (from the j = 0 workgroups, which read x anyway)
__global__ __launch_bounds__(256) void k_dots(long len, const double *__restrict__ x, const double *__restrict__ ys,
                                              long ystride, double *__restrict__ part, double *__restrict__ pmax)
{
    __shared__ double red[4], redm[4];
    const int j = blockIdx.y, nchunk = gridDim.x;
    const long per = (len + nchunk - 1) / nchunk;
    const long b = (long) blockIdx.x * per, e = b + per < len ? b + per : len;
    const double *y = ys + (size_t) j * ystride;
    const bool wantmax = pmax != nullptr && j == 0;   // uniform over the workgroup
    double s = 0.0, m = 0.0;
    for (long k = b + threadIdx.x; k < e; k += blockDim.x) {
        const double xv = x[k];
        s += xv * y[k];
        if (wantmax) m = fmax(m, fabs(xv));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        m = fmax(m, __shfl_xor(m, off));
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = s;
        redm[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[(size_t) j * nchunk + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        if (wantmax) pmax[blockIdx.x] = fmax(fmax(redm[0], redm[1]), fmax(redm[2], redm[3]));
    }
}

// x[0 .. n) = v (grid-stride)
__global__ void k_fill(long n, double *__restrict__ x, double v)
{
    for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x) x[i] = v;
}

// T = 3 I - T (n x n): the Newton-Schulz step Q <- Q (3 I - Q^T Q) / 2 of the orthogonal polar factor
__global__ void k_ns_shift(int n, double *__restrict__ T)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) n * n) return;
    const int i = (int) (idx / n), j = (int) (idx % n);
    T[idx] = (i == j ? 3.0 : 0.0) - T[idx];
}

// max |T - I| (n x n), one workgroup
__global__ __launch_bounds__(1024) void k_dev_identity(int n, const double *__restrict__ T, double *__restrict__ out)
{
    __shared__ double red[16];
    double m = 0.0;
    for (long e = threadIdx.x; e < (long) n * n; e += blockDim.x) {
        const int i = (int) (e / n), j = (int) (e % n);
        m = fmax(m, fabs(T[e] - (i == j ? 1.0 : 0.0)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int) (blockDim.x >> 6); ++w) m = fmax(m, red[w]);
        out[0] = m;
    }
}

// PM[i][s] = C1[i][s] + C0[i][s], PM[i][n_occ + s] = C1[i][s] - C0[i][s], in FP32 (nbf x 2 n_occ)
__global__ void k_pm32(int nbf, int n_occ, const double *__restrict__ C1, const double *__restrict__ C0,
                       float *__restrict__ PM)
{
    const long e = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long) nbf * n_occ) return;
    const int i = (int) (e / n_occ), s = (int) (e % n_occ);
    const double a = C1[e], b = C0[e];
    PM[(size_t) i * 2 * n_occ + s] = (float) (a + b);
    PM[(size_t) i * 2 * n_occ + n_occ + s] = (float) (a - b);
}

struct Coefs {
    double c[YGPU_MAX_DIIS];
};

__global__ void k_lincomb(long len, const double *__restrict__ ys, long ystride, int ny, Coefs cf,
                          double *__restrict__ out)
{
    const long k = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= len) return;
    double s = 0.0;
    for (int j = 0; j < ny; ++j) s += cf.c[j] * ys[(size_t) j * ystride + k];
    out[k] = s;
}

}  // namespace

int yg_antisym(hipStream_t s, int n, const double *A, double *out)
{
    const long tot = (long) n * n;
    hipLaunchKernelGGL(k_antisym, dim3((tot + 255) / 256), dim3(256), 0, s, n, A, out);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_symmetrize(hipStream_t s, int n, double *A, long lda)
{
    const long tot = (long) n * n;
    hipLaunchKernelGGL(k_symmetrize, dim3((tot + 255) / 256), dim3(256), 0, s, n, A, lda);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_delta(hipStream_t s, long n, const double *D, double *Dprev, double *dD)
{
    hipLaunchKernelGGL(k_delta, dim3((n + 255) / 256), dim3(256), 0, s, n, D, Dprev, dD);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_accum(hipStream_t s, long n, double *J, double *Jprev)
{
    hipLaunchKernelGGL(k_accum, dim3((n + 255) / 256), dim3(256), 0, s, n, J, Jprev);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_maxabs(hipStream_t s, long n, const double *x, double *out)
{
    hipLaunchKernelGGL(k_maxabs, dim3(1), dim3(1024), 0, s, n, x, out);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_dots(hipStream_t s, long len, const double *x, const double *ys, long ystride, int ny, double *part, int nchunk,
            double *pmax)
{
    if (ny <= 0) return YGPU_OK;
    hipLaunchKernelGGL(k_dots, dim3(nchunk, ny), dim3(256), 0, s, len, x, ys, ystride, part, pmax);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_lincomb(hipStream_t s, long len, const double *ys, long ystride, int ny, const double *coef, double *out)
{
    Coefs cf{};
    if (ny > YGPU_MAX_DIIS) {
        ygpu_set_error(YGPU_ERR_INVALID, "linear combination of %d > %d vectors", ny, YGPU_MAX_DIIS);
        return YGPU_ERR_INVALID;
    }
    for (int j = 0; j < ny; ++j) cf.c[j] = coef[j];
    hipLaunchKernelGGL(k_lincomb, dim3((len + 255) / 256), dim3(256), 0, s, len, ys, ystride, ny, cf, out);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

// column blocks of the pass-1 grid: JWAVES tiles of 64 lanes x cpl columns
static long tempj_colblocks(long naux_pad, bool fp32)
{
    const int C = fp32 ? cpl<float>() : cpl<double>();
    return (naux_pad + (long) JWAVES * 64 * C - 1) / ((long) JWAVES * 64 * C);
}

long yg_jbuild_groups(long n_up, long naux_pad, bool fp32)
{
    // row groups: about 2048 workgroups in total, at least 64 rows (one load of weights) per group
    const long ngroup = std::max<long>(1, std::min<long>(2048 / tempj_colblocks(naux_pad, fp32), n_up / 64));
    return std::min<long>(ngroup, 65535);
}

int yg_jbuild_ws(ygpu_ctx *c, hipStream_t s, const double *dD, double *dJ, const float *B32, double *part,
                 double *tempJ, double *wts, long ngroup)
{
    const long ld = c->naux_pad, n_up = c->n_up;
    const int nbf = c->nbf;
    const long ncb = tempj_colblocks(ld, B32 != nullptr), rpg = (n_up + ngroup - 1) / ngroup;
    YG_HIP(hipEventRecord(c->ev[0], s));
    hipLaunchKernelGGL(k_pair_weights, dim3((unsigned) ((n_up + 255) / 256)), dim3(256), 0, s, n_up,
                       c->d_up_i.as<int>(), c->d_up_j.as<int>(), dD, nbf, wts);
    YG_HIP(hipGetLastError());
    if (B32)
        hipLaunchKernelGGL(k_tempj_part<float>, dim3((unsigned) ncb, (unsigned) ngroup), dim3(64 * JWAVES), 0, s, B32,
                           ld, n_up, rpg, wts, part);
    else
        hipLaunchKernelGGL(k_tempj_part<double>, dim3((unsigned) ncb, (unsigned) ngroup), dim3(64 * JWAVES), 0, s,
                           c->d_B.as<double>(), ld, n_up, rpg, wts, part);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tempj_reduce, dim3((unsigned) ((ld + TJR_COLS - 1) / TJR_COLS)), dim3(TJR_COLS * TJR_SEG), 0, s,
                       part, (int) ngroup, ld, tempJ);
    YG_HIP(hipGetLastError());
    YG_HIP(hipEventRecord(c->ev[1], s));
    // J entries of screened-out pairs stay zero (a fill kernel, not a memset node)
    YG_TRY(yg_fill(s, (long) nbf * nbf, dJ, 0.0));
    // two workgroups of 16 wavefronts per CU (temp_J in LDS: ld doubles each), grid-stride rows
    const long ngroups = (n_up + JM_ROWS - 1) / JM_ROWS, wpb = JM_THREADS / 64;
    const long nblk = std::max<long>(1, std::min<long>(512, (ngroups + wpb - 1) / wpb));
    if (B32)
        hipLaunchKernelGGL(k_jmat<float>, dim3(nblk), dim3(JM_THREADS), ld * sizeof(double), s, B32, ld, n_up,
                           c->d_up_i.as<int>(), c->d_up_j.as<int>(), tempJ, nbf, dJ);
    else
        hipLaunchKernelGGL(k_jmat<double>, dim3(nblk), dim3(JM_THREADS), ld * sizeof(double), s,
                           c->d_B.as<double>(), ld, n_up, c->d_up_i.as<int>(), c->d_up_j.as<int>(), tempJ, nbf, dJ);
    YG_HIP(hipGetLastError());
    YG_HIP(hipEventRecord(c->ev[2], s));
    return YGPU_OK;
}

int yg_jbuild(ygpu_ctx *c, hipStream_t s, const double *dD, double *dJ, const float *B32, float *t_tempJ_ms,
              float *t_Jmat_ms)
{
    const long ld = c->naux_pad;
    const long ngroup = yg_jbuild_groups(c->n_up, ld, B32 != nullptr);
    YG_TRY(c->d_tempJpart.ensure((size_t) ngroup * ld * sizeof(double)));
    YG_TRY(c->d_tempJ.ensure((size_t) ld * sizeof(double)));
    YG_TRY(c->d_jw.ensure((size_t) c->n_up * sizeof(double)));
    YG_TRY(yg_jbuild_ws(c, s, dD, dJ, B32, c->d_tempJpart.as<double>(), c->d_tempJ.as<double>(), c->d_jw.as<double>(),
                        ngroup));
    if (t_tempJ_ms || t_Jmat_ms) {
        YG_HIP(hipEventSynchronize(c->ev[2]));
        if (t_tempJ_ms) YG_HIP(hipEventElapsedTime(t_tempJ_ms, c->ev[0], c->ev[1]));
        if (t_Jmat_ms) YG_HIP(hipEventElapsedTime(t_Jmat_ms, c->ev[1], c->ev[2]));
    }
    return YGPU_OK;
}

int yg_fill(hipStream_t s, long n, double *x, double v)
{
    if (n <= 0) return YGPU_OK;
    const long nb = std::min<long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, s, n, x, v);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_fock_assemble(hipStream_t s, int nbf, const double *H, double *J, const double *K, double *F)
{
    const long tot = (long) nbf * nbf;
    hipLaunchKernelGGL(k_fock, dim3((tot + 255) / 256), dim3(256), 0, s, nbf, H, J, K, F);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_energy_blocks() { return EN_BLOCKS; }

int yg_energy_terms(hipStream_t s, int nbf, const double *D, const double *H, const double *F, const double *Dr,
                    const double *Fr, double *part)
{
    hipLaunchKernelGGL(k_energy_terms, dim3(EN_BLOCKS), dim3(256), 0, s, (long) nbf * nbf, D, H, F, Dr, Fr, part);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_scale_rows(hipStream_t s, int n, int ncol, double *A, int lda, const double *w, int mode)
{
    const long tot = (long) n * ncol;
    if (tot == 0) return YGPU_OK;
    hipLaunchKernelGGL(k_scale_rows, dim3((tot + 255) / 256), dim3(256), 0, s, n, ncol, A, lda, w, mode);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_gather_cols(hipStream_t s, int nbf, int n_occ, const double *C, const int *idx, double *Cocc)
{
    const long tot = (long) nbf * n_occ;
    if (tot == 0) return YGPU_OK;
    hipLaunchKernelGGL(k_gather_cols, dim3((tot + 255) / 256), dim3(256), 0, s, nbf, n_occ, C, idx, Cocc);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_transpose(hipStream_t s, int rows, int cols, const double *A, double *AT)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, rows, cols, A, AT);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_ns_shift(hipStream_t s, int n, double *T)
{
    const long tot = (long) n * n;
    hipLaunchKernelGGL(k_ns_shift, dim3((tot + 255) / 256), dim3(256), 0, s, n, T);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_dev_identity(hipStream_t s, int n, const double *T, double *out)
{
    hipLaunchKernelGGL(k_dev_identity, dim3(1), dim3(1024), 0, s, n, T, out);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_pm32(hipStream_t s, int nbf, int n_occ, const double *C1, const double *C0, float *PM)
{
    const long tot = (long) nbf * n_occ;
    if (tot == 0) return YGPU_OK;
    hipLaunchKernelGGL(k_pm32, dim3((tot + 255) / 256), dim3(256), 0, s, nbf, n_occ, C1, C0, PM);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}
