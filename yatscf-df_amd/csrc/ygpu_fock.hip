// ygpu_fock.hip -- bandwidth-bound J build and the small elementwise kernels.
//
// J (build_Fock.c:37-130):
//   pass 1 "aux. J tensor": temp_J[p] = sum_r w_r B[r][p] over the packed rows
//     r = (i <= j), w_r = D_ii on the diagonal and 2 D_ij off it; each
//     workgroup streams a contiguous band of rows once (coalesced row reads,
//     one column per lane) into per-band partial sums, reduced in a fixed order
//     (deterministic; replaces the per-thread buffers + tree of :15-33, :61-98);
//   pass 2 "J matrix": J_ij = B[r] . temp_J with temp_J staged in LDS, one
//     wavefront per row, cross-lane reduction with shuffles.
// Screened pairs keep J = 0 (quirk A7 of the reference made explicit).
#include "ygpu_internal.h"

#include <algorithm>

namespace {

constexpr int JCOLS = 256;     // columns per pass-1 workgroup (one per lane)

__global__ __launch_bounds__(JCOLS) void k_tempj_part(const double *__restrict__ B, long ld, long n_up,
                                                      const int *__restrict__ up_i, const int *__restrict__ up_j,
                                                      const double *__restrict__ D, int nbf, long rows_per_band,
                                                      double *__restrict__ part)
{
    const long col = (long) blockIdx.x * JCOLS + threadIdx.x;
    const long r0 = (long) blockIdx.y * rows_per_band;
    const long r1 = r0 + rows_per_band < n_up ? r0 + rows_per_band : n_up;
    double acc0 = 0.0, acc1 = 0.0;
    if (col < ld) {
        long r = r0;
        for (; r + 1 < r1; r += 2) {
            const int i0 = up_i[r], j0 = up_j[r], i1 = up_i[r + 1], j1 = up_j[r + 1];
            const double w0 = (i0 == j0 ? 1.0 : 2.0) * D[(size_t) i0 * nbf + j0];
            const double w1 = (i1 == j1 ? 1.0 : 2.0) * D[(size_t) i1 * nbf + j1];
            acc0 += w0 * B[(size_t) r * ld + col];
            acc1 += w1 * B[(size_t)(r + 1) * ld + col];
        }
        if (r < r1) {
            const int i0 = up_i[r], j0 = up_j[r];
            acc0 += (i0 == j0 ? 1.0 : 2.0) * D[(size_t) i0 * nbf + j0] * B[(size_t) r * ld + col];
        }
        part[(size_t) blockIdx.y * ld + col] = acc0 + acc1;
    }
}

__global__ void k_tempj_reduce(const double *__restrict__ part, int nband, long ld, double *__restrict__ tempJ)
{
    const long col = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ld) return;
    double s = 0.0;
    for (int b = 0; b < nband; ++b) s += part[(size_t) b * ld + col];
    tempJ[col] = s;
}

__global__ __launch_bounds__(256) void k_jmat(const double *__restrict__ B, long ld, long n_up,
                                              const int *__restrict__ up_i, const int *__restrict__ up_j,
                                              const double *__restrict__ tempJ, int nbf, int rows_per_wave,
                                              double *__restrict__ J)
{
    extern __shared__ double tj[];
    for (long c = threadIdx.x; c < ld; c += blockDim.x) tj[c] = tempJ[c];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long rbase = ((long) blockIdx.x * 4 + w) * rows_per_wave;
    for (int q = 0; q < rows_per_wave; ++q) {
        const long r = rbase + q;
        if (r >= n_up) break;
        const double *row = B + (size_t) r * ld;
        double s = 0.0;
        for (long c = lane; c < ld; c += 64) s += row[c] * tj[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) J[(size_t) up_i[r] * nbf + up_j[r]] = s;
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

}  // namespace

int yg_jbuild(ygpu_ctx *c, const double *dD, double *dJ, float *t_tempJ_ms, float *t_Jmat_ms)
{
    const long ld = c->naux_pad, n_up = c->n_up;
    const int nbf = c->nbf;
    const int ncolb = (int) ((ld + JCOLS - 1) / JCOLS);
    // bands: enough workgroups to cover the chip several times, >= 16 rows each
    long nband = std::max<long>(1, std::min<long>(4096 / ncolb, (n_up + 15) / 16));
    const long rpb = (n_up + nband - 1) / nband;
    nband = (n_up + rpb - 1) / rpb;
    YG_TRY(c->d_tempJpart.ensure((size_t) nband * ld * sizeof(double)));
    YG_TRY(c->d_tempJ.ensure((size_t) ld * sizeof(double)));
    YG_HIP(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_tempj_part, dim3(ncolb, nband), dim3(JCOLS), 0, c->stream, c->d_B.as<double>(), ld, n_up,
                       c->d_up_i.as<int>(), c->d_up_j.as<int>(), dD, nbf, rpb, c->d_tempJpart.as<double>());
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tempj_reduce, dim3((ld + 255) / 256), dim3(256), 0, c->stream, c->d_tempJpart.as<double>(),
                       (int) nband, ld, c->d_tempJ.as<double>());
    YG_HIP(hipGetLastError());
    YG_HIP(hipEventRecord(c->ev[1], c->stream));
    YG_HIP(hipMemsetAsync(dJ, 0, (size_t) nbf * nbf * sizeof(double), c->stream));
    const int rpw = 4;
    const long nblk = (n_up + 4L * rpw - 1) / (4L * rpw);
    hipLaunchKernelGGL(k_jmat, dim3(nblk), dim3(256), ld * sizeof(double), c->stream, c->d_B.as<double>(), ld, n_up,
                       c->d_up_i.as<int>(), c->d_up_j.as<int>(), c->d_tempJ.as<double>(), nbf, rpw, dJ);
    YG_HIP(hipGetLastError());
    YG_HIP(hipEventRecord(c->ev[2], c->stream));
    if (t_tempJ_ms || t_Jmat_ms) {
        YG_HIP(hipEventSynchronize(c->ev[2]));
        if (t_tempJ_ms) YG_HIP(hipEventElapsedTime(t_tempJ_ms, c->ev[0], c->ev[1]));
        if (t_Jmat_ms) YG_HIP(hipEventElapsedTime(t_Jmat_ms, c->ev[1], c->ev[2]));
    }
    return YGPU_OK;
}

int yg_fock_assemble(hipStream_t s, int nbf, const double *H, double *J, const double *K, double *F)
{
    const long tot = (long) nbf * nbf;
    hipLaunchKernelGGL(k_fock, dim3((tot + 255) / 256), dim3(256), 0, s, nbf, H, J, K, F);
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
