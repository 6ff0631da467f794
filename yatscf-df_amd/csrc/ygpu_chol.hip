// ygpu_chol.hip -- Cholesky factor of the Coulomb metric and its inverse.
//
// The reference fits with J^{-1/2} from an eigendecomposition of J_PQ
// (build_DF_tensor.c:205-237).  B B^T = (mn|P) J^{-1} (Q|kl) holds for any W
// with W^T W = J^{-1}, so the DF tensor can equally be B = (mn|P) W^T with
// W = L^{-1}, J = L L^T: the Coulomb and exchange matrices are the same, the
// factor costs n^3/3 flops instead of an eigensolve, and W is lower
// triangular, so the contraction needs half the flops of the full product.
//
// Blocked right-looking Cholesky with 64-wide blocks (LAPACK dpotrf's lower
// variant on the row-major matrix): one workgroup factors the diagonal block
// in LDS and inverts it; the panel below is solved by a product with that
// inverse and the trailing matrix updated, both on the FP64 MFMA GEMM of
// ygpu_gemm.hip.  W = L^{-1} then follows block row by block row:
//   W_ii = L_ii^{-1},  W[i, 0:i] = -W_ii (L[i, 0:i] W[0:i, 0:i]).
#include "ygpu_internal.h"

#include <algorithm>

namespace {

constexpr int NB = 64;

// Factor the NB x NB diagonal block at (k0, k0) of the row-major A (leading dimension lda) in
// place: its lower triangle becomes L_kk (strictly upper part zeroed), and Linv (NB x NB,
// row-major) receives L_kk^{-1}.  Rows / columns at or beyond n keep an identity block.  A
// pivot that is not positive sets *info to its global column + 1 (first failure kept) and is
// replaced by 1 so that the remaining work stays finite.
// One wavefront, lane i owning row i in registers (Cholesky-Crout by rows): at column j every
// lane forms a_ij - sum_{k<j} L_ik L_jk with row j of L read from LDS as a broadcast, lane j
// takes the square root and the lanes below divide by it.  Then lane c solves L x = e_c by
// forward substitution for column c of L^{-1}, again against broadcast rows of L.
__global__ __launch_bounds__(64) void k_potrf_diag(int n, long lda, double *__restrict__ A, int k0,
                                                   double *__restrict__ Linv, int *__restrict__ info)
{
    __shared__ double Ls[NB][NB + 1];
    const int i = threadIdx.x;
    // the block through LDS (row q read by all lanes, one column each), then row i to registers;
    // the factorisation overwrites Ls[j][k] (k <= j) with L before reading it
    for (int q = 0; q < NB; ++q) {
        double v = 0.0;
        if (k0 + q < n && k0 + i < n) v = (i <= q) ? A[(size_t) (k0 + q) * lda + k0 + i] : 0.0;
        else if (q == i) v = 1.0;
        Ls[q][i] = v;
    }
    __syncthreads();
    double r[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) r[j] = Ls[i][j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double s = r[j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= r[k] * Ls[j][k];
        if (i == j) {
            if (!(s > 0.0)) {
                if (k0 + j < n) atomicCAS(info, 0, k0 + j + 1);
                s = 1.0;
            }
            s = sqrt(s);
            r[j] = s;
            Ls[j][j] = s;
        }
        __syncthreads();
        if (i > j) {
            r[j] = s / Ls[j][j];
            Ls[i][j] = r[j];
        }
        __syncthreads();
    }
    // L back to A row by row (lane = column); Ls holds L on and below the diagonal
    for (int q = 0; q < NB; ++q)
        if (k0 + q < n && k0 + i < n) A[(size_t) (k0 + q) * lda + k0 + i] = (i <= q) ? Ls[q][i] : 0.0;
    // column c = i of L^{-1}: x_q = (delta_qc - sum_{p<q} L_qp x_p) / L_qq (x_p = 0 for p < c)
    double x[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        double s = (q == i) ? 1.0 : 0.0;
#pragma unroll
        for (int p = 0; p < q; ++p) s -= Ls[q][p] * x[p];
        x[q] = s / Ls[q][q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) Linv[q * NB + i] = x[q];
}

// W's diagonal blocks from the per-block inverses; every other element of rows [0, npad) and
// columns [0, npad) zero (W is lower triangular; padding rows and columns carry identity)
__global__ void k_trtri_init(int npad, long ldw, double *__restrict__ W, const double *__restrict__ Linv)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) npad * npad) return;
    const int i = (int) (idx / npad), j = (int) (idx % npad);
    double v = 0.0;
    if (i / NB == j / NB) {
        const int b = i / NB;
        v = Linv[(size_t) b * NB * NB + (i % NB) * NB + (j % NB)];
    }
    W[(size_t) i * ldw + j] = v;
}

}  // namespace

int yg_potrf_lower(hipStream_t s, int n, int npad, long lda, double *A, double *Linv, int *info, double *T)
{
    if (npad % NB || npad < n) {
        ygpu_set_error(YGPU_ERR_INVALID, "Cholesky: padded order %d not a multiple of %d or below %d", npad, NB, n);
        return YGPU_ERR_INVALID;
    }
    const int nblk = npad / NB;
    for (int b = 0; b < nblk; ++b) {
        const int k0 = b * NB;
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, n, lda, A, k0, Linv + (size_t) b * NB * NB, info);
        YG_HIP(hipGetLastError());
        const int rest = npad - k0 - NB;
        if (rest <= 0) break;
        double *panel = A + (size_t) (k0 + NB) * lda + k0;
        // L_21 = A_21 L_11^{-T} into T (rest x NB), then back into the panel's place in A
        YG_TRY(yg_dgemm(s, 0, 1, rest, NB, NB, 1.0, panel, (int) lda, Linv + (size_t) b * NB * NB, NB, 0.0, T, NB));
        YG_HIP(hipMemcpy2DAsync(panel, lda * sizeof(double), T, NB * sizeof(double), NB * sizeof(double), rest,
                                hipMemcpyDeviceToDevice, s));
        // trailing matrix A_22 -= L_21 L_21^T
        double *a22 = A + (size_t) (k0 + NB) * lda + k0 + NB;
        YG_TRY(yg_dgemm(s, 0, 1, rest, rest, NB, -1.0, T, NB, T, NB, 1.0, a22, (int) lda));
    }
    return YGPU_OK;
}

int yg_trtri_lower(hipStream_t s, int npad, long lda, const double *L, const double *Linv, double *W, long ldw,
                   double *T)
{
    if (npad % NB) {
        ygpu_set_error(YGPU_ERR_INVALID, "triangular inverse: order %d not a multiple of %d", npad, NB);
        return YGPU_ERR_INVALID;
    }
    const long tot = (long) npad * npad;
    hipLaunchKernelGGL(k_trtri_init, dim3((unsigned) ((tot + 255) / 256)), dim3(256), 0, s, npad, ldw, W, Linv);
    YG_HIP(hipGetLastError());
    const int nblk = npad / NB;
    for (int b = 1; b < nblk; ++b) {
        const int r0 = b * NB;
        // T = L[b, 0:b] W[0:b, 0:b]  (NB x r0), then W[b, 0:b] = -W_bb T
        YG_TRY(yg_dgemm(s, 0, 0, NB, r0, r0, 1.0, L + (size_t) r0 * lda, (int) lda, W, (int) ldw, 0.0, T, r0));
        YG_TRY(yg_dgemm(s, 0, 0, NB, r0, NB, -1.0, Linv + (size_t) b * NB * NB, NB, T, r0, 0.0, W + (size_t) r0 * ldw,
                        (int) ldw));
    }
    return YGPU_OK;
}
