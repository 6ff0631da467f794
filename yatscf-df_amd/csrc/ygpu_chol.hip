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
// variant on the row-major matrix): one wavefront factors the diagonal block
// in LDS and inverts it; the panel below is solved by a product with that
// inverse and the trailing matrix updated, both on the FP64 MFMA GEMM of
// ygpu_gemm.hip.  W = L^{-1} then follows block row by block row:
//   W_ii = L_ii^{-1},  W[i, 0:i] = -W_ii (L[i, 0:i] W[0:i, 0:i]),
// the products of a block row split along k so that the whole GPU works on them.
// (A recursive variant on halves, L21 = A21 L11^{-1 T} with the explicit inverse
// of the leading half, measured ten times the error in W J W^T - I on a metric of
// condition 4e10: the triangular solves stay with the 64 x 64 blocks.)
#include "ygpu_internal.h"

#include <algorithm>

namespace {

constexpr int NB = 64;

// Factor the NB x NB diagonal block at (k0, k0) of the row-major A (leading dimension lda) in
// place: its lower triangle becomes L_kk (strictly upper part zeroed), and Linv (NB x NB,
// leading dimension ldl) receives L_kk^{-1} (zero above the diagonal).  Rows / columns at or
// beyond n keep an identity block.  A pivot that is not positive sets *info to its global column
// + 1 (first failure kept) and is replaced by 1 so that the remaining work stays finite.
// One wavefront, lane i owning row i in registers (Cholesky-Crout by rows): at column j every
// lane forms a_ij - sum_{k<j} L_ik L_jk with row j of L read from LDS as a broadcast, lane j
// takes the square root and the lanes below divide by it.  Then lane c solves L x = e_c by
// forward substitution for column c of L^{-1}, again against broadcast rows of L.  (The same
// loops over LDS rows instead of registers measured 125 against 73 us per block.)
__global__ __launch_bounds__(64) void k_potrf_diag(int n, long lda, double *__restrict__ A, int k0,
                                                   double *__restrict__ Linv, long ldl, int *__restrict__ info)
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
    for (int q = 0; q < NB; ++q) Linv[(size_t) q * ldl + i] = x[q];
}

}  // namespace

size_t yg_chol_inv_ws(int npad) { return (size_t) npad * NB; }

int yg_chol_inv_lower(hipStream_t s, int n, int npad, long lda, double *A, double *W, long ldw, int *info, double *T,
                      DBuf *ws)
{
    if (npad % NB || npad < n || n < 1 || lda % 8 || ldw % 8) {
        ygpu_set_error(YGPU_ERR_INVALID, "Cholesky: order %d padded to %d (a multiple of %d), leading dimensions %ld / %ld",
                       n, npad, NB, lda, ldw);
        return YGPU_ERR_INVALID;
    }
    // W lower triangular: zero above the diagonal blocks, which are written by the block factors
    YG_HIP(hipMemset2DAsync(W, ldw * sizeof(double), 0, npad * sizeof(double), npad, s));
    const int nblk = npad / NB;
    for (int b = 0; b < nblk; ++b) {
        const int k0 = b * NB;
        double *wbb = W + (size_t) k0 * ldw + k0;
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, n, lda, A, k0, wbb, ldw, info);
        YG_HIP(hipGetLastError());
        const int rest = npad - k0 - NB;
        if (rest <= 0) break;
        double *panel = A + (size_t) (k0 + NB) * lda + k0;
        // L_21 = A_21 L_11^{-T} into T (rest x NB), then back into the panel's place in A
        YG_TRY(yg_dgemm(s, 0, 1, rest, NB, NB, 1.0, panel, (int) lda, wbb, (int) ldw, 0.0, T, NB));
        YG_HIP(hipMemcpy2DAsync(panel, lda * sizeof(double), T, NB * sizeof(double), NB * sizeof(double), rest,
                                hipMemcpyDeviceToDevice, s));
        // trailing matrix A_22 -= L_21 L_21^T
        double *a22 = A + (size_t) (k0 + NB) * lda + k0 + NB;
        YG_TRY(yg_dgemm(s, 0, 1, rest, rest, NB, -1.0, T, NB, T, NB, 1.0, a22, (int) lda));
    }
    // W block row by block row: T = L[b, 0:b] W[0:b, 0:b] (NB x r0; the zeros of W above its
    // diagonal make column c of the product start at k = c), then W[b, 0:b] = -W_bb T.  The
    // products have r0 / NB output tiles and depth r0: split along k over the GPU (partials in ws)
    for (int b = 1; b < nblk; ++b) {
        const int r0 = b * NB;
        double *wb = W + (size_t) r0 * ldw;
        YG_TRY(yg_dgemm(s, 0, 0, NB, r0, r0, 1.0, A + (size_t) r0 * lda, (int) lda, W, (int) ldw, 0.0, T, r0, ws, 1024));
        YG_TRY(yg_dgemm(s, 0, 0, NB, r0, NB, -1.0, wb + r0, (int) ldw, T, r0, 0.0, wb, (int) ldw));
    }
    return YGPU_OK;
}
