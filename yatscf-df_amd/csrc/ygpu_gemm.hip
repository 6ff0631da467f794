// ygpu_gemm.hip -- FP64 MFMA GEMM family for gfx950.
//
// One tile engine serves every matrix product of the hot path:
//   B = pqA * J^{-1/2}                       (build_DF_tensor.c:239-254)
//   X_i = C_occ(partners of i)^T B(rows of i) (build_Fock.c:201-285, "aux. K tensor")
//   K = X X^T, upper tiles, split-K          (build_Fock.c:317-341, "K matrix")
//   C = X C0, D = C_occ C_occ^T               (build_density.c:56-68)
//   U diag U^T of the inverse square roots    (build_DF_tensor.c:205-237)
//   the DIIS products                         (DIIS.c:34-107)
// Workgroup tile 64x64, k-step 16, 256 threads = 4 wavefronts; each wavefront
// owns a 32x32 sub-tile computed as 2x2 v_mfma_f64_16x16x4f64 blocks from LDS.
// Operands are staged global -> registers -> LDS with the next k-step's loads
// issued before the current step's MFMAs.
#include "ygpu_internal.h"

#include <algorithm>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16, NTHR = 256, PAD = 2;

enum { MODE_N = 0, MODE_T = 1, MODE_GATHER_T = 2, MODE_GATHER_N = 3 };

// element (m,k) of A / (k,n) of B for the four storage modes
struct Operand {
    const double *p;
    long ld;
    const int *rows;   // gather: row index per k
};

// A tile: As[k][m].  N: A[m*ld+k].  T: A[k*ld+m].  GATHER_T: A[rows[k]*ld+m]
template <int MODE>
__device__ __forceinline__ void load_a(const Operand &A, int M, int K, int m0, int k0, double (&r)[4])
{
    const int t = threadIdx.x;
    if (MODE == MODE_N) {
        const int m = m0 + (t >> 2), kq = k0 + (t & 3) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = (m < M && kq + q < K) ? A.p[(size_t) m * A.ld + kq + q] : 0.0;
    } else {
        const int k = k0 + (t >> 4), mq = m0 + (t & 15) * 4;
        const bool kin = k < K;
        size_t row = 0;
        if (kin) row = (MODE == MODE_GATHER_T) ? (size_t) A.rows[k] : (size_t) k;
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = (kin && mq + q < M) ? A.p[row * A.ld + mq + q] : 0.0;
    }
}

template <int MODE>
__device__ __forceinline__ void store_a(double (*As)[BM + PAD], const double (&r)[4])
{
    const int t = threadIdx.x;
    if (MODE == MODE_N) {
        const int m = t >> 2, kq = (t & 3) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) As[kq + q][m] = r[q];
    } else {
        const int k = t >> 4, mq = (t & 15) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) As[k][mq + q] = r[q];
    }
}

// B tile: Bs[k][n].  N: B[k*ld+n].  T: B[n*ld+k].  GATHER_N: B[rows[k]*ld+n]
template <int MODE>
__device__ __forceinline__ void load_b(const Operand &B, int N, int K, int n0, int k0, double (&r)[4])
{
    const int t = threadIdx.x;
    if (MODE == MODE_T) {
        const int n = n0 + (t >> 2), kq = k0 + (t & 3) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = (n < N && kq + q < K) ? B.p[(size_t) n * B.ld + kq + q] : 0.0;
    } else {
        const int k = k0 + (t >> 4), nq = n0 + (t & 15) * 4;
        const bool kin = k < K;
        size_t row = 0;
        if (kin) row = (MODE == MODE_GATHER_N) ? (size_t) B.rows[k] : (size_t) k;
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = (kin && nq + q < N) ? B.p[row * B.ld + nq + q] : 0.0;
    }
}

template <int MODE>
__device__ __forceinline__ void store_b(double (*Bs)[BN + PAD], const double (&r)[4])
{
    const int t = threadIdx.x;
    if (MODE == MODE_T) {
        const int n = t >> 2, kq = (t & 3) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) Bs[kq + q][n] = r[q];
    } else {
        const int k = t >> 4, nq = (t & 15) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) Bs[k][nq + q] = r[q];
    }
}

// Core: accumulate a 64x64 tile over k in [kbeg, kend)
template <int MA, int MB>
__device__ __forceinline__ void tile_mma(const Operand &A, const Operand &B, int M, int N, int K, int m0, int n0,
                                         int kbeg, int kend, v4d (&acc)[2][2], bool zero = true)
{
    __shared__ double As[BK][BM + PAD];
    __shared__ double Bs[BK][BN + PAD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    if (zero) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    double ra[4], rb[4];
    if (kbeg < kend) {
        load_a<MA>(A, M, kend, m0, kbeg, ra);
        load_b<MB>(B, N, kend, n0, kbeg, rb);
    }
    (void) K;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        __syncthreads();
        store_a<MA>(As, ra);
        store_b<MB>(Bs, rb);
        __syncthreads();
        if (k0 + BK < kend) {
            load_a<MA>(A, M, kend, m0, k0 + BK, ra);
            load_b<MB>(B, N, kend, n0, k0 + BK, rb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), c = lane & 15;
            const double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
            const double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
}

// D[i][j] layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l, element v ->
// row l/16 + 4*v, col l%16.  Operands: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k.
template <class F>
__device__ __forceinline__ void tile_epilogue(int m0, int n0, const v4d (&acc)[2][2], F &&store)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + i * 16 + (lane >> 4) + 4 * v;
                const int n = n0 + wn + j * 16 + (lane & 15);
                store(m, n, acc[i][j][v]);
            }
}

template <int MA, int MB>
__global__ __launch_bounds__(NTHR) void k_gemm(Operand A, Operand B, int M, int N, int K, double alpha,
                                               double beta, double *__restrict__ C, long ldc)
{
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    v4d acc[2][2];
    tile_mma<MA, MB>(A, B, M, N, K, m0, n0, 0, K, acc);
    tile_epilogue(m0, n0, acc, [&](int m, int n, double v) {
        if (m < M && n < N) {
            double *c = C + (size_t) m * ldc + n;
            *c = (beta == 0.0) ? alpha * v : alpha * v + beta * *c;
        }
    });
}

// Split-K form for GEMMs with few output tiles (the n x n products of the SCF at n of a few
// hundred, 25 tiles at n = 260): workgroup (x, y, z) accumulates its tile over the k range
// [z kchunk, (z + 1) kchunk) and stores it to the partial P[z] (m x n, leading dimension n);
// k_gemm_reduce adds the partials in z order, so the result does not depend on scheduling.
template <int MA, int MB>
__global__ __launch_bounds__(NTHR) void k_gemm_part(Operand A, Operand B, int M, int N, int K, int kchunk,
                                                    double *__restrict__ P)
{
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    v4d acc[2][2];
    tile_mma<MA, MB>(A, B, M, N, K, m0, n0, kb, ke, acc);
    double *Pz = P + (size_t) blockIdx.z * M * N;
    tile_epilogue(m0, n0, acc, [&](int m, int n, double v) {
        if (m < M && n < N) Pz[(size_t) m * N + n] = v;
    });
}

__global__ void k_gemm_reduce(const double *__restrict__ P, int nsplit, int M, int N, double alpha, double beta,
                              double *__restrict__ C, long ldc)
{
    const long mn = (long) M * N;
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= mn) return;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += P[(size_t) z * mn + idx];
    double *c = C + (size_t) (idx / N) * ldc + idx % N;
    *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
}

constexpr int XKCH64 = 256;      // gather indices staged in LDS per k chunk

// X_i = Cocc[partners]^T * B[up rows]: workgroup t takes column tile t % ncolt, row tile
// (t / ncolt) % nrowt of basis function i = t / (nrowt ncolt), so consecutive workgroups read
// neighbouring segments of the same B rows (measured faster than grouping the row tiles of
// one B panel on an XCD).  The gather indices of each k chunk are staged in LDS: a stage's
// loads then wait on one memory latency, not on an index load followed by a data load.
// Five workgroups per CU requested: the accumulators then stay in VGPRs (80 in all) and six
// wavefronts per SIMD hide the gather latency (16 % faster than four at C10H22 and C96H24).
__global__ __launch_bounds__(NTHR, 5) void k_xbuild(const int *__restrict__ xptr, const int *__restrict__ xup,
                                                 const int *__restrict__ xpartner, const double *__restrict__ Cocc,
                                                 const double *__restrict__ Bup, double *__restrict__ X, int n_occ,
                                                 int naux_pad, int ncolt, int nrowt)
{
    const long t = blockIdx.x;
    const int ct = (int) (t % ncolt), rt = (int) ((t / ncolt) % nrowt), i = (int) (t / ((long) nrowt * ncolt));
    const int kb = xptr[i], klen = xptr[i + 1] - kb;
    const int m0 = rt * BM, n0 = ct * BN;
    __shared__ int s_pa[XKCH64], s_up[XKCH64];
    v4d acc[2][2];
    for (int kc = 0; kc < klen || kc == 0; kc += XKCH64) {
        const int kce = min(klen, kc + XKCH64);
        if (kc > 0) __syncthreads();            // the previous chunk's loads are done with the indices
        for (int k = kc + (int) threadIdx.x; k < kce; k += NTHR) {
            s_pa[k - kc] = xpartner[kb + k];
            s_up[k - kc] = xup[kb + k];
        }
        __syncthreads();
        Operand A{Cocc, n_occ, s_pa};
        Operand B{Bup, naux_pad, s_up};
        tile_mma<MODE_GATHER_T, MODE_GATHER_N>(A, B, n_occ, naux_pad, kce - kc, m0, n0, 0, kce - kc, acc, kc == 0);
        if (kce >= klen) break;
    }
    double *Xi = X + (size_t) i * n_occ * naux_pad;
    tile_epilogue(m0, n0, acc, [&](int m, int n, double v) {
        if (m < n_occ && n < naux_pad) Xi[(size_t) m * naux_pad + n] = v;
    });
}

// ---- X build with 160 x 128 tiles: X_i = Cocc[partners of i]^T B[up rows of i] as above, the
// 160 rows (occupied orbitals) and 128 columns (aux functions) of a workgroup split over four
// wavefronts of 80 x 64 (5 x 4 blocks of v_mfma_f64_16x16x4f64, 80 accumulators each in
// doubles).  2.2x the flops per loaded byte of the 64 x 64 engine, and 300 occupied orbitals
// pad to 320, not 384.  Both operands are gathered rows with the m / n index contiguous, so a
// stage is copied into LDS without a transpose: thread t moves 10 doubles of A row t / 16
// (columns 10 (t % 16) ..) and 8 of B row t / 16, the next stage's loads issued before the
// current stage's products.  LDS rows are padded to 128 bytes past a multiple of 256, so the
// two 16-lane halves of an operand read fall in different banks.  The gather indices of each
// chunk of XK160 k are staged in LDS first.
constexpr int XT_M = 160, XT_N = 128, XT_BK = 16, XT_PADM = 16, XT_PADN = 16, XK160 = 1024;

__global__ __launch_bounds__(NTHR, 2) void k_xbuild160(const int *__restrict__ xptr, const int *__restrict__ xup,
                                                    const int *__restrict__ xpartner,
                                                    const double *__restrict__ Cocc, const double *__restrict__ Bup,
                                                    double *__restrict__ X, int n_occ, int naux_pad, int ncolt,
                                                    int nrowt)
{
    __shared__ int s_pa[XK160], s_up[XK160];
    __shared__ double As[XT_BK][XT_M + XT_PADM];
    __shared__ double Bs[XT_BK][XT_N + XT_PADN];
    const long t = blockIdx.x;
    const int ct = (int) (t % ncolt), rt = (int) ((t / ncolt) % nrowt), i = (int) (t / ((long) nrowt * ncolt));
    const int kb = xptr[i], klen = xptr[i + 1] - kb;
    const int m0 = rt * XT_M, n0 = ct * XT_N;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 80, wn = (w & 1) * 64;
    const int lk = tid >> 4, lj = tid & 15;
    const int am = m0 + 10 * lj, bn = n0 + 8 * lj;
    v4d acc[5][4];
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    double ra[10], rb[8];
    // stage k0 (relative to the chunk start kc) of a chunk of kn k
    auto load = [&](int k0, int kn) {
        const int k = k0 + lk;
        const bool kin = k < kn;
        const double *pa = Cocc + (size_t) (kin ? s_pa[k] : 0) * n_occ;
        const double *pb = Bup + (size_t) (kin ? s_up[k] : 0) * naux_pad;
#pragma unroll
        for (int q = 0; q < 10; ++q) ra[q] = (kin && am + q < n_occ) ? pa[am + q] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) rb[q] = (kin && bn + q < naux_pad) ? pb[bn + q] : 0.0;
    };
    for (int kc = 0; kc < klen; kc += XK160) {
        const int kn = min(XK160, klen - kc);
        if (kc > 0) __syncthreads();            // the previous chunk's loads are done with the indices
        for (int k = tid; k < kn; k += NTHR) {
            s_pa[k] = xpartner[kb + kc + k];
            s_up[k] = xup[kb + kc + k];
        }
        __syncthreads();
        load(0, kn);
        for (int k0 = 0; k0 < kn; k0 += XT_BK) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 10; ++q) As[lk][10 * lj + q] = ra[q];
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[lk][8 * lj + q] = rb[q];
            __syncthreads();
            if (k0 + XT_BK < kn) load(k0 + XT_BK, kn);
#pragma unroll
            for (int kk = 0; kk < XT_BK; kk += 4) {
                const int kr = kk + (lane >> 4), c = lane & 15;
                double a[5], b[4];
#pragma unroll
                for (int x = 0; x < 5; ++x) a[x] = As[kr][wm + 16 * x + c];
#pragma unroll
                for (int y = 0; y < 4; ++y) b[y] = Bs[kr][wn + 16 * y + c];
#pragma unroll
                for (int x = 0; x < 5; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
            }
        }
    }
    double *Xi = X + (size_t) i * n_occ * naux_pad;
#pragma unroll
    for (int x = 0; x < 5; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * x + (lane >> 4) + 4 * v, n = n0 + wn + 16 * y + (lane & 15);
                if (m < n_occ && n < naux_pad) Xi[(size_t) m * naux_pad + n] = acc[x][y][v];
            }
}

// upper tiles of X X^T; blockIdx.x = linear upper-tile id, blockIdx.z = k split
__global__ __launch_bounds__(NTHR) void k_syrk_part(const double *__restrict__ X, int nbf, long kdim, int ntb,
                                                    long kchunk, double *__restrict__ part)
{
    int t = blockIdx.x, bi = 0;
    while (t >= ntb - bi) { t -= ntb - bi; ++bi; }
    const int bj = bi + t;
    const int m0 = bi * BM, n0 = bj * BN;
    const long kb = (long) blockIdx.z * kchunk;
    const long ke = std::min<long>(kdim, kb + kchunk);
    // offset the operands so the k range starts at 0 (int k indices inside the tile engine)
    Operand A{X + kb, kdim, nullptr};
    Operand B{X + kb, kdim, nullptr};
    v4d acc[2][2];
    tile_mma<MODE_N, MODE_T>(A, B, nbf, nbf, (int)(ke - kb), m0, n0, 0, (int)(ke - kb), acc);
    double *P = part + (size_t) blockIdx.z * nbf * nbf;
    tile_epilogue(m0, n0, acc, [&](int m, int n, double v) {
        if (m < nbf && n < nbf) P[(size_t) m * nbf + n] = v;
    });
}

// ---- SYRK with 128x128 tiles: 4 wavefronts, each a 64x64 sub-tile of 4x4 MFMA blocks.
// Twice the flops per loaded byte of the 64x64 engine; the k range is split in chunks
// and workgroups are ordered chunk-major and XCD-contiguous, so the tiles that read the
// same panels of one chunk run on one XCD at the same time.
constexpr int S_T = 128, S_BK = 16, S_PAD = 2;

__device__ __forceinline__ long xcd_remap64(long b, long total)
{
    const long per = total / 8, rem = total % 8, x = b % 8, slot = b / 8;
    return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

// acc += A(rows m0.., k in [kb, ke)) B(rows n0.., same k)^T with both operands stored as
// k-contiguous rows (lda, ldb); rows at or beyond arows / brows read as zero.  k extents and
// row starts are multiples of 8 doubles, so a thread's 8 values are all inside or all outside.
__device__ __forceinline__ void tile128_nt(const double *__restrict__ A, long lda, int arows, int m0,
                                           const double *__restrict__ B, long ldb, int brows, int n0, long kb,
                                           long ke, v4d (&acc)[4][4])
{
    __shared__ double As[S_BK][S_T + S_PAD];
    __shared__ double Bs[S_BK][S_T + S_PAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    // loader: thread t moves 8 consecutive k of row t/2 (half t%2) of both panels
    const int lr = t >> 1, lk = (t & 1) * 8;
    const bool ina = m0 + lr < arows, inb = n0 + lr < brows;
    const double *pa = A + (size_t) (ina ? m0 + lr : 0) * lda;
    const double *pb = B + (size_t) (inb ? n0 + lr : 0) * ldb;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
    double ra[8], rb[8];
    auto load = [&](long k0) {
        const long k = k0 + lk;
        const bool kin = k < ke;
        if (ina && kin) {
            const double2 *p2 = (const double2 *) (pa + k);
#pragma unroll
            for (int q = 0; q < 4; ++q) { const double2 v = p2[q]; ra[2 * q] = v.x; ra[2 * q + 1] = v.y; }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) ra[q] = 0.0;
        }
        if (inb && kin) {
            const double2 *p2 = (const double2 *) (pb + k);
#pragma unroll
            for (int q = 0; q < 4; ++q) { const double2 v = p2[q]; rb[2 * q] = v.x; rb[2 * q + 1] = v.y; }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) rb[q] = 0.0;
        }
    };
    // 16-row / 16-column blocks of this wavefront's 64 x 64 part that reach inside the matrix
    const int ilim = min(4, max(0, (arows - m0 - wm + 15) / 16));
    const int jlim = min(4, max(0, (brows - n0 - wn + 15) / 16));
    const bool wfull = ilim == 4 && jlim == 4;
    if (kb < ke) load(kb);
    for (long k0 = kb; k0 < ke; k0 += S_BK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            As[lk + q][lr] = ra[q];
            Bs[lk + q][lr] = rb[q];
        }
        __syncthreads();
        if (k0 + S_BK < ke) load(k0 + S_BK);
        if (wfull) {
#pragma unroll
            for (int kk = 0; kk < S_BK; kk += 4) {
                const int kr = kk + (lane >> 4), c = lane & 15;
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = As[kr][wm + 16 * i + c];
                    b[i] = Bs[kr][wn + 16 * i + c];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        } else if (ilim > 0 && jlim > 0) {
            // a wavefront on the ragged edge: only its 16 x 16 blocks that hold a row < arows and
            // a column < brows (the others stay zero and are never stored)
#pragma unroll
            for (int kk = 0; kk < S_BK; kk += 4) {
                const int kr = kk + (lane >> 4), c = lane & 15;
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = As[kr][wm + 16 * i + c];
                    b[i] = Bs[kr][wn + 16 * i + c];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (i < ilim && j < jlim)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
    }
}

template <class F>
__device__ __forceinline__ void tile128_epilogue(const v4d (&acc)[4][4], F &&store)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                store(wm + 16 * i + (lane >> 4) + 4 * v, wn + 16 * j + (lane & 15), acc[i][j][v]);
}

__global__ __launch_bounds__(NTHR) void k_syrk128(const double *__restrict__ X, int nbf, long kdim, int ntb,
                                                  int ntiles, long kchunk, long total, double *__restrict__ part)
{
    const long w_id = xcd_remap64(blockIdx.x, total);
    int tile = (int) (w_id % ntiles);
    const long z = w_id / ntiles;
    int bi = 0;
    while (tile >= ntb - bi) { tile -= ntb - bi; ++bi; }
    const int bj = bi + tile;
    const int m0 = bi * S_T, n0 = bj * S_T;
    const long kb = z * kchunk, ke = std::min<long>(kdim, kb + kchunk);
    v4d acc[4][4];
    tile128_nt(X, kdim, nbf, m0, X, kdim, nbf, n0, kb, ke, acc);
    double *P = part + (size_t) z * nbf * nbf;
    tile128_epilogue(acc, [&](int r, int c, double v) {
        const int m = m0 + r, n = n0 + c;
        if (m < nbf && n < nbf) P[(size_t) m * nbf + n] = v;
    });
}

// C [M][N] (ldc) = A [M][K] B[N][K]^T, 128x128 tiles, column tiles of one row panel
// consecutive on one XCD.  lower_b: B is lower triangular (B[n][k] = 0 for k > n), so the
// column tile [n0, n0 + 128) needs only k < n0 + 128
__global__ __launch_bounds__(NTHR) void k_gemm128_nt(const double *__restrict__ A, long lda,
                                                     const double *__restrict__ B, long ldb, int M, int N, long K,
                                                     int ncolt, long total, int lower_b, double *__restrict__ C,
                                                     long ldc)
{
    const long w_id = xcd_remap64(blockIdx.x, total);
    const int ct = (int) (w_id % ncolt), rt = (int) (w_id / ncolt);
    const int m0 = rt * S_T, n0 = ct * S_T;
    const long ke = lower_b ? std::min<long>(K, n0 + S_T) : K;
    v4d acc[4][4];
    tile128_nt(A, lda, M, m0, B, ldb, N, n0, 0, ke, acc);
    tile128_epilogue(acc, [&](int r, int c, double v) {
        const int m = m0 + r, n = n0 + c;
        if (m < M && n < N) C[(size_t) m * ldc + n] = v;
    });
}

// K[m][n] = K[n][m] = sum_z part[z][m][n] for m <= n
__global__ void k_syrk_reduce(const double *__restrict__ part, int nsplit, int nbf, double *__restrict__ K)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) nbf * nbf) return;
    const int m = (int)(idx / nbf), n = (int)(idx % nbf);
    if (m > n) return;
    double s = 0.0;
    for (int z = 0; z < nsplit; ++z) s += part[(size_t) z * nbf * nbf + idx];
    K[(size_t) m * nbf + n] = s;
    K[(size_t) n * nbf + m] = s;
}

}  // namespace

// ===========================================================================
template <int MA, int MB>
static void launch_gemm_part(hipStream_t s, dim3 grid, const Operand &a, const Operand &b, int m, int n, int k,
                             int kchunk, double *P)
{
    hipLaunchKernelGGL((k_gemm_part<MA, MB>), grid, dim3(NTHR), 0, s, a, b, m, n, k, kchunk, P);
}

int yg_dgemm(hipStream_t s, int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
             const double *B, int ldb, double beta, double *C, int ldc, DBuf *ws, int wg_target)
{
    if (m <= 0 || n <= 0) return YGPU_OK;
    dim3 grid((n + BN - 1) / BN, (m + BM - 1) / BM);
    Operand a{A, lda, nullptr}, b{B, ldb, nullptr};
    // fewer than wg_target / 4 output tiles: split k into chunks of at least two stages, up to
    // about wg_target workgroups in all, when a workspace is given
    const int tiles = (int) (grid.x * grid.y);
    if (ws && tiles < wg_target / 4 && k >= 4 * BK) {
        int nsplit = std::min(k / (2 * BK), (wg_target + tiles - 1) / tiles);
        int kchunk = (k + nsplit - 1) / nsplit;
        kchunk = (kchunk + BK - 1) / BK * BK;
        nsplit = (k + kchunk - 1) / kchunk;
        if (nsplit > 1) {
            const long mn = (long) m * n;
            YG_TRY(ws->ensure((size_t) nsplit * mn * sizeof(double)));
            double *P = ws->as<double>();
            dim3 g3(grid.x, grid.y, nsplit);
            if (!ta && !tb) launch_gemm_part<MODE_N, MODE_N>(s, g3, a, b, m, n, k, kchunk, P);
            else if (!ta && tb) launch_gemm_part<MODE_N, MODE_T>(s, g3, a, b, m, n, k, kchunk, P);
            else if (ta && !tb) launch_gemm_part<MODE_T, MODE_N>(s, g3, a, b, m, n, k, kchunk, P);
            else launch_gemm_part<MODE_T, MODE_T>(s, g3, a, b, m, n, k, kchunk, P);
            YG_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned) ((mn + 255) / 256)), dim3(256), 0, s, P, nsplit, m, n,
                               alpha, beta, C, (long) ldc);
            YG_HIP(hipGetLastError());
            return YGPU_OK;
        }
    }
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<MODE_N, MODE_N>), grid, dim3(NTHR), 0, s, a, b, m, n, k, alpha, beta, C, (long) ldc);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<MODE_N, MODE_T>), grid, dim3(NTHR), 0, s, a, b, m, n, k, alpha, beta, C, (long) ldc);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<MODE_T, MODE_N>), grid, dim3(NTHR), 0, s, a, b, m, n, k, alpha, beta, C, (long) ldc);
    else hipLaunchKernelGGL((k_gemm<MODE_T, MODE_T>), grid, dim3(NTHR), 0, s, a, b, m, n, k, alpha, beta, C, (long) ldc);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_dgemm_nt128(hipStream_t s, int m, int n, long k, const double *A, long lda, const double *B, long ldb,
                   double *C, long ldc, bool lower_b)
{
    if (m <= 0 || n <= 0) return YGPU_OK;
    if (k % 8 || lda % 8 || ldb % 8) {
        ygpu_set_error(YGPU_ERR_INVALID, "yg_dgemm_nt128: k and leading dimensions must be multiples of 8");
        return YGPU_ERR_INVALID;
    }
    const int ncolt = (n + S_T - 1) / S_T, nrowt = (m + S_T - 1) / S_T;
    const long total = (long) ncolt * nrowt;
    if (total > 0x7fffffffL) {
        ygpu_set_error(YGPU_ERR_INVALID, "yg_dgemm_nt128: %ld workgroups exceed the grid limit", total);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_gemm128_nt, dim3((unsigned) total), dim3(NTHR), 0, s, A, lda, B, ldb, m, n, k, ncolt, total,
                       (int) lower_b, C, ldc);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_xbuild(hipStream_t s, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup, const int *xpartner,
              const double *Cocc, const double *B, double *X, int max_row_len)
{
    (void) max_row_len;
    if (nbf <= 0 || n_occ <= 0) return YGPU_OK;
    // 160 x 128 tiles from 64 occupied orbitals up (fewer pad rows than 64 x 64 tiles waste below)
    const bool big = n_occ > 64;
    const int tm = big ? XT_M : BM, tn = big ? XT_N : BN;
    const int ncolt = (naux_pad + tn - 1) / tn, nrowt = (n_occ + tm - 1) / tm;
    const long total = (long) ncolt * nrowt * nbf;
    if (total > 0x7fffffffL) {
        ygpu_set_error(YGPU_ERR_INVALID, "X build: %ld workgroups exceed the grid limit", total);
        return YGPU_ERR_INVALID;
    }
    if (big)
        hipLaunchKernelGGL(k_xbuild160, dim3((unsigned) total), dim3(NTHR), 0, s, xptr, xup, xpartner, Cocc, B, X,
                           n_occ, naux_pad, ncolt, nrowt);
    else
        hipLaunchKernelGGL(k_xbuild, dim3((unsigned) total), dim3(NTHR), 0, s, xptr, xup, xpartner, Cocc, B, X, n_occ,
                           naux_pad, ncolt, nrowt);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_syrk_upper_split(int nbf, long kdim, long *kchunk_out)
{
    long kchunk;
    if (nbf >= 256) {
        // 128x128 tiles at two workgroups per CU (512 resident on 256 CUs); at most 8192
        // workgroups (16 rounds), with chunks of at least 1024 along k.  The finer split keeps
        // more workgroups in flight at the tail of the launch: at C96H24 (78 tiles, k = 2.0M)
        // 105 chunks against 26 at a 2048-workgroup cap, an FP64 solve 5.45 s against 5.55 s
        // (SYRK 89 against 93 ms per call), for 2.0 GB of partials instead of 0.51 GB.  The
        // split count is rounded down, so the last round is never a nearly empty extra one.
        const int ntb = (nbf + S_T - 1) / S_T;
        const int ntiles = ntb * (ntb + 1) / 2;
        const int nsplit = (int) std::max<long>(1, std::min<long>(8192 / ntiles, (kdim + 1023) / 1024));
        kchunk = (kdim + nsplit - 1) / nsplit;
        kchunk = (kchunk + S_BK - 1) / S_BK * S_BK;
    } else {
        const int ntb = (nbf + BM - 1) / BM;
        const int ntiles = ntb * (ntb + 1) / 2;
        // at most 2048 workgroups (two rounds at four per CU), chunks of at least 512 along k
        const int nsplit = (int) std::max<long>(1, std::min<long>(2048 / ntiles, (kdim + 511) / 512));
        kchunk = (kdim + nsplit - 1) / nsplit;
        kchunk = (kchunk + BK - 1) / BK * BK;
    }
    *kchunk_out = kchunk;
    return (int) ((kdim + kchunk - 1) / kchunk);
}

int yg_syrk_upper_ws(hipStream_t s, int nbf, long kdim, const double *X, double *K, double *part, long kchunk)
{
    if (nbf <= 0) return YGPU_OK;
    const int nsplit = (int) ((kdim + kchunk - 1) / kchunk);
    const long tot = (long) nbf * nbf;
    YG_TRY(yg_syrk_upper_parts(s, nbf, kdim, X, part, kchunk));
    hipLaunchKernelGGL(k_syrk_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, part, nsplit, nbf, K);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_syrk_upper_parts(hipStream_t s, int nbf, long kdim, const double *X, double *part, long kchunk)
{
    if (nbf <= 0) return YGPU_OK;
    const int nsplit = (int) ((kdim + kchunk - 1) / kchunk);
    if (nbf >= 256) {
        if (kdim % 8 || kchunk % S_BK) {
            ygpu_set_error(YGPU_ERR_INVALID, "SYRK: k extent %ld / chunk %ld not a multiple of 8 / 16", kdim, kchunk);
            return YGPU_ERR_INVALID;
        }
        const int ntb = (nbf + S_T - 1) / S_T;
        const int ntiles = ntb * (ntb + 1) / 2;
        const long total = (long) ntiles * nsplit;
        hipLaunchKernelGGL(k_syrk128, dim3((unsigned) total), dim3(NTHR), 0, s, X, nbf, kdim, ntb, ntiles, kchunk,
                           total, part);
        YG_HIP(hipGetLastError());
    } else {
        if (kchunk % BK || kchunk > 0x7fffffffL) {
            ygpu_set_error(YGPU_ERR_INVALID, "SYRK: chunk %ld not a multiple of 16", kchunk);
            return YGPU_ERR_INVALID;
        }
        const int ntb = (nbf + BM - 1) / BM;
        const int ntiles = ntb * (ntb + 1) / 2;
        hipLaunchKernelGGL(k_syrk_part, dim3(ntiles, 1, nsplit), dim3(NTHR), 0, s, X, nbf, kdim, ntb, kchunk, part);
        YG_HIP(hipGetLastError());
    }
    return YGPU_OK;
}

int yg_syrk_upper(hipStream_t s, int nbf, long kdim, const double *X, double *K, DBuf &part)
{
    if (nbf <= 0) return YGPU_OK;
    long kchunk = 0;
    const int nsplit = yg_syrk_upper_split(nbf, kdim, &kchunk);
    YG_TRY(part.ensure((size_t) nsplit * nbf * nbf * sizeof(double)));
    return yg_syrk_upper_ws(s, nbf, kdim, X, K, part.as<double>(), kchunk);
}
