// ygpu_eigen.hip -- the eigenproblem of the SCF density build for small n.
//
// build_density.c:54 diagonalises F' = X^T F X (dsyev) every iteration.  For
// n <= YG_EIG_NMAX_SMALL the device does it as
//   1. k_sytrd_reg: Householder reduction to tridiagonal form, F' = Q T Q^T, by one
//      workgroup of 512 threads with the whole lower triangle resident in registers
//      (LAPACK dsytd2 steps, lower variant: reflector k annihilates column k below the
//      subdiagonal and the trailing block takes the symmetric rank-2 update
//      A -= v w^T + w v^T).  The n - 2 dependent steps then touch no memory but LDS:
//      a step is one symmetric matrix-vector product and one rank-2 update on
//      registers, three workgroup barriers, and the reflector written out for step 3;
//   2. the eigenpairs of T: for the density all eigenvalues by Sturm multisection and the
//      n_occ lowest eigenvectors by inverse iteration (k_tridiag_*, below); for S^{-1/2}
//      and J^{-1/2} all eigenpairs by rocSOLVER dstedc;
//   3. k_backtransform_cols: u = Q z = H_0 (H_1 (... H_{n-3} z)) only for the
//      eigenvectors needed, one wavefront each with the vector in registers, reflectors
//      staged through LDS (also for the blocked reduction's reflectors, n <= 2048).
// rocSOLVER's dsyevd spends most of its time at these sizes in the
// latency-bound panel kernels of its own tridiagonal reduction.
//
// Register layout of k_sytrd_reg: thread (a, b), a < 16 (row class), b < 32 (column
// class), owns the elements (i, j) = (a + 16 ri, b + 32 rj) for the 81 slot pairs
// ri < 17, rj < 9, ri >= 2 rj, which covers the lower triangle of n <= 272 exactly
// once.  Slots with ri = 2 rj or ri = 2 rj + 1 straddle the diagonal: depending on
// (a, b) they hold a strictly lower element, a diagonal one, or an upper one.  Upper
// slots hold 0 and stay 0; diagonal elements are held halved, so that the symmetric
// product (row i collects A(i, j) v_j, column j collects A(i, j) v_i over every held
// element) needs no masks.  Elements outside n hold 0 and stay 0 (the reflector and w
// vanish there).  Columns j <= k are finished after step k: the steps run in five phases
// whose arithmetic skips the finished blocks of 32 columns (phase R = 0, 2, 4, 6, 8 covers
// 32 R - 1 <= k < 32 R + 63 and touches blocks rj >= R only).  Fewer, wider phases keep the
// 81 slots and the phases' shared LDS addresses inside 256 VGPRs without spills.
#include "ygpu_internal.h"

#include <cstdlib>

namespace {

constexpr int RTHR = 512;        // tridiagonal reduction: one workgroup
constexpr int NRI = 17, NRJ = 9; // row / column slots per thread
constexpr int NSLOT = 81;
constexpr int NREG = 288;        // LDS vector length: row indices < 16 * 17, column indices < 32 * 9
constexpr int SRS = 33, SCS = 17;// LDS row strides of the matrix-vector partial sums
static_assert(YG_EIG_NMAX_SMALL <= 16 * NRI, "register tiling covers n <= 272");

__device__ constexpr int slot_of(int ri, int rj) { return rj * NRI - rj * (rj - 1) + (ri - 2 * rj); }

// DPP move of a double within rows of 16 lanes (full permutations only: every lane has a source)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_d(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// sum over the wavefront, the same bits in every lane: butterflies within each row of 16
// (xor 1, xor 2, half-row mirror, row mirror; each pairs equal partial sums, so all lanes
// of a row agree), then the four row sums read out of lanes 0, 16, 32, 48
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<0xB1>(v);                        // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);                        // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);                       // row_half_mirror
    v += dpp_d<0x140>(v);                       // row_mirror
    return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}

// four wave_sum reductions with their steps interleaved (the same operations per value)
__device__ __forceinline__ void wave_sum4(double (&v)[4])
{
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] += dpp_d<0xB1>(v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] += dpp_d<0x4E>(v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] += dpp_d<0x141>(v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] += dpp_d<0x140>(v[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = (lane_d(v[u], 0) + lane_d(v[u], 16)) + (lane_d(v[u], 32) + lane_d(v[u], 48));
}

// xs[i] = A(i, k) for k <= i < n from the threads of column class k % 32 (slot column KRJ)
template <int KRJ>
__device__ __forceinline__ void extract_col(const double (&A)[NSLOT], int a, int b, int k, int n, double *xs)
{
    if (b != (k & 31)) return;
    // i = a + 16 ri against k and n through per-step offsets, so that no per-slot index is
    // kept live across steps
    const int ka = k - a, na = n - a;
    double *x = xs + a;
#pragma unroll
    for (int ri = 2 * KRJ; ri < NRI; ++ri) {
        if (16 * ri >= ka && 16 * ri < na) x[16 * ri] = (16 * ri == ka) ? 2.0 * A[slot_of(ri, KRJ)] : A[slot_of(ri, KRJ)];
    }
}

__device__ __forceinline__ void extract(const double (&A)[NSLOT], int a, int b, int k, int n, double *xs)
{
    switch (k >> 5) {                           // uniform over the workgroup
    case 0: extract_col<0>(A, a, b, k, n, xs); break;
    case 1: extract_col<1>(A, a, b, k, n, xs); break;
    case 2: extract_col<2>(A, a, b, k, n, xs); break;
    case 3: extract_col<3>(A, a, b, k, n, xs); break;
    case 4: extract_col<4>(A, a, b, k, n, xs); break;
    case 5: extract_col<5>(A, a, b, k, n, xs); break;
    case 6: extract_col<6>(A, a, b, k, n, xs); break;
    case 7: extract_col<7>(A, a, b, k, n, xs); break;
    default: extract_col<8>(A, a, b, k, n, xs); break;
    }
}

// partial sums of A v over the column blocks RJA <= rj < RJB (rows i < 32 RJA hold no
// element of those blocks); ACC adds the row sums to those of an earlier column range
template <int RJA, int RJB, bool ACC>
__device__ __forceinline__ void matvec_cols(const double (&A)[NSLOT], int a, int b, const double *vs, double *Srow,
                                            double *Scol)
{
    double vb[NRJ], colp[NRJ];
#pragma unroll
    for (int rj = RJA; rj < RJB; ++rj) {
        vb[rj] = vs[b + 32 * rj];
        colp[rj] = 0.0;
    }
#pragma unroll
    for (int ri = 2 * RJA; ri < NRI; ++ri) {
        const double va = vs[a + 16 * ri];
        double rowp = 0.0;
#pragma unroll
        for (int rj = RJA; rj < RJB && rj <= ri / 2; ++rj) {
            const double x = A[slot_of(ri, rj)];
            rowp += x * vb[rj];
            colp[rj] += x * va;
        }
        double *dst = Srow + (a + 16 * ri) * SRS + b;
        if (ACC) *dst += rowp;
        else *dst = rowp;
    }
#pragma unroll
    for (int rj = RJA; rj < RJB; ++rj) Scol[(b + 32 * rj) * SCS + a] = colp[rj];
}

// A -= v w^T + w v^T on the column blocks RJA <= rj < RJB; f0 / f1 scale the
// diagonal-straddling slots (1 strictly lower, 1/2 diagonal, 0 upper)
template <int RJA, int RJB>
__device__ __forceinline__ void update_cols(double (&A)[NSLOT], int a, int b, const double *vs, const double *ps,
                                            double c, double f0, double f1)
{
    double vb[NRJ], wb[NRJ];
#pragma unroll
    for (int rj = RJA; rj < RJB; ++rj) {
        vb[rj] = vs[b + 32 * rj];
        wb[rj] = ps[b + 32 * rj] + c * vb[rj];
    }
#pragma unroll
    for (int ri = 2 * RJA; ri < NRI; ++ri) {
        const double va = vs[a + 16 * ri];
        const double wa = ps[a + 16 * ri] + c * va;
#pragma unroll
        for (int rj = RJA; rj < RJB && rj <= ri / 2; ++rj) {
            const double u = va * wb[rj] + wa * vb[rj];
            if (ri == 2 * rj) A[slot_of(ri, rj)] -= f0 * u;
            else if (ri == 2 * rj + 1) A[slot_of(ri, rj)] -= f1 * u;
            else A[slot_of(ri, rj)] -= u;
        }
    }
}

// the live column blocks R <= rj < 9 in at most two ranges, [R, 5) and [5, 9), so that at
// most five column values of v (and of w) are held at a time
template <int R>
__device__ __forceinline__ void matvec_part(const double (&A)[NSLOT], int a, int b, const double *vs, double *Srow,
                                            double *Scol)
{
    if constexpr (R < 5) {
        matvec_cols<R, 5, false>(A, a, b, vs, Srow, Scol);
        matvec_cols<5, NRJ, true>(A, a, b, vs, Srow, Scol);
    } else {
        matvec_cols<R, NRJ, false>(A, a, b, vs, Srow, Scol);
    }
}

template <int R>
__device__ __forceinline__ void update_part(double (&A)[NSLOT], int a, int b, const double *vs, const double *ps,
                                            double c, double f0, double f1)
{
    if constexpr (R < 5) {
        update_cols<R, 5>(A, a, b, vs, ps, c, f0, f1);
        update_cols<5, NRJ>(A, a, b, vs, ps, c, f0, f1);
    } else {
        update_cols<R, NRJ>(A, a, b, vs, ps, c, f0, f1);
    }
}

struct SytrdLds {
    double xs[NREG], vs[NREG], ps[NREG];
    double red[RTHR / 64];
    double sc, sbeta, sdiag;
    double Scol[NREG * SCS];                    // Scol[j][a]: column-class partial sums of (A v)_j
    double Srow[NREG * SRS];                    // Srow[i][b]: row-class partial sums of (A v)_i
};

// steps k0 <= k < k1, all with the first live column block R = (k + 1) / 32
template <int R>
__device__ __forceinline__ void sytrd_phase(double (&A)[NSLOT], SytrdLds &L, int k0, int k1, int n, double *Ag,
                                            double *d, double *e, double *tau, double f0, double f1)
{
    const int t = threadIdx.x, a = t >> 5, b = t & 31, lane = t & 63;
    for (int k = k0; k < k1; ++k) {
        // column k lies in block R - 1 (first step), R or R + 1
        if (R > 0 && (k >> 5) < R) extract_col<(R > 0 ? R - 1 : 0)>(A, a, b, k, n, L.xs);
        else if ((k >> 5) == R) extract_col<R>(A, a, b, k, n, L.xs);
        else extract_col<(R < NRJ - 1 ? R + 1 : R)>(A, a, b, k, n, L.xs);
        __syncthreads();
        if (t < 64) {
            // dlarfg on x = A(k+1:n, k): H x = beta e_1, H = I - tau v v^T, v_0 = 1; the LDS
            // reads of each loop are issued together (fixed trip count)
            constexpr int NQ = (NREG + 63) / 64;
            const double alpha = L.xs[k + 1];
            double xv[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int i = k + 2 + t + 64 * q;
                xv[q] = (i < n) ? L.xs[i] : 0.0;
            }
            double ss = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) ss += xv[q] * xv[q];
            ss = wave_sum(ss);
            double tk = 0.0, beta = alpha, scale = 0.0;
            if (ss > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            // v over [k, NREG): entries below k are zero since earlier steps (v_{k-1} had its
            // unit at k, cleared here)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int i = k + t + 64 * q;
                if (i < NREG) L.vs[i] = (i == k + 1) ? 1.0 : (i > k + 1 && i < n) ? L.xs[i] * scale : 0.0;
            }
            if (t == 0) {
                L.sdiag = L.xs[k];
                L.sbeta = beta;
                L.sc = tk;
            }
        }
        __syncthreads();
        const double tk = L.sc;
        matvec_part<R>(A, a, b, L.vs, L.Srow, L.Scol);
        __syncthreads();
        // p = tau A22 v on the trailing indices (0 elsewhere), and p . v
        double pv = 0.0;
        if (t < NREG) {
            double p = 0.0;
            if (t > k && t < n) {
                double s0 = 0.0, s1 = 0.0;
#pragma unroll
                for (int q = 0; q < 32; q += 2) {
                    s0 += L.Srow[t * SRS + q];
                    s1 += L.Srow[t * SRS + q + 1];
                }
#pragma unroll
                for (int q = 0; q < 16; q += 2) {
                    s0 += L.Scol[t * SCS + q];
                    s1 += L.Scol[t * SCS + q + 1];
                }
                p = tk * (s0 + s1);
            }
            L.ps[t] = p;
            pv = p * L.vs[t];
        }
        pv = wave_sum(pv);
        if (lane == 0) L.red[t >> 6] = pv;
        __syncthreads();
        double vp = 0.0;
#pragma unroll
        for (int q = 0; q < RTHR / 64; ++q) vp += L.red[q];
        const double c = -0.5 * tk * vp;        // w = p + c v
        // the reflector and the step's scalars leave for memory here, so that the barrier
        // of the next step (which waits for this wave's stores) finds them done
        if (t > k && t < n) Ag[(size_t) k * n + t] = L.vs[t];
        if (t == 0) {
            d[k] = L.sdiag;
            e[k] = L.sbeta;
            tau[k] = tk;
        }
        update_part<R>(A, a, b, L.vs, L.ps, c, f0, f1);
    }
}

// In place on the row-major n x n matrix A (leading dimension n; the upper triangle is
// read, as dsyevd with fill_upper does).  Outputs: d[0..n), e[0..n-1), tau[0..n-1); the
// reflector k (v with v_0 = 1, over indices k+1..n-1) is left in row k right of the diagonal.
__global__ __launch_bounds__(RTHR) void k_sytrd_reg(int n, double *__restrict__ Ag, double *__restrict__ d,
                                                    double *__restrict__ e, double *__restrict__ tau)
{
    __shared__ SytrdLds L;
    const int t = threadIdx.x, a = t >> 5, b = t & 31;
    // diagonal-straddling slots: ri = 2 rj gives i - j = a - b, ri = 2 rj + 1 gives 16 + a - b
    const double f0 = (a > b) ? 1.0 : (a == b) ? 0.5 : 0.0;
    const double f1 = (b < a + 16) ? 1.0 : (b == a + 16) ? 0.5 : 0.0;

    double A[NSLOT];
#pragma unroll
    for (int rj = 0; rj < NRJ; ++rj) {
#pragma unroll
        for (int ri = 2 * rj; ri < NRI; ++ri) {
            const int i = a + 16 * ri, j = b + 32 * rj;
            double x = 0.0;
            if (i < n && j < n && i >= j) x = Ag[(size_t) j * n + i];
            A[slot_of(ri, rj)] = (i == j) ? 0.5 * x : x;
        }
    }
    __syncthreads();                            // every load is done before row k is overwritten

    // phase R: steps 32 R - 1 <= k < 32 R + 63 (and k < n - 2), R = 0, 2, 4, 6, 8
#define YG_PHASE(R) sytrd_phase<R>(A, L, (R) == 0 ? 0 : 32 * (R) - 1, min(n - 2, 32 * (R) + 63), n, Ag, d, e, tau, f0, f1)
    YG_PHASE(0);
    YG_PHASE(2);
    YG_PHASE(4);
    YG_PHASE(6);
    YG_PHASE(8);
#undef YG_PHASE
    // the last 2 x 2 block
    if (n >= 2) {
        extract(A, a, b, n - 2, n, L.xs);
        __syncthreads();
        if (t == 0) {
            d[n - 2] = L.xs[n - 2];
            e[n - 2] = L.xs[n - 1];
            tau[n - 2] = 0.0;
        }
        __syncthreads();
    }
    extract(A, a, b, n - 1, n, L.xs);
    __syncthreads();
    if (t == 0) d[n - 1] = L.xs[n - 1];
}

constexpr int BLV = 4;           // back-transform above the one-workgroup size: vectors per workgroup

// Z <- Q Z for the first nvec columns of Z (column j = z_j, leading dimension ldz; the
// dstedc layout) with Q = H_0 H_1 ... H_{n-3} of the blocked reduction, whose reflector k
// sits in row k of A (leading dimension lda) at columns k + 2 .. n - 1 with its unit
// element at k + 1 implicit (LAPACK dsytrd's lower layout of the column-major view).  In
// place: each wavefront reads its vector into registers (element l + 64 m in lane l,
// m < MB) before anything is written.  The reflectors are staged in LDS BK at a time,
// zero outside their support and padded to 64 MB columns, so that a reflector step reads
// all MB chunks without branches; the next block's loads are held in registers (thread
// t: columns t + 256 u) while the current one is applied.
template <int MB, int BK>
__global__ __launch_bounds__(64 * BLV) void k_backtransform_cols(int n, int nvec, const double *__restrict__ A,
                                                                 long lda, const double *__restrict__ tau,
                                                                 double *__restrict__ Z, long ldz)
{
    extern __shared__ double sV[];              // [BK][64 MB]
    __shared__ double stau[BK];
    constexpr int NT = 64 * BLV, UC = 64 * MB / NT, LN = 64 * MB;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int j = blockIdx.x * BLV + wv;
    const int jj = min(j, nvec - 1);            // surplus wavefronts repeat the last vector, store nothing
    double r[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
        const int c = lane + 64 * m;
        r[m] = (c < n) ? Z[(size_t) jj * ldz + c] : 0.0;
    }
    // block at ktop holds reflectors ktop - q, q < BK: sV[q][c] = v_k(c) (1 at k + 1, 0 at c <= k)
    double pre[BK][UC];
    auto load_block = [&](int ktop) {
#pragma unroll
        for (int q = 0; q < BK; ++q) {
            const int k = ktop - q;
            const double *row = A + (size_t) max(k, 0) * lda;
#pragma unroll
            for (int u = 0; u < UC; ++u) {
                const int c = t + NT * u;
                double v = 0.0;
                if (k >= 0 && c < n) v = (c == k + 1) ? 1.0 : (c > k + 1) ? row[c] : 0.0;
                pre[q][u] = v;
            }
        }
    };
    const int ktop0 = n - 3;
    if (ktop0 >= 0) load_block(ktop0);
    for (int ktop = ktop0; ktop >= 0; ktop -= BK) {
        __syncthreads();                        // the previous block's reflectors are applied
#pragma unroll
        for (int q = 0; q < BK; ++q)
#pragma unroll
            for (int u = 0; u < UC; ++u) sV[q * LN + t + NT * u] = pre[q][u];
        if (t < BK) stau[t] = (ktop - t >= 0) ? tau[ktop - t] : 0.0;
        __syncthreads();
        if (ktop - BK >= 0) load_block(ktop - BK);
        const int qn = min(BK, ktop + 1);
        for (int q = 0; q < qn; ++q) {
            const double *v = sV + q * LN + lane;
            double vm[MB], s4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < MB; ++m) vm[m] = v[64 * m];
#pragma unroll
            for (int m = 0; m < MB; ++m) s4[m & 3] += vm[m] * r[m];
            const double f = stau[q] * wave_sum((s4[0] + s4[1]) + (s4[2] + s4[3]));
#pragma unroll
            for (int m = 0; m < MB; ++m) r[m] -= f * vm[m];
        }
    }
    if (j < nvec)
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int c = lane + 64 * m;
            if (c < n) Z[(size_t) j * ldz + c] = r[m];
        }
}

// ---------------------------------------------------------------------------------
// Eigenpairs of the symmetric tridiagonal T = tridiag(e, d, e) (n <= YG_EIG_NMAX_SMALL):
// all eigenvalues by Sturm-count multisection, the lowest nvec eigenvectors by inverse
// iteration, orthogonalised within clusters.  The steps follow LAPACK's dstebz (Sturm
// counts with the pivmin guard, Gershgorin start interval) and dstein (shift separation
// 10 eps |lambda| inside runs of equal eigenvalues, clusters of eigenvalues closer than
// 1e-3 ||T||, Gram-Schmidt inside a cluster); the parallel form is this file's own: 256
// trial points per workgroup and round, one lane per vector in the inverse iteration.

constexpr int NBIS = 256;        // bisection: one workgroup per eigenvalue
constexpr int NORTH = 256;       // cluster orthogonalisation: one workgroup

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}

struct TriBounds {
    double gl, gu, tnorm, pivmin;
};

// Gershgorin interval (widened as dstebz does), ||T|| and the Sturm pivot floor, by one
// wavefront (every lane returns them)
__device__ TriBounds tri_bounds(int n, const double *d, const double *e)
{
    const int lane = threadIdx.x & 63;
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double el = (i > 0) ? fabs(e[i - 1]) : 0.0, er = (i < n - 1) ? fabs(e[i]) : 0.0;
        gl = fmin(gl, d[i] - el - er);
        gu = fmax(gu, d[i] + el + er);
        emax2 = fmax(emax2, er * er);
    }
    TriBounds b;
    b.gl = wave_min(gl);
    b.gu = wave_max(gu);
    b.tnorm = fmax(fabs(b.gl), fabs(b.gu));
    b.pivmin = DBL_MIN * fmax(1.0, wave_max(emax2));
    const double widen = 2.0 * b.tnorm * DBL_EPSILON * n + 2.0 * b.pivmin;
    b.gl -= widen;
    b.gu += widen;
    return b;
}

// 1/q to full double precision: the hardware reciprocal and two Newton steps (the
// recurrence below waits on this chain, which is half as long as IEEE division's)
__device__ __forceinline__ double recip(double q)
{
    double r = __builtin_amdgcn_rcp(q);
    double t = fma(-q, r, 1.0);
    r = fma(r, t, r);
    t = fma(-q, r, 1.0);
    return fma(r, t, r);
}

// number of eigenvalues of T below x (|q| >= pivmin keeps 1/q finite)
__device__ __forceinline__ int sturm_count(int n, const double *d, const double *e2, double x, double pivmin)
{
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    int cnt = q < 0.0;
    // unrolled so that the loads of d and e2 are issued ahead of the dependent chain
#pragma unroll 8
    for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - e2[i - 1] * recip(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// w[j], j < n, ascending: workgroup j narrows [lo, hi] around lambda_j by testing the
// 256 interior points lo + (hi - lo) (k + 1) / 257 per round, one per thread
__global__ __launch_bounds__(NBIS) void k_tridiag_bisect(int n, const double *__restrict__ d,
                                                         const double *__restrict__ e, double *__restrict__ w)
{
    __shared__ double sd[YG_EIG_NMAX_SMALL], se2[YG_EIG_NMAX_SMALL];
    __shared__ int below[2][NBIS / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = t; i < n; i += NBIS) {
        sd[i] = d[i];
        se2[i] = (i < n - 1) ? e[i] * e[i] : 0.0;
    }
    __syncthreads();
    const TriBounds bd = tri_bounds(n, sd, e);
    const int j = blockIdx.x;
    double lo = bd.gl, hi = bd.gu;
    // dstebz's stopping rule: relative 2 ulp, or absolute ulp ||T|| (so that eigenvalues near 0
    // stop too); 257^7 > 2 / ulp, so at most 7 rounds
    const double atol = DBL_EPSILON * bd.tnorm;
    for (int r = 0; r < 16; ++r) {
        if (hi - lo <= fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), atol) + DBL_MIN) break;   // uniform
        const double x = lo + (hi - lo) * ((double) (t + 1) / (double) (NBIS + 1));
        const int c = sturm_count(n, sd, se2, x, bd.pivmin);
        const int Lw = __popcll(__ballot(c <= j));
        if (lane == 0) below[r & 1][wv] = Lw;
        __syncthreads();                        // double-buffered: one barrier per round
        int L = 0;
#pragma unroll
        for (int q = 0; q < NBIS / 64; ++q) L += below[r & 1][q];   // points at or below lambda_j: 0..L-1
        const double nlo = (L > 0) ? lo + (hi - lo) * ((double) L / (double) (NBIS + 1)) : lo;
        const double nhi = (L < NBIS) ? lo + (hi - lo) * ((double) (L + 1) / (double) (NBIS + 1)) : hi;
        lo = nlo;
        hi = nhi;
    }
    if (t == 0) w[j] = 0.5 * (lo + hi);
}

__device__ __forceinline__ double start_value(int i, int j)
{
    unsigned h = (unsigned) i * 2654435761u ^ ((unsigned) j * 40503u + 12345u) * 2246822519u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    return (double) h * (2.0 / 4294967296.0) - 1.0;
}

constexpr int IVL = 16;          // inverse iteration: vectors per workgroup (lanes 0..15 of one wavefront)

// Inverse iteration, lane l < IVL of workgroup b on eigenvalue j = IVL b + l (lanes past nvec
// repeat the last one and store nothing).  T - sigma I = L D L^T without pivoting, tiny pivots
// replaced by +-eps ||T||; three solves from a seeded start vector, each rescaled by its
// largest element; the result normalised into column j of Zt (n x nvp).  The factors and the
// iterate live in LDS, so the sequential sweeps wait on LDS latency only.
__global__ __launch_bounds__(64) void k_tridiag_invit(int n, int nvec, int nvp, const double *__restrict__ d,
                                                      const double *__restrict__ e, const double *__restrict__ w,
                                                      double *__restrict__ Zt)
{
    __shared__ double zs[YG_EIG_NMAX_SMALL * IVL], ls[YG_EIG_NMAX_SMALL * IVL], ds[YG_EIG_NMAX_SMALL * IVL];
    __shared__ double sd[YG_EIG_NMAX_SMALL], se[YG_EIG_NMAX_SMALL];
    const int lane = threadIdx.x, j = blockIdx.x * IVL + lane;
    for (int i = lane; i < n; i += 64) {
        sd[i] = d[i];
        se[i] = (i < n - 1) ? e[i] : 0.0;
    }
    __syncthreads();
    const TriBounds bd = tri_bounds(n, sd, se);   // a wavefront-wide reduction: every lane takes part
    if (lane >= IVL) return;
    const int jj = min(j, nvec - 1);
    // dstein: inside a run of (nearly) equal eigenvalues the shifts are spread by 10 eps |lambda|
    const double pt = 10.0 * DBL_EPSILON, floor_ = bd.tnorm * DBL_EPSILON;
    int m = 0;
    while (jj - m - 1 >= 0 && w[jj - m] - w[jj - m - 1] < pt * fmax(fabs(w[jj - m]), floor_)) ++m;
    const double sig = w[jj] + m * pt * fmax(fabs(w[jj]), floor_);
    const double ptiny = DBL_EPSILON * bd.tnorm + DBL_MIN;
    double D = sd[0] - sig;
    if (fabs(D) < ptiny) D = (D < 0.0) ? -ptiny : ptiny;
    double id = recip(D);
    ds[lane] = id;
    zs[lane] = start_value(0, jj);
#pragma unroll 4
    for (int i = 1; i < n; ++i) {
        const double ei = se[i - 1];
        const double l = ei * id;
        ls[(i - 1) * IVL + lane] = l;
        D = (sd[i] - sig) - l * ei;
        if (fabs(D) < ptiny) D = (D < 0.0) ? -ptiny : ptiny;
        id = recip(D);
        ds[i * IVL + lane] = id;
        zs[i * IVL + lane] = start_value(i, jj);
    }
    double s = 1.0;
    for (int it = 0; it < 3; ++it) {
        // forward: u_i = s x_i - l_{i-1} u_{i-1}, kept as u_i / D_i
        double u = s * zs[lane];
        zs[lane] = u * ds[lane];
#pragma unroll 8
        for (int i = 1; i < n; ++i) {
            u = s * zs[i * IVL + lane] - ls[(i - 1) * IVL + lane] * u;
            zs[i * IVL + lane] = u * ds[i * IVL + lane];
        }
        // back: y_i = u_i / D_i - l_i y_{i+1}
        double y = zs[(n - 1) * IVL + lane], amax = fabs(y);
#pragma unroll 8
        for (int i = n - 2; i >= 0; --i) {
            y = zs[i * IVL + lane] - ls[i * IVL + lane] * y;
            zs[i * IVL + lane] = y;
            amax = fmax(amax, fabs(y));
        }
        s = 1.0 / amax;
    }
    double ss = 0.0;
    for (int i = 0; i < n; ++i) {
        const double v = s * zs[i * IVL + lane];
        ss += v * v;
    }
    const double f = s / sqrt(ss);
    if (j < nvec)
        for (int i = 0; i < n; ++i) Zt[(size_t) i * nvp + j] = f * zs[i * IVL + lane];
}

constexpr int CLMAX = 64;        // clusters up to this size are orthogonalised in LDS

// max_i |(T z - lambda z)_i| of z (element i at z[i * stride]) over wavefront lanes
__device__ __forceinline__ double tri_residual(int n, const double *__restrict__ d, const double *__restrict__ e,
                                               double lam, const double *z, long stride, int lane)
{
    double r = 0.0;
    for (int i = lane; i < n; i += 64) {
        double tz = (d[i] - lam) * z[(size_t) i * stride];
        if (i > 0) tz += e[i - 1] * z[(size_t) (i - 1) * stride];
        if (i < n - 1) tz += e[i] * z[(size_t) (i + 1) * stride];
        r = fmax(r, fabs(tz));
    }
    return wave_max(r);
}

// Gram-Schmidt (two passes) inside each cluster of eigenvalues closer than 1e-3 ||T||:
// workgroup j works only if eigenvalue j starts a cluster, on that cluster, and writes its
// rows Z[j'][i] = Zt[i][j'].  Every vector of the cluster is then checked, |T z - lambda z| <=
// rtol ||T|| (inverse iteration has no convergence test of its own): the first failing vector j
// sets *info = j + 1
__global__ __launch_bounds__(NORTH) void k_tridiag_orth(int n, int nvec, int nvp, const double *__restrict__ d,
                                                        const double *__restrict__ e, const double *__restrict__ w,
                                                        double *__restrict__ Zt, double *__restrict__ Z,
                                                        int *__restrict__ info, double rtol)
{
    __shared__ double cs[CLMAX * YG_EIG_NMAX_SMALL];
    __shared__ double sdot[CLMAX];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int j0 = blockIdx.x;
    const TriBounds bd = tri_bounds(n, d, e);
    const double ortol = 1e-3 * bd.tnorm;
    if (j0 > 0 && w[j0] - w[j0 - 1] < ortol) return;   // not a cluster start (uniform)
    int j1 = j0 + 1;
    while (j1 < nvec && w[j1] - w[j1 - 1] < ortol) ++j1;
    const int csz = j1 - j0;
    if (csz <= CLMAX) {
        for (long idx = t; idx < (long) csz * n; idx += NORTH) {
            const int v = (int) (idx / n), i = (int) (idx % n);
            cs[idx] = Zt[(size_t) i * nvp + j0 + v];
        }
        __syncthreads();
        if (wv == 0) {
            // classical Gram-Schmidt applied twice, the vector in registers (element lane + 64 q):
            // the dots of a pass against z_0 .. z_{v-1} are independent of each other
            constexpr int NQ = (YG_EIG_NMAX_SMALL + 63) / 64;
            for (int v = 1; v < csz; ++v) {
                double *zv = cs + (size_t) v * n;
                double x[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int i = lane + 64 * q;
                    x[q] = (i < n) ? zv[i] : 0.0;
                }
                for (int pass = 0; pass < 2; ++pass) {
                    for (int k0 = 0; k0 < v; k0 += 4) {
                        double dot[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const double *zk = cs + (size_t) min(k0 + u, v - 1) * n;
                            dot[u] = 0.0;
#pragma unroll
                            for (int q = 0; q < NQ; ++q) {
                                const int i = lane + 64 * q;
                                dot[u] += (i < n) ? x[q] * zk[i] : 0.0;
                            }
                        }
                        wave_sum4(dot);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (lane == 0 && k0 + u < v) sdot[k0 + u] = dot[u];
                    }
                    for (int k = 0; k < v; ++k) {
                        const double *zk = cs + (size_t) k * n;
                        const double dk = sdot[k];
#pragma unroll
                        for (int q = 0; q < NQ; ++q) {
                            const int i = lane + 64 * q;
                            if (i < n) x[q] -= dk * zk[i];
                        }
                    }
                }
                double ss = 0.0;
#pragma unroll
                for (int q = 0; q < NQ; ++q) ss += x[q] * x[q];
                const double f = 1.0 / sqrt(wave_sum(ss));
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int i = lane + 64 * q;
                    if (i < n) zv[i] = f * x[q];
                }
            }
        }
        __syncthreads();
        for (int v = wv; v < csz; v += NORTH / 64) {
            const double r = tri_residual(n, d, e, w[j0 + v], cs + (size_t) v * n, 1, lane);
            if (lane == 0 && !(r <= rtol * bd.tnorm)) atomicCAS(info, 0, j0 + v + 1);
        }
        for (long idx = t; idx < (long) csz * n; idx += NORTH) Z[(size_t) j0 * n + idx] = cs[idx];
        return;
    }
    // a larger cluster: the same in place in Zt by wavefront 0, visible to itself through fences
    if (wv == 0) {
        for (int j = j0 + 1; j < j1; ++j) {
            for (int pass = 0; pass < 2; ++pass) {
                for (int k = j0; k < j; ++k) {
                    double dot = 0.0;
                    for (int i = lane; i < n; i += 64) dot += Zt[(size_t) i * nvp + j] * Zt[(size_t) i * nvp + k];
                    dot = wave_sum(dot);
                    for (int i = lane; i < n; i += 64) Zt[(size_t) i * nvp + j] -= dot * Zt[(size_t) i * nvp + k];
                    __threadfence();
                }
            }
            double ss = 0.0;
            for (int i = lane; i < n; i += 64) ss += Zt[(size_t) i * nvp + j] * Zt[(size_t) i * nvp + j];
            const double f = 1.0 / sqrt(wave_sum(ss));
            for (int i = lane; i < n; i += 64) Zt[(size_t) i * nvp + j] *= f;
            __threadfence();
        }
    }
    __syncthreads();
    for (int v = wv; v < csz; v += NORTH / 64) {
        const double r = tri_residual(n, d, e, w[j0 + v], Zt + j0 + v, nvp, lane);
        if (lane == 0 && !(r <= rtol * bd.tnorm)) atomicCAS(info, 0, j0 + v + 1);
    }
    for (long idx = t; idx < (long) csz * n; idx += NORTH) {
        const int v = (int) (idx / n), i = (int) (idx % n);
        Z[(size_t) (j0 + v) * n + i] = Zt[(size_t) i * nvp + j0 + v];
    }
}

}  // namespace

int yg_sytrd_small(hipStream_t s, int n, double *A, double *d, double *e, double *tau)
{
    if (n < 1 || n > YG_EIG_NMAX_SMALL) {
        ygpu_set_error(YGPU_ERR_INVALID, "small tridiagonal reduction: n = %d outside 1..%d", n, YG_EIG_NMAX_SMALL);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_sytrd_reg, dim3(1), dim3(RTHR), 0, s, n, A, d, e, tau);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_backtransform_cols(hipStream_t s, int n, int nvec, const double *A, long lda, const double *tau, double *Z,
                          long ldz)
{
    if (n < 1 || n > YG_SYTRD_BLOCKED_NMAX || nvec < 0 || nvec > n || lda < n || ldz < n) {
        ygpu_set_error(YGPU_ERR_INVALID, "back-transform: n = %d, nvec = %d, lda = %ld, ldz = %ld out of range", n,
                       nvec, lda, ldz);
        return YGPU_ERR_INVALID;
    }
    if (nvec == 0) return YGPU_OK;
    const int mbn = (n + 63) / 64;
    const dim3 grid((nvec + BLV - 1) / BLV), blk(64 * BLV);
    auto lds = [](int mb, int bk) { return (size_t) bk * 64 * mb * sizeof(double); };
    if (mbn <= 8)
        hipLaunchKernelGGL((k_backtransform_cols<8, 16>), grid, blk, lds(8, 16), s, n, nvec, A, lda, tau, Z, ldz);
    else if (mbn <= 16)
        hipLaunchKernelGGL((k_backtransform_cols<16, 16>), grid, blk, lds(16, 16), s, n, nvec, A, lda, tau, Z, ldz);
    else if (mbn <= 24)
        hipLaunchKernelGGL((k_backtransform_cols<24, 4>), grid, blk, lds(24, 4), s, n, nvec, A, lda, tau, Z, ldz);
    else
        hipLaunchKernelGGL((k_backtransform_cols<32, 4>), grid, blk, lds(32, 4), s, n, nvec, A, lda, tau, Z, ldz);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_tridiag_eig(hipStream_t s, int n, int nvec, const double *d, const double *e, double *w, double *Z,
                   double *ws, int *info)
{
    // the residual bound; YGPU_TRIDIAG_RESID in the environment replaces it (0 sends every
    // small-route build to the general eigensolver, as the tests of that route do)
    double rtol = YG_TRIDIAG_RESID;
    if (const char *v = getenv("YGPU_TRIDIAG_RESID")) rtol = atof(v);
    if (n < 1 || n > YG_EIG_NMAX_SMALL || nvec < 1 || nvec > n) {
        ygpu_set_error(YGPU_ERR_INVALID, "tridiagonal eigensolver: n = %d, nvec = %d out of range", n, nvec);
        return YGPU_ERR_INVALID;
    }
    const int nvp = (nvec + 63) / 64 * 64;
    double *Zt = ws;
    hipLaunchKernelGGL(k_tridiag_bisect, dim3(n), dim3(NBIS), 0, s, n, d, e, w);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tridiag_invit, dim3((nvec + IVL - 1) / IVL), dim3(64), 0, s, n, nvec, nvp, d, e, w, Zt);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tridiag_orth, dim3(nvec), dim3(NORTH), 0, s, n, nvec, nvp, d, e, w, Zt, Z, info, rtol);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}
