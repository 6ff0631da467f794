// ygpu_sytrd.hip -- blocked Householder tridiagonalisation for the mid-size symmetric
// eigenproblems of the SCF: J^{-1/2} (build_DF_tensor.c:205-237, dsyev of the metric) and
// the density build above the one-workgroup size (build_density.c:54, dsyev of F').
//
// LAPACK dsytrd / dlatrd, lower variant, panels of NB columns.  Inside a panel the trailing
// block keeps its panel-start values A0 and the transformations so far are held as
// A = A0 - V W^T - W V^T (V the reflectors, W = tau A v corrected); a column i is
//   1. finished from the previous column's product y: w = tau y - (tau^2 / 2)(y . v) v,
//   2. brought up to date, x = A0(:, i) - V W(i, :)^T - W V(i, :)^T,
//   3. reduced by a reflector (dlarfg) that zeroes x below the subdiagonal,
//   4. multiplied into the trailing block: y = A0 v - V (W^T v) - W (V^T v).
// One kernel per column (k_latrd_step): each of its workgroups repeats steps 1-3 and the
// small products W^T v, V^T v for all rows (loads issued up front, values shared through
// LDS only), then takes its share of the step-4 GEMV rows, one wavefront per row.  After a panel the
// trailing block takes the rank-2 NB update in one pass over it (k_panel_update, which also
// finishes the panel's last W column and writes the reflectors back).  V and W are stored
// column-major (column l contiguous over the rows), so that the column kernel's per-l loads
// across a wavefront are coalesced.  The reflectors are
// left where LAPACK leaves them for the lower variant (v(i+2:n) in A(i+2:n, i) of the
// column-major view, which is row i of this row-major storage), where LAPACK's dormtr and
// this library's yg_backtransform_cols read them.  The input's lower triangle (the upper triangle of the column-major view that
// dsyevd with fill_upper reads) is mirrored first.
#include "ygpu_internal.h"

#include <algorithm>

namespace {

constexpr int NB = 8;            // panel width (4, 6, 12 and 16 measured slower at n = 300..1560)
constexpr int ST = 512;          // column step: threads per workgroup (four rows each at n = 2048)
constexpr int LGRID = 128;       // workgroups of the column step (each repeats the column work, then a
                                 // 1/128 share of the trailing GEMV; 64 and 256 measured slower at n = 1560)
constexpr int UT = 64;           // trailing update: C tile edge

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

// wavefront sum, the same bits in every lane (row-of-16 butterflies, then lanes 0/16/32/48)
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}

// sum over each row of 16 lanes (every lane of the row receives it)
__device__ __forceinline__ double row16_sum(double v)
{
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v;
}

// workgroup sum over NT threads, returned to all; red holds NT / 64 values
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) s += red[q];
    __syncthreads();
    return s;
}

// A(r, c) = A(c, r) from the lower triangle (row-major r > c) into the upper
__global__ void k_mirror_lower(int n, long lda, double *__restrict__ A)
{
    const long tot = (long) n * n;
    for (long idx = (long) blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (long) gridDim.x * blockDim.x) {
        const int r = (int) (idx / n), c = (int) (idx % n);
        if (c > r) A[(size_t) r * lda + c] = A[(size_t) c * lda + r];
    }
}

// One column i = p + J of a panel, by every workgroup of the grid (specialised on the panel
// position J, so that only the J - 1 stored columns of V and W are loaded).  Each workgroup computes
// the column's steps 1-3 and the products W^T v, V^T v for all rows (the same operations in
// the same order, so every workgroup holds the same v), then its share of the trailing
// GEMV y = A0 v - V (W^T v) - W (V^T v).  Workgroup 0 also writes W(:, J - 1), V(:, J) and
// d, e, tau.  vin / yin are the previous column's v and y, vout / yout this column's
// (separate buffers: other workgroups still read the inputs).  Row i of A receives the
// reflector only at the end of the panel (k_panel_update).  Thread t owns rows
// i + t + ST k (k < 4); V and W are read once per row: with v = 1 at row i + 1 and
// v = x scale below, W^T v = scale (W^T x)_{>= i+2} + W(i + 1, :), accumulated in the same
// pass that forms x.
template <int J>
__global__ __launch_bounds__(ST) void k_latrd_step(int n, long lda, const double *__restrict__ A, int i,
                                                   double *__restrict__ V, double *__restrict__ W,
                                                   const double *__restrict__ vin, const double *__restrict__ yin,
                                                   double *__restrict__ vout, double *__restrict__ yout,
                                                   double *__restrict__ d, double *__restrict__ e,
                                                   double *__restrict__ tau)
{
    constexpr int RK = YG_SYTRD_BLOCKED_NMAX / ST;   // rows per thread
    constexpr int NS = ST / 16;                      // row-of-16 partial sums per workgroup
    __shared__ double red[ST / 64];
    __shared__ double red2[NS][2 * NB];
    __shared__ double sv[YG_SYTRD_BLOCKED_NMAX], sw[YG_SYTRD_BLOCKED_NMAX];   // v_i and W(:, J - 1)
    __shared__ double svi[NB], swi[NB], sv1[NB], sw1[NB], sab[2 * NB];       // V, W rows i and i + 1
    __shared__ double sx[2];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const bool lead = blockIdx.x == 0;
    constexpr int JP = J - 1;                   // the previous column (V(:, JP) = vin, W(:, JP) from yin)
    constexpr int NL = J > 1 ? J - 1 : 1, NP = J > 0 ? J : 1;
    // ---- loads that do not depend on this kernel's results
    double a_ir[RK], y_r[RK], vp[RK], wp[RK];
#pragma unroll
    for (int k = 0; k < RK; ++k) {
        const int r = i + t + ST * k;
        const bool in = r < n;
        a_ir[k] = in ? A[(size_t) i * lda + r] : 0.0;
        y_r[k] = (in && J > 0) ? yin[r] : 0.0;
        vp[k] = (in && J > 0) ? vin[r] : 0.0;
        wp[k] = 0.0;
    }
    if (t < JP) {
        svi[t] = V[(size_t) t * n + i];
        swi[t] = W[(size_t) t * n + i];
    }
    const double tp = (J > 0) ? tau[i - 1] : 0.0;
    // ---- step 1: W(:, JP) = tau y - (tau^2 / 2)(y . v) v from the previous column
    if constexpr (J > 0) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < RK; ++k) s += y_r[k] * vp[k];
        s = block_sum<ST>(s, red);              // also publishes svi, swi
        const double c = -0.5 * tp * tp * s;
#pragma unroll
        for (int k = 0; k < RK; ++k) {
            const int r = i + t + ST * k;
            wp[k] = tp * y_r[k] + c * vp[k];
            if (r < n) {
                sw[r] = wp[k];
                if (lead) W[(size_t) JP * n + r] = wp[k];
            }
        }
        if (t == 0) {
            svi[JP] = vp[0];                    // V(i, JP) (= 1)
            swi[JP] = wp[0];                    // W(i, JP)
        }
        __syncthreads();
    }
    // ---- step 2: x_r = A0(i, r) - sum_l (V(r, l) W(i, l) + W(r, l) V(i, l)), r >= i, and the
    // partial sums of W^T x, V^T x over r >= i + 2
    double x[RK], ss = 0.0;
    double pa[NP], pb[NP];
#pragma unroll
    for (int l = 0; l < NP; ++l) pa[l] = pb[l] = 0.0;
#pragma unroll
    for (int k = 0; k < RK; ++k) {
        const int r = i + t + ST * k;
        double v = a_ir[k];
        if (r < n) {
            // all of the row's panel loads in flight at once
            double vl[NL], wl[NL];
#pragma unroll
            for (int l = 0; l < JP; ++l) {
                vl[l] = V[(size_t) l * n + r];
                wl[l] = W[(size_t) l * n + r];
            }
#pragma unroll
            for (int l = 0; l < JP; ++l) v -= vl[l] * swi[l] + wl[l] * svi[l];
            if constexpr (J > 0) v -= vp[k] * swi[JP] + wp[k] * svi[JP];
            if (r == i) sx[0] = v;
            if (r == i + 1) {
                sx[1] = v;
#pragma unroll
                for (int l = 0; l < JP; ++l) {
                    sv1[l] = vl[l];
                    sw1[l] = wl[l];
                }
                if constexpr (J > 0) {
                    sv1[JP] = vp[k];
                    sw1[JP] = wp[k];
                }
            }
            if (r >= i + 2) {
                ss += v * v;
#pragma unroll
                for (int l = 0; l < JP; ++l) {
                    pa[l] += wl[l] * v;
                    pb[l] += vl[l] * v;
                }
                if constexpr (J > 0) {
                    pa[JP] += wp[k] * v;
                    pb[JP] += vp[k] * v;
                }
            }
        }
        x[k] = v;
    }
    // row-of-16 sums of the partial products (their last adds go through LDS)
#pragma unroll
    for (int l = 0; l < J; ++l) {
        pa[l] = row16_sum(pa[l]);
        pb[l] = row16_sum(pb[l]);
    }
    if ((lane & 15) == 0) {
        const int slot = wv * 4 + (lane >> 4);
#pragma unroll
        for (int l = 0; l < J; ++l) {
            red2[slot][l] = pa[l];
            red2[slot][NB + l] = pb[l];
        }
    }
    ss = block_sum<ST>(ss, red);                // its barriers also publish sx, sv1, sw1, red2
    // ---- step 3: dlarfg
    const double diag = sx[0], alpha = sx[1];
    double tk = 0.0, beta = alpha, scale = 0.0;
    if (ss > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        tk = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    double v_r[RK];
#pragma unroll
    for (int k = 0; k < RK; ++k) {
        const int r = i + t + ST * k;
        v_r[k] = (r == i + 1) ? 1.0 : (r >= i + 2 && r < n) ? x[k] * scale : 0.0;
        if (r < n) sv[r] = v_r[k];
    }
    if (t < 2 * NB) {
        const int l = t % NB;
        double s = 0.0;
        if (l < J) {
#pragma unroll
            for (int q = 0; q < NS; ++q) s += red2[q][t];
            s = scale * s + ((t < NB) ? sw1[l] : sv1[l]);
        }
        sab[t] = s;
    }
    __syncthreads();                            // sv and sab complete
    // ---- this workgroup's rows of the trailing GEMV: wavefront g of the grid takes rows
    // i + 1 + g, i + 1 + g + (ST / 64) G, ...
    const int G = gridDim.x;
    for (int r = i + 1 + blockIdx.x * (ST / 64) + wv; r < n; r += G * (ST / 64)) {
        const double *row = A + (size_t) r * lda;
        double s0 = 0.0, s1 = 0.0;
        int c = i + 1 + lane;
#pragma unroll 4
        for (; c + 64 < n; c += 128) {
            s0 += row[c] * sv[c];
            s1 += row[c + 64] * sv[c + 64];
        }
        if (c < n) s0 += row[c] * sv[c];
        double s = s0 + s1;
        if (lane < J) {
            const double vl = (lane < JP) ? V[(size_t) lane * n + r] : vin[r];
            const double wl = (lane < JP) ? W[(size_t) lane * n + r] : sw[r];
            s -= vl * sab[lane] + wl * sab[NB + lane];
        }
        s = wave_sum(s);
        if (lane == 0) yout[r] = s;
    }
    // ---- workgroup 0: this column's outputs
    if (lead) {
#pragma unroll
        for (int k = 0; k < RK; ++k) {
            const int r = i + t + ST * k;
            if (r < n) {
                vout[r] = v_r[k];
                if (r >= i + 1) V[(size_t) J * n + r] = v_r[k];
            }
        }
        if (t == 0) {
            d[i] = diag;
            e[i] = beta;
            tau[i] = tk;
        }
    }
}

// End of a panel p .. p + nbe - 1 and its trailing update, in one launch.  Every workgroup
//   * forms W(:, nbe - 1) = tau y - (tau^2 / 2)(y . v) v of the panel's last column (v = vin,
//     y = yin, rows >= q = p + nbe; the dot product repeated per workgroup) where its tile needs it,
//   * updates its 64 x 64 tile of the trailing block C = A(q:, q:) -= V W^T + W V^T (V, W
//     column-major with leading dimension n, 16 x 16 threads, a 4 x 4 block per thread, the
//     tile's rows and columns of V and W staged in LDS),
//   * writes its share of the panel's reflectors into their rows of A (e on the subdiagonal, v
//     beyond; rows p .. q - 1, which the update does not touch and which the column kernels left
//     alone while other workgroups still read them).
__global__ __launch_bounds__(256) void k_panel_update(int n, long lda, double *__restrict__ A, int p, int nbe,
                                                      const double *__restrict__ V, const double *__restrict__ W,
                                                      const double *__restrict__ vin, const double *__restrict__ yin,
                                                      const double *__restrict__ e, const double *__restrict__ tau)
{
    __shared__ double vr[NB][UT], wr[NB][UT], vc[NB][UT], wc[NB][UT];
    __shared__ double red[4];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int q = p + nbe, m = n - q;
    const int r0 = blockIdx.y * UT, c0 = blockIdx.x * UT;
    const int rb = r0 + 4 * ty, cb = c0 + 4 * tx;
    double *C = A + (size_t) q * lda + q;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            acc[a][b] = (rb + a < m && cb + b < m) ? C[(size_t) (rb + a) * lda + cb + b] : 0.0;
    // W(:, nbe - 1) from the last column's y and v
    const double tp = tau[q - 1];
    double sdot = 0.0;
    for (int r = q + t; r < n; r += 256) sdot += yin[r] * vin[r];
    sdot = block_sum<256>(sdot, red);
    const double cl = -0.5 * tp * tp * sdot;
    const int lw = nbe - 1;
    for (int idx = t; idx < nbe * UT; idx += 256) {
        const int l = idx / UT, k = idx % UT;
        const int rr = q + r0 + k, rc = q + c0 + k;
        const bool rin = r0 + k < m, cin = c0 + k < m;
        vr[l][k] = rin ? V[(size_t) l * n + rr] : 0.0;
        vc[l][k] = cin ? V[(size_t) l * n + rc] : 0.0;
        if (l == lw) {
            wr[l][k] = rin ? tp * yin[rr] + cl * vin[rr] : 0.0;
            wc[l][k] = cin ? tp * yin[rc] + cl * vin[rc] : 0.0;
        } else {
            wr[l][k] = rin ? W[(size_t) l * n + rr] : 0.0;
            wc[l][k] = cin ? W[(size_t) l * n + rc] : 0.0;
        }
    }
    __syncthreads();
#pragma unroll 4
    for (int l = 0; l < nbe; ++l) {
        double va[4], wa[4], vb[4], wb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            va[u] = vr[l][4 * ty + u];
            wa[u] = wr[l][4 * ty + u];
            vb[u] = vc[l][4 * tx + u];
            wb[u] = wc[l][4 * tx + u];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] -= va[a] * wb[b] + wa[a] * vb[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (rb + a < m && cb + b < m) C[(size_t) (rb + a) * lda + cb + b] = acc[a][b];
    // this workgroup's share of the reflector rows
    const long G = (long) gridDim.x * gridDim.y, wg = (long) blockIdx.y * gridDim.x + blockIdx.x;
    const long len = n - p, tot = (long) nbe * len;
    for (long idx = wg * 256 + t; idx < tot; idx += G * 256) {
        const int j = (int) (idx / len), r = p + (int) (idx % len), ic = p + j;
        if (r > ic) A[(size_t) ic * lda + r] = (r == ic + 1) ? e[ic] : V[(size_t) j * n + r];
    }
}

// the trailing 2 x 2 block after the last reflector
__global__ void k_latrd_tail(int n, long lda, double *__restrict__ A, double *__restrict__ d, double *__restrict__ e,
                             double *__restrict__ tau)
{
    if (threadIdx.x != 0) return;
    d[n - 2] = A[(size_t) (n - 2) * lda + n - 2];
    e[n - 2] = A[(size_t) (n - 1) * lda + n - 2];
    d[n - 1] = A[(size_t) (n - 1) * lda + n - 1];
    tau[n - 2] = 0.0;
    A[(size_t) (n - 2) * lda + n - 1] = e[n - 2];
}

// k_latrd_step<J> for the panel position j (0 <= j < NB)
template <int J>
void launch_step(hipStream_t s, int j, int n, long lda, const double *A, int i, double *V, double *W,
                 const double *vin, const double *yin, double *vout, double *yout, double *d, double *e, double *tau)
{
    if (j == J)
        hipLaunchKernelGGL(k_latrd_step<J>, dim3(LGRID), dim3(ST), 0, s, n, lda, A, i, V, W, vin, yin, vout, yout, d,
                           e, tau);
    else if constexpr (J + 1 < NB)
        launch_step<J + 1>(s, j, n, lda, A, i, V, W, vin, yin, vout, yout, d, e, tau);
}

}  // namespace

size_t yg_sytrd_blocked_ws(int n) { return ((size_t) 2 * n * NB + 4 * (size_t) n) * sizeof(double); }

int yg_sytrd_blocked(hipStream_t s, int n, long lda, double *A, double *d, double *e, double *tau, double *ws)
{
    if (n < 3 || n > YG_SYTRD_BLOCKED_NMAX || lda < n) {
        ygpu_set_error(YGPU_ERR_INVALID, "blocked tridiagonal reduction: n = %d (3..%d), lda = %ld", n,
                       YG_SYTRD_BLOCKED_NMAX, lda);
        return YGPU_ERR_INVALID;
    }
    double *V = ws, *W = ws + (size_t) n * NB;
    double *vb[2] = {W + (size_t) n * NB, W + (size_t) n * NB + n};
    double *yb[2] = {vb[1] + n, vb[1] + 2 * (size_t) n};
    hipLaunchKernelGGL(k_mirror_lower, dim3(std::min<long>(((long) n * n + 255) / 256, 4096)), dim3(256), 0, s, n,
                       lda, A);
    YG_HIP(hipGetLastError());
    // no clearing of V and W per panel: column l of a panel starting at p is written for rows
    // >= p + l + 1 before any kernel reads it, and every read is of such rows
    const int last = n - 3;                     // reflectors 0 .. n - 3
    int cur = 0;                                // vb[cur], yb[cur]: the previous column's v and y
    for (int p = 0; p <= last; p += NB) {
        const int nbe = std::min(NB, last - p + 1);
        for (int j = 0; j < nbe; ++j) {
            const int i = p + j;
            launch_step<0>(s, j, n, lda, A, i, V, W, vb[cur], yb[cur], vb[cur ^ 1], yb[cur ^ 1], d, e, tau);
            YG_HIP(hipGetLastError());
            cur ^= 1;
        }
        // the panel's last W column, trailing update of rows / columns >= q and the reflector rows
        const int q = p + nbe, m = n - q;
        const int nt = (m + UT - 1) / UT;
        hipLaunchKernelGGL(k_panel_update, dim3(nt, nt), dim3(256), 0, s, n, lda, A, p, nbe, (const double *) V,
                           (const double *) W, (const double *) vb[cur], (const double *) yb[cur],
                           (const double *) e, (const double *) tau);
        YG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_latrd_tail, dim3(1), dim3(64), 0, s, n, lda, A, d, e, tau);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}
