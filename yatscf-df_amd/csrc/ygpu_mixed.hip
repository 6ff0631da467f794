// ygpu_mixed.hip -- reduced-precision K build for the mixed-precision SCF phase.
//
// The two K products of build_Fock.c:201-347,
//   X_i = C_occ(partners of i)^T B(rows of i)   ("aux. K tensor", :249-285)
//   K   = X X^T                                  ("K matrix", :317-341)
// run as BF16-split products on the gfx950 matrix cores: every FP32 operand
// x is split into NP BF16 parts, p_0 = bf16(x), p_1 = bf16(x - p_0), ..., and
// a*b is formed from the part products p_a * q_b with a + b < NP on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation.  NP = 2 ("BF16x3"): three
// products, ~16 significant bits per operand (the dropped lo*lo term is
// ~2^-18 relative).  NP = 3 ("BF16x6"): six products; the three parts hold an
// FP32 value exactly and the dropped terms are ~2^-24 relative, so the
// products are as accurate as FP32 ones, at 2.5 PF/s / 6 against the 157
// TF/s of the FP32 matrix instructions.
//
// Tile engine: WM x WN wavefronts (up to 16), each owning a 64x64 output
// sub-tile as 2x2 MFMA blocks of 32x32; the workgroup tile is (64*WM) x
// (64*WN).  Operands are staged global -> registers (FP32) -> split -> LDS as
// k-contiguous rows of 8 BF16 values, so every MFMA operand is one 16-byte
// LDS read; the next k-stage's global loads are issued before the current
// stage's MFMAs.  The X build takes all occupied rows (up to 256) in one
// workgroup so every B tile is read once per basis function.
//
// Layouts of v_mfma_f32_32x32x16_bf16 on gfx950 (measured): lane l holds
// A[l%32][k] and B[k][l%32] for the eight k of half l/32 of the 16-deep step;
// output element v of lane l is D[8*(v/4) + 4*(l/32) + v%4][l%32].
#include "ygpu_internal.h"

#include <algorithm>
#include <type_traits>

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int NTHR = 256;     // SYRK workgroup (2 x 2 wavefronts)
// k per LDS stage (BK = 32 or 64: two or four 16-deep MFMA steps); rows are stored as
// BK / 8 groups of 8 BF16 values plus one group of padding
template <int BK> struct Stage {
    static constexpr int KV = BK / 8;
    static constexpr int KVP = KV + 1;
};
// SYRK stage depth: 64 for two parts (4 wavefronts have the registers for it); 32 for three,
// which keeps the stage's LDS (6 arrays) small enough for two workgroups per CU
template <int NP> constexpr int syrk_bk() { return NP == 2 ? 64 : 32; }
constexpr int XBK = 32;   // X-build stage depth (12-16 wavefronts per workgroup)
constexpr int XNBUF = 1;  // X-build LDS stages (two measured no faster: the loads, not the split, set the pace)

// x = p[0] + p[1] + ... (+ the remainder below the last part), each part the bf16 rounding
// (round to nearest even) of what the earlier parts leave; the subtractions are exact in FP32
template <int NP>
__device__ __forceinline__ void splitN(const float (&x)[8], v8bf (&p)[NP])
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float r = x[e];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const __bf16 h = (__bf16) r;
            p[q][e] = h;
            r -= (float) h;
        }
    }
}

// acc[i][j] += A(64 rows of this wave) * B(64 cols of this wave) over k in [kbeg, kend).
// Loader contract: coords(idx, row, kv) names the tile row and 8-group of the
// idx-th (row, kv) slot (idx < tile rows * KV); load(row, k8, kend, v) fills the
// 8 values k8 .. k8+7 of that row (zero beyond kend or outside the matrix), or for loaders
// whose data is stored pre-split (kPresplit) loadp(row, k8, kend, parts) its NP BF16 parts.
// NBUF = 1: one LDS stage, split and MFMA phases separated by two barriers per stage.
// NBUF = 2: two LDS stages; stage s + 1 is split into one buffer while the MFMAs of
// stage s read the other, with one barrier per stage, so the split of some wavefronts
// overlaps the matrix-core work of others.
// mrows / ncols: rows and columns of the tile inside the matrix; with SKIP a wavefront whose
// 64 x 64 part lies wholly outside issues no MFMAs (its accumulators stay zero and are never
// stored).  The SYRK keeps SKIP off: the test costs it the registers of its second workgroup per CU.
// Part products are issued smallest first: for NP = 2 lo*hi, hi*lo, hi*hi; for NP = 3 the
// products of part-index sum 2, then 1, then 0.
template <int NP, int WM, int WN, int BK, int NBUF, bool SKIP, class LA, class LB>
__device__ __forceinline__ void bfs_tile(const LA &la, const LB &lb, int kbeg, int kend, v16f (&acc)[2][2],
                                         int mrows, int ncols, bool zero = true)
{
    constexpr int KV = Stage<BK>::KV, KVP = Stage<BK>::KVP;
    constexpr int NT = 64 * WM * WN;
    constexpr int BMt = 64 * WM, BNt = 64 * WN;
    constexpr int SA = BMt * KV, SB = BNt * KV;                    // (row, 8-group) slots per stage
    constexpr int GA = (SA + NT - 1) / NT, GB = (SB + NT - 1) / NT;
    __shared__ v8bf As[NBUF][NP][BMt][KVP], Bs[NBUF][NP][BNt][KVP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w / WN) * 64, wn = (w % WN) * 64;
    if (zero) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    }
    // slot g of this thread is idx = threadIdx.x + NT g; slots beyond SA / SB are idle
    int rowa[GA], kva[GA], rowb[GB], kvb[GB];
    bool ina[GA], inb[GB];
#pragma unroll
    for (int g = 0; g < GA; ++g) {
        const int idx = threadIdx.x + NT * g;
        ina[g] = idx < SA;
        la.coords(ina[g] ? idx : 0, rowa[g], kva[g]);
    }
#pragma unroll
    for (int g = 0; g < GB; ++g) {
        const int idx = threadIdx.x + NT * g;
        inb[g] = idx < SB;
        lb.coords(inb[g] ? idx : 0, rowb[g], kvb[g]);
    }
    // FP32 staging registers (split into parts at the LDS store), or for pre-split data the
    // BF16 parts themselves
    float ra[GA][8], rb[GB][8];
    v8bf rap[GA][NP], rbp[GB][NP];
    auto load_stage = [&](int k0) {
#pragma unroll
        for (int g = 0; g < GA; ++g)
            if (ina[g]) {
                if constexpr (LA::kPresplit) la.loadp(rowa[g], k0 + 8 * kva[g], kend, rap[g]);
                else la.load(rowa[g], k0 + 8 * kva[g], kend, ra[g]);
            }
#pragma unroll
        for (int g = 0; g < GB; ++g)
            if (inb[g]) {
                if constexpr (LB::kPresplit) lb.loadp(rowb[g], k0 + 8 * kvb[g], kend, rbp[g]);
                else lb.load(rowb[g], k0 + 8 * kvb[g], kend, rb[g]);
            }
    };
    auto split_stage = [&](int buf) {
#pragma unroll
        for (int g = 0; g < GA; ++g)
            if (ina[g]) {
                if constexpr (!LA::kPresplit) splitN<NP>(ra[g], rap[g]);
#pragma unroll
                for (int q = 0; q < NP; ++q) As[buf][q][rowa[g]][kva[g]] = rap[g][q];
            }
#pragma unroll
        for (int g = 0; g < GB; ++g)
            if (inb[g]) {
                if constexpr (!LB::kPresplit) splitN<NP>(rb[g], rbp[g]);
#pragma unroll
                for (int q = 0; q < NP; ++q) Bs[buf][q][rowb[g]][kvb[g]] = rbp[g][q];
            }
    };
    // a wavefront with no row or no column inside the matrix skips its matrix-core work
    const bool wskip = SKIP && (mrows <= wm || ncols <= wn);
    auto mfma_stage = [&](int buf) {
        if (wskip) return;
#pragma unroll
        for (int s = 0; s < BK / 16; ++s) {
            const int kv = 2 * s + (lane >> 5), r = lane & 31;
            v8bf a[2][NP], b[2][NP];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    a[i][q] = As[buf][q][wm + 32 * i + r][kv];
                    b[i][q] = Bs[buf][q][wn + 32 * i + r][kv];
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int sum = NP - 1; sum >= 0; --sum)
#pragma unroll
                        for (int pa = sum; pa >= 0; --pa)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i][pa], b[j][sum - pa], acc[i][j],
                                                                               0, 0, 0);
        }
    };
    if (kbeg >= kend) return;
    load_stage(kbeg);
    if constexpr (NBUF == 1) {
        for (int k0 = kbeg; k0 < kend; k0 += BK) {
            __syncthreads();
            split_stage(0);
            __syncthreads();
            if (k0 + BK < kend) load_stage(k0 + BK);
            mfma_stage(0);
        }
    } else {
        split_stage(0);
        if (kbeg + BK < kend) load_stage(kbeg + BK);
        __syncthreads();
        int cur = 0;
        for (int k0 = kbeg; k0 < kend; k0 += BK) {
            // buffer cur ^ 1 was last read by the MFMAs before the previous barrier
            if (k0 + BK < kend) {
                split_stage(cur ^ 1);
                if (k0 + 2 * BK < kend) load_stage(k0 + 2 * BK);
            }
            mfma_stage(cur);
            __syncthreads();
            cur ^= 1;
        }
    }
}

template <int WN, class F>
__device__ __forceinline__ void bfs_epilogue(const v16f (&acc)[2][2], F &&store)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w / WN) * 64, wn = (w % WN) * 64;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                store(wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), wn + 32 * j + (lane & 31), acc[i][j][v]);
}

// workgroup id -> logical tile id such that consecutive logical ids run on one
// XCD (the hardware deals workgroups round-robin over the 8 XCDs), so tiles
// that share an operand panel share that XCD's L2
__device__ __forceinline__ long xcd_remap(long b, long total)
{
    const long per = total / 8, rem = total % 8, x = b % 8, slot = b / 8;
    return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

// ---- X build: A(m,k) = Cocc32[partner[k]][m], B(k,n) = B32[up[k]][n]; the gather
// indices of the current k chunk sit in LDS (pa[k - koff], up[k - koff]), so a stage's data
// loads wait on one memory latency, not on an index load followed by a data load
template <int BMt>
struct XLoadA {
    static constexpr bool kPresplit = false;
    const float *cocc; const int *pa; int koff, n_occ, m0;
    __device__ void coords(int idx, int &row, int &kv) const { row = idx % BMt; kv = idx / BMt; }
    __device__ void load(int row, int k8, int kend, float (&v)[8]) const {
        const int m = m0 + row;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k8 + e;
            v[e] = (k < kend && m < n_occ) ? cocc[(size_t) pa[k - koff] * n_occ + m] : 0.f;
        }
    }
};

template <int BNt>
struct XLoadB {
    static constexpr bool kPresplit = false;
    const float *b32; const int *up; int koff; long ld; int n0;
    __device__ void coords(int idx, int &row, int &kv) const { row = idx % BNt; kv = idx / BNt; }
    __device__ void load(int row, int k8, int kend, float (&v)[8]) const {
        const long n = n0 + row;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k8 + e;
            v[e] = (k < kend && n < ld) ? b32[(size_t) up[k - koff] * ld + n] : 0.f;
        }
    }
};

constexpr int XKCH = 256;    // gather indices staged in LDS per k chunk (one index latency per 8 stages)

// X build: (64 WM occupied rows) x (64 WN aux columns) per workgroup of WM x WN wavefronts;
// X stored pre-split, part q at Xp + q * xpart
template <int NP, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_xbuild_bf(const int *__restrict__ xptr,
                                                            const int *__restrict__ xup,
                                                            const int *__restrict__ xpartner,
                                                            const float *__restrict__ Cocc32,
                                                            const float *__restrict__ B32, __bf16 *__restrict__ Xp,
                                                            size_t xpart, int n_occ, int naux_pad, int ncolt,
                                                            int nrowt, long total)
{
    // column tiles fastest: consecutive workgroups read neighbouring segments of the same B
    // rows (measured faster than grouping the row tiles of one panel on an XCD)
    const long t = blockIdx.x;
    const int ct = (int) (t % ncolt), rt = (int) ((t / ncolt) % nrowt), i = (int) (t / ((long) nrowt * ncolt));
    const int kb = xptr[i], klen = xptr[i + 1] - kb;
    const int m0 = rt * 64 * WM, n0 = ct * 64 * WN;
    __shared__ int s_pa[XKCH], s_up[XKCH];
    v16f acc[2][2];
    for (int kc = 0; kc < klen || kc == 0; kc += XKCH) {
        const int kce = min(klen, kc + XKCH);
        if (kc > 0) __syncthreads();            // the previous chunk's loads are done with the indices
        for (int k = kc + (int) threadIdx.x; k < kce; k += 64 * WM * WN) {
            s_pa[k - kc] = xpartner[kb + k];
            s_up[k - kc] = xup[kb + k];
        }
        __syncthreads();
        XLoadA<64 * WM> la{Cocc32, s_pa, kc, n_occ, m0};
        XLoadB<64 * WN> lb{B32, s_up, kc, naux_pad, n0};
        bfs_tile<NP, WM, WN, XBK, XNBUF, true>(la, lb, kc, kce, acc, n_occ - m0, naux_pad - n0, kc == 0);
        if (kce >= klen) break;
    }
    // X stored pre-split for the SYRK, exactly the parts the SYRK's own split of the FP32
    // value would give
    const size_t xo = (size_t) i * n_occ * naux_pad;
    bfs_epilogue<WN>(acc, [&](int r, int c, float v) {
        const int m = m0 + r, n = n0 + c;
        if (m < n_occ && n < naux_pad) {
            float rem = v;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const __bf16 h = (__bf16) rem;
                Xp[(size_t) q * xpart + xo + (size_t) m * naux_pad + n] = h;
                rem -= (float) h;
            }
        }
    });
}

// ---- SYRK: A(m,k) = X[m][k], B(k,n) = X[n][k], both k-contiguous, X pre-split into NP BF16
// arrays by the X build: a stage moves 16-byte groups straight into LDS.  Rows are ld apart
// (ld = kdim for the SYRK's X; L_i for the blocks of the pre-split X build)
template <int NP, int BK>
struct SLoad {
    static constexpr bool kPresplit = true;
    const __bf16 *xp; size_t xpart; long ld; int nbf, r0;
    __device__ void coords(int idx, int &row, int &kv) const
    {
        kv = idx % Stage<BK>::KV;
        row = idx / Stage<BK>::KV;
    }
    __device__ void loadp(int row, int k8, int kend, v8bf (&p)[NP]) const {
        const int r = r0 + row;
        if (r < nbf && k8 < kend) {
#pragma unroll
            for (int q = 0; q < NP; ++q) p[q] = *(const v8bf *) (xp + (size_t) q * xpart + (size_t) r * ld + k8);
        } else {
#pragma unroll
            for (int q = 0; q < NP; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) p[q][e] = (__bf16) 0.f;
        }
    }
};

// ---- two-part X build from panel-major operand blocks.  For basis function i with klen_i pair
// rows, L_i = klen_i rounded up to 32 and kofs_i = sum_{j<i} L_j.  Block i of a source with R
// rows (rows r of the operand, k = the pair rows of i) holds the two BF16 parts of
// src[idx[xptr_i + kk]][r] at dst[q * dpart + kofs_i * R + (kk / 32) * 32 R + 32 r + kk % 32]
// (zero for kk >= klen_i): 32-deep k panels, each panel the R rows one after the other.  B (idx =
// the pair rows) is laid out once per DF tensor, C_occ (idx = the partners) once per build.  A
// 32-deep stage of any row range is then one contiguous stretch of memory, which the X build
// reads with whole cache lines (k-contiguous rows read 32 or 64 bytes per row and kept the
// texture path, not the matrix cores, busy).
__global__ __launch_bounds__(256) void k_gather_split_t(const float *__restrict__ src, long ld, int R,
                                                        const int *__restrict__ idx, const int *__restrict__ xptr,
                                                        const int *__restrict__ kofs, const int2 *__restrict__ ktiles,
                                                        int subs, __bf16 *__restrict__ dst, size_t dpart)
{
    // workgroup (64-deep k tile (i, k0) of the list, subs x 64 operand rows in 64-row passes)
    const int2 kt = ktiles[blockIdx.x];
    const int i = kt.x, k0 = kt.y;
    const int kb = xptr[i], klen = xptr[i + 1] - kb, L = (klen + 31) & ~31;
    __bf16 *d0 = dst + (size_t) kofs[i] * R, *d1 = d0 + dpart;
    __shared__ float tile[64][65];
    __shared__ int rows[64];
    if (threadIdx.x < 64) rows[threadIdx.x] = k0 + (int) threadIdx.x < klen ? idx[kb + k0 + threadIdx.x] : -1;
    __syncthreads();
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // store mapping: lane t writes the 8-value group g = t % 4 of row t / 4 (of 64), so the 64 lanes
    // of a wavefront store 16 rows x 64 bytes, one contiguous KiB of a panel
    const int g = threadIdx.x & 3, rw = threadIdx.x >> 2;
    for (int sub = 0; sub < subs; ++sub) {
        const int r0 = (blockIdx.y * subs + sub) * 64;
        if (r0 >= R) break;
#pragma unroll 4
        for (int kk = ty; kk < 64; kk += 4) {
            const int row = rows[kk];
            tile[kk][tx] = (row >= 0 && r0 + tx < R) ? src[(size_t) row * ld + r0 + tx] : 0.f;
        }
        __syncthreads();
        const int r = r0 + rw;
        if (r < R) {
#pragma unroll
            for (int pp = 0; pp < 2; ++pp) {
                const int k = k0 + 32 * pp;
                if (k >= L) break;
                v8bf h, l;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = tile[32 * pp + 8 * g + e][rw];
                    h[e] = (__bf16) x;
                    l[e] = (__bf16) (x - (float) h[e]);
                }
                const size_t off = (size_t) (k >> 5) * R * 32 + (size_t) r * 32 + 8 * g;
                *(v8bf *) (d0 + off) = h;
                *(v8bf *) (d1 + off) = l;
            }
        }
        __syncthreads();
    }
}

#define YG_LDSP(p) ((__attribute__((address_space(3))) void *) (p))

// LDS ring of the panel engine: NS stages of one 32-deep k panel of the (64 BI)-row A range and
// the (32 BJ WN)-row B range, two BF16 parts each, in 16-byte slots; 2 WN wavefronts
template <int BI, int BJ, int WN> struct PsStage {
    static constexpr int NW = 2 * WN, BM = 64 * BI, BN = 32 * BJ * WN;
    static constexpr int SA = 2 * BM * 4, SB = 2 * BN * 4, SLOTS = SA + SB, PW = SLOTS / (64 * NW);
    static_assert(SA % (64 * NW) == 0 && SB % (64 * NW) == 0 && BN % 16 == 0,
                  "each operand part of a stage is whole wavefront copies of every wavefront");
};

// f(std::integral_constant<int, t>) for the t in -1 .. sizeof...(Ts) - 2 equal to the uniform t
template <class F, int... Ts>
__device__ __forceinline__ void ps_cut_dispatch(int t, F &&f, std::integer_sequence<int, Ts...>)
{
    (void) ((t == Ts - 1 ? (f(std::integral_constant<int, Ts - 1>{}), 0) : 0) + ...);
}

// column block of wavefront w's register tile.  The hardware places wavefront w of a workgroup on
// SIMD w % 4, so wavefronts w and w + WN (WN = 4) share a SIMD.  With mir (uniform) the second row
// half takes the column blocks in reverse, so a SIMD holds column blocks b and WN - 1 - b: the
// register tiles that the upper-triangle product skips (wholly below the diagonal, at the low
// column end of the second row half; beyond the columns, at the high end of both) fall on
// different SIMDs.  (Tiles without skips run ~2 % faster without mir.)
template <int WN> __device__ __forceinline__ int ps_colblk(int w, bool mir)
{
    return (!mir || w < WN) ? w % WN : 2 * WN - 1 - w;
}

// Panel engine (ring: NS stages of PsStage<BI, BJ, WN>::SLOTS slots in LDS): acc = A B^T for one
// (64 BI) x (32 BJ WN) tile (rows m0.. of A, n0.. of B) over the 32-deep k panels 0 .. nst-1.
// A and B are panel-major: panel p of an operand with R rows holds the R rows one after the
// other, 32 k each, at p * 32 R; its two BF16 parts are apart / bpart elements apart.  2 x WN
// wavefronts, each with a (32 BI) x (32 BJ) register tile of 32x32 MFMA blocks (WN = 2, BI = 5,
// BJ = 3: 240 accumulators, one wavefront per SIMD; WN = 4, BJ = 2: 160, two per SIMD), so every
// 16-byte LDS operand read feeds BJ (or BI) block products.  A stage is copied global -> LDS by
// the texture path's direct loads (global_load_lds_dwordx4, no registers), each wavefront
// instruction one contiguous KiB; NS stages in a ring, the copy of stage s + NS - 1 issued while
// stage s is multiplied.  In LDS a row's four 8-k groups sit at 4 r + (g XOR (r >> 1) & 3), which spreads the
// rows of one operand read over all banks.  Rows beyond an operand are read from its last row, and
// with two stages a copy of 16 rows wholly beyond it is not issued, its slots holding zeros: either
// way only results that are not stored see them.  Wavefront w's register tile is row half w / WN, column block
// ps_colblk<WN>(w, mir).  The wavefront multiplies only its 32x32 blocks (x, y) with x - y < cut
// (cut uniform; cut >= BI: all, the blocks wholly below the diagonal of a register tile whose first
// row is 32 (cut - 1) before its first column excluded); with cut < 0 it copies its share of every
// stage and meets every barrier but reads no operands and multiplies nothing.  Blocks not
// multiplied stay 0.
template <int BI, int BJ, int WN, int NS>
__device__ __forceinline__ void ps_tile(v8bf *ring, const __bf16 *__restrict__ A,
                                        size_t apart, int RA, int m0, const __bf16 *__restrict__ B, size_t bpart,
                                        int RB, int n0, int nst, bool mir, int cut, v16f (&acc)[BI][BJ])
{
    using P = PsStage<BI, BJ, WN>;
    constexpr int BM = P::BM, BN = P::BN, SA = P::SA, PW = P::PW, NW = P::NW;
    static_assert((NS - 2) * PW <= 63, "vmcnt range");
    // the wavefront index as a scalar: the LDS destination and operand panel of each of its
    // copies are uniform, only the row and group within the panel vary by lane
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (w / WN) * 32 * BI, wn = ps_colblk<WN>(w, mir) * 32 * BJ;
    // wavefront copy j of a stage fills slots 64 j .. 64 j + 63: the first SA / (64 NW) of each
    // wavefront's copies are A's, the rest B's.  Slot t = 64 q + lane of an operand part holds
    // row t / 4 = 16 q + lane / 4 and, at position t % 4 = lane % 4, its 8-k group
    // (lane % 4) XOR ((row / 2) % 4) = (lane % 4) XOR ((lane / 8) % 4)
    const int lr = lane >> 2, lg = 8 * ((lane & 3) ^ ((lane >> 3) & 3));
    // with two stages (every wait then for all copies in flight) a copy of 16 rows wholly beyond
    // its operand is not issued
    auto beyond = [&](int jj) {
        const int j = NW * jj + w;
        if (NS != 2) return false;
        if (jj < SA / (64 * NW)) return m0 + 16 * (j % (BM / 16)) >= RA;
        return n0 + 16 * ((j - SA / 64) % (BN / 16)) >= RB;
    };
    auto issue = [&](int slot, int p) {
        v8bf *dst = ring + slot * P::SLOTS;
#pragma unroll
        for (int jj = 0; jj < PW; ++jj) {
            const int j = NW * jj + w;
            const __bf16 *base;
            int row;
            if (beyond(jj)) continue;
            if (jj < SA / (64 * NW)) {
                const int part = j / (BM / 16), q = j % (BM / 16);
                base = A + part * apart + (size_t) p * RA * 32;
                row = min(m0 + 16 * q + lr, RA - 1);
            } else {
                const int j2 = j - SA / 64, part = j2 / (BN / 16), q = j2 % (BN / 16);
                base = B + part * bpart + (size_t) p * RB * 32;
                row = min(n0 + 16 * q + lr, RB - 1);
            }
            __builtin_amdgcn_global_load_lds((const void *) (base + (unsigned) (row * 32 + lg)), YG_LDSP(dst + 64 * j),
                                             16, 0, 0);
        }
    };
    // the copies not issued leave zeros in their slots (written once, here; the first stage's
    // wait and barrier order them before any read): stale bits there would feed the products
    // arbitrary operands, which cost clock
#pragma unroll
    for (int jj = 0; jj < PW; ++jj)
        if (beyond(jj))
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) ring[sl * P::SLOTS + 64 * (NW * jj + w) + lane] = v8bf{};
#pragma unroll
    for (int a = 0; a < BI; ++a)
#pragma unroll
        for (int b = 0; b < BJ; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    const int rl = lane & 31, hk = lane >> 5;
    for (int p = 0; p < NS - 1 && p < nst; ++p) issue(p, p);
    // the stage loop for one cut T (a copy of the loop per value: no tests inside)
    auto stages = [&](auto tc) {
        constexpr int T = decltype(tc)::value;
        for (int st = 0; st < nst; ++st) {
            // stage st landed (the later stages issued so far may still be in flight), every LDS
            // read of this wavefront done, then the workgroup: the ring slot written next was read
            // by all wavefronts in the previous stage
            const int later = min(NS - 2, nst - 1 - st);
            asm volatile("" ::: "memory");
            switch (later) {
#define YG_WAIT(n) \
    case n: __builtin_amdgcn_s_waitcnt((((n) * PW) & 15) | ((((n) * PW) >> 4) << 14) | (7 << 4)); break;
                YG_WAIT(0) YG_WAIT(1) YG_WAIT(2) YG_WAIT(3)
#undef YG_WAIT
            }
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (st + NS - 1 < nst) issue((st + NS - 1) % NS, st + NS - 1);
            if constexpr (T < 0) continue;
            const v8bf *sb = ring + (st % NS) * P::SLOTS;
            const v8bf *sa0 = sb, *sa1 = sb + BM * 4, *sb0 = sb + SA, *sb1 = sb0 + BN * 4;
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                v8bf a[BI][2], b[BJ][2];
#pragma unroll
                for (int x = 0; x < BI; ++x) {
                    if (x - (BJ - 1) >= T) continue;
                    const int r = wm + 32 * x + rl, off = r * 4 + ((2 * ks + hk) ^ ((r >> 1) & 3));
                    a[x][0] = sa0[off];
                    a[x][1] = sa1[off];
                }
#pragma unroll
                for (int y = 0; y < BJ; ++y) {
                    const int r = wn + 32 * y + rl, off = r * 4 + ((2 * ks + hk) ^ ((r >> 1) & 3));
                    b[y][0] = sb0[off];
                    b[y][1] = sb1[off];
                }
                // lo*hi, hi*lo, hi*hi: the order of the other two-part products
#pragma unroll
                for (int x = 0; x < BI; ++x)
#pragma unroll
                    for (int y = 0; y < BJ; ++y) {
                        if (x - y >= T) continue;
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x][1], b[y][0], acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x][0], b[y][1], acc[x][y], 0, 0, 0);
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x][0], b[y][0], acc[x][y], 0, 0, 0);
                    }
            }
        }
    };
    ps_cut_dispatch(std::min(cut, BI), stages, std::make_integer_sequence<int, BI + 2>{});
}

// store(m, n, v) for every element of this wavefront's part of a ps_tile result (m, n relative
// to the workgroup tile)
template <int BI, int BJ, int WN, class F>
__device__ __forceinline__ void ps_epilogue(const v16f (&acc)[BI][BJ], bool mir, F &&store)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w / WN) * 32 * BI, wn = ps_colblk<WN>(w, mir) * 32 * BJ, rl = lane & 31, hk = lane >> 5;
#pragma unroll
    for (int x = 0; x < BI; ++x)
#pragma unroll
        for (int y = 0; y < BJ; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                store(wm + 32 * x + 8 * (v >> 2) + 4 * hk + (v & 3), wn + 32 * y + rl, acc[x][y][v]);
}

// X_i (nrow x naux_pad) = Cp_i Bt_i^T over the L_i = round32(klen_i) pair rows of basis function
// i (k beyond klen_i is zero in both blocks).  X is stored as two BF16 parts (xpart apart),
// panel-major over the basis functions for the products that follow: element k = m naux_pad + n
// of function i at (k / 32) * 32 nbf + 32 i + k % 32, so a 32-deep k panel of any range of
// functions is one contiguous stretch.
template <int BI, int BJ, int WN, int NS>
__global__ __launch_bounds__(128 * WN, 1) void k_xbuild_ps(const int *__restrict__ xptr, const int *__restrict__ kofs,
                                                      const int *__restrict__ order, const __bf16 *__restrict__ Cp,
                                                      size_t cpart, const __bf16 *__restrict__ Bt, size_t btpart,
                                                      __bf16 *__restrict__ Xp, size_t xpart, int nrow, int naux_pad,
                                                      int ncolt, int nrowt, int nbf)
{
    using P = PsStage<BI, BJ, WN>;
    __shared__ __attribute__((aligned(16))) v8bf ring[NS][P::SLOTS];
    // basis functions in order of decreasing row length (order[]), dealt round-robin over the 8
    // XCDs (the hardware deals workgroup b to XCD b % 8), every tile of the o-th function on
    // XCD o % 8: the row lengths, which vary 4-5x between functions and run in long stretches
    // along the basis, even out over the XCDs, the longest run first, and the column tiles of a
    // function and the row tiles of a column tile share its blocks in one L2
    const long q = blockIdx.x / 8;
    const int tiles = ncolt * nrowt;
    const int o = (int) (q / tiles) * 8 + (int) (blockIdx.x % 8);
    if (o >= nbf) return;
    const int i = order[o];
    const int rem = (int) (q % tiles), ct = rem / nrowt, rt = rem % nrowt;
    const int L = (xptr[i + 1] - xptr[i] + 31) & ~31;
    const size_t ko = (size_t) kofs[i];
    const int m0 = rt * P::BM, n0 = ct * P::BN;
    v16f acc[BI][BJ];
    ps_tile<BI, BJ, WN, NS>(&ring[0][0], Cp + ko * nrow, cpart, nrow, m0, Bt + ko * naux_pad, btpart, naux_pad, n0,
                            L / 32, false, BI, acc);
    // the tile leaves through the (now free) ring: each wavefront writes one 32-row block of its
    // accumulators as FP32 rows into its own part of LDS, then reads them back as runs of 8
    // columns, splits them and stores 16 bytes of each part per lane (four lanes fill one 32-k
    // group of a row), instead of one 2-byte store per element and part
    constexpr int CW = 32 * BJ, LS = CW + 8;   // row stride: the two half-wavefronts' rows 4 apart land in different banks
    static_assert(P::NW * 32 * LS * 4 <= (int) sizeof(ring), "epilogue staging fits the ring");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, rl = lane & 31, hk = lane >> 5;
    const int wm = (w / WN) * 32 * BI, wn = (w % WN) * CW;
    float *stg = (float *) &ring[0][0] + w * 32 * LS;
    __bf16 *X0 = Xp + 32 * (size_t) i;
    __syncthreads();
#pragma unroll
    for (int x = 0; x < BI; ++x) {
#pragma unroll
        for (int y = 0; y < BJ; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                stg[(8 * (v >> 2) + 4 * hk + (v & 3)) * LS + 32 * y + rl] = acc[x][y][v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        constexpr int NCH = CW / 8;             // runs of 8 columns in a row
#pragma unroll
        for (int t0 = 0; t0 < 32 * NCH; t0 += 64) {
            const int t = t0 + lane, r = t / NCH, c = t % NCH;
            const int m = m0 + wm + 32 * x + r, n = n0 + wn + 8 * c;
            if (t < 32 * NCH && m < nrow && n < naux_pad) {
                const float4 u0 = *(const float4 *) (stg + r * LS + 8 * c);
                const float4 u1 = *(const float4 *) (stg + r * LS + 8 * c + 4);
                const float u[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
                v8bf h, l;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    h[e] = (__bf16) u[e];
                    l[e] = (__bf16) (u[e] - (float) h[e]);
                }
                const long k = (long) m * naux_pad + n;
                const size_t off = (size_t) (k >> 5) * 32 * nbf + (k & 31);
                *(v8bf *) (X0 + off) = h;
                *(v8bf *) (X0 + xpart + off) = l;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// FP32 partials of A B^T over k chunks, A and B panel-major with R rows (the products of the
// two-part X: the SYRK X X^T and the difference product X^M X^P^T).  Tiles of (64 BI) x
// (32 BJ WN); with UPPER only the tiles (rt, ct) holding an element m <= n (rt BM < (ct + 1) BN),
// else all nrowt x ncolt.  Chunk z covers panels [z cp, min(npan, (z + 1) cp)) and writes its
// tiles' elements (m, n < R; with UPPER those of the register tiles holding an element m <= n)
// to part[z] (R x R, FP32).  The chunks are dealt to the XCDs (the hardware
// deals workgroup b to XCD b % 8), every tile of chunk z on XCD z % 8 one after the other, so the
// tiles of a chunk run side by side and read its panels from the one L2 (nchunk is a multiple of 8).
template <int BI, int BJ, int WN, int NS, bool UPPER>
__global__ __launch_bounds__(128 * WN, 1) void k_pgemm_ps(const __bf16 *__restrict__ A, const __bf16 *__restrict__ B,
                                                     size_t xpart, int R, long npan, long cp, int ncolt, int ntiles,
                                                     float *__restrict__ part)
{
    using P = PsStage<BI, BJ, WN>;
    __shared__ __attribute__((aligned(16))) v8bf ring[NS][P::SLOTS];
    constexpr int BM = P::BM, BN = P::BN;
    const long s = blockIdx.x / 8;
    const long z = (s / ntiles) * 8 + blockIdx.x % 8;
    int tile = (int) (s % ntiles), rt = 0, ct = 0;
    if (UPPER) {
        // tile < ntiles = the count over rt < nrowt of the host's same sum: ends at such an rt
        for (;; ++rt) {
            const int c0 = rt * BM / BN, cnt = max(0, ncolt - c0);
            if (tile < cnt) {
                ct = c0 + tile;
                break;
            }
            tile -= cnt;
        }
    } else {
        rt = tile / ncolt;
        ct = tile % ncolt;
    }
    const long p0 = z * cp, p1 = std::min(npan, p0 + cp);
    const int nst = p1 > p0 ? (int) (p1 - p0) : 0;
    const size_t po = (size_t) (p0 < npan ? p0 : 0) * 32 * R;
    const int m0 = rt * BM, n0 = ct * BN;
    // UPPER: a wavefront whose register tile lies beyond the rows or columns or wholly below the
    // diagonal (first row past its last column) stores nothing, so it multiplies nothing, and one
    // whose register tile crosses the diagonal multiplies only the 32x32 blocks holding an element
    // m <= n (the reduction reads no other); the column blocks are mirrored in the tiles that have
    // such register tiles.  (Skipping the few register tiles of the full product beyond the rows
    // or columns did not pay.)
    bool mir = false, act = true;
    int cut = BI;
    if constexpr (UPPER) {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        auto live = [&](int a, int b) {
            const int ms = m0 + a * 32 * BI, ns = n0 + b * 32 * BJ;
            return ms < R && ns < R && ms < ns + 32 * BJ;
        };
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < WN; ++b) mir = mir || !live(a, b);
        const int a = w / WN, b = ps_colblk<WN>(w, mir);
        act = live(a, b);
        // block (x, y) holds an element m <= n iff its first row is before its last column:
        // x - y < (ns - ms) / 32 + 1 (ms, ns multiples of 32; >= 0 for a live register tile)
        cut = act ? (n0 + b * 32 * BJ - m0 - a * 32 * BI) / 32 + 1 : -1;
    }
    v16f acc[BI][BJ];
    ps_tile<BI, BJ, WN, NS>(&ring[0][0], A + po, xpart, R, m0, B + po, xpart, R, n0, nst, mir, cut, acc);
    if (!act) return;
    float *Pz = part + (size_t) z * R * R;
    ps_epilogue<BI, BJ, WN>(acc, mir, [&](int r, int c, float v) {
        const int m = m0 + r, n = n0 + c;
        if (m < R && n < R) Pz[(size_t) m * R + n] = v;
    });
}

constexpr int ST = 128;   // SYRK tile edge (2 x 2 wavefronts)

// upper 128x128 tiles of X X^T over k chunks; part[z] is FP32 per chunk
template <int NP>
__global__ __launch_bounds__(NTHR) void k_syrk_bf(const __bf16 *__restrict__ Xp, size_t xpart, int nbf, long kdim,
                                                  int ntb, int ntiles, long kchunk, long total,
                                                  float *__restrict__ part)
{
    const long t = xcd_remap(blockIdx.x, total);
    int tile = (int) (t % ntiles);
    const long z = t / ntiles;
    int bi = 0;
    while (tile >= ntb - bi) { tile -= ntb - bi; ++bi; }
    const int bj = bi + tile;
    const long kb = z * kchunk, ke = std::min<long>(kdim, kb + kchunk);
    constexpr int BK = syrk_bk<NP>();
    SLoad<NP, BK> la{Xp + kb, xpart, kdim, nbf, bi * ST};
    SLoad<NP, BK> lb{Xp + kb, xpart, kdim, nbf, bj * ST};   // X rows are kdim apart
    const int m0 = bi * ST, n0 = bj * ST;
    v16f acc[2][2];
    bfs_tile<NP, 2, 2, BK, 1, false>(la, lb, 0, (int) (ke - kb), acc, nbf - m0, nbf - n0);
    float *P = part + (size_t) z * nbf * nbf;
    bfs_epilogue<2>(acc, [&](int r, int c, float v) {
        const int m = m0 + r, n = n0 + c;
        if (m < nbf && n < nbf) P[(size_t) m * nbf + n] = v;
    });
}

// K[m][n] = K[n][m] = sum_z part[z][m][n] (in FP64) for m <= n inside upper tiles
__global__ void k_syrk_bf_reduce(const float *__restrict__ part, int nsplit, int nbf, double *__restrict__ K)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) nbf * nbf) return;
    const int m = (int) (idx / nbf), n = (int) (idx % nbf);
    if (m > n) return;
    double s0 = 0.0, s1 = 0.0;
    int z = 0;
    for (; z + 1 < nsplit; z += 2) {
        s0 += (double) part[(size_t) z * nbf * nbf + idx];
        s1 += (double) part[(size_t) (z + 1) * nbf * nbf + idx];
    }
    if (z < nsplit) s0 += (double) part[(size_t) z * nbf * nbf + idx];
    K[(size_t) m * nbf + n] = s0 + s1;
    K[(size_t) n * nbf + m] = s0 + s1;
}

// ---- difference-density exchange: G = A B^T over all 128x128 tiles of the n x n result, A and
// B rows of pre-split BF16 matrices (rows ld apart, parts xpart apart), k chunks of FP32
// partials in part[z]
template <int NP>
__global__ __launch_bounds__(NTHR) void k_gemm_bf_nt(const __bf16 *__restrict__ A, const __bf16 *__restrict__ B,
                                                     size_t xpart, long ld, int n, long kdim, int ntb, long kchunk,
                                                     long total, float *__restrict__ part)
{
    const long t = xcd_remap(blockIdx.x, total);
    const int ntiles = ntb * ntb;
    const int tile = (int) (t % ntiles);
    const long z = t / ntiles;
    const int bi = tile / ntb, bj = tile % ntb;
    const long kb = z * kchunk, ke = std::min<long>(kdim, kb + kchunk);
    constexpr int BK = syrk_bk<NP>();
    SLoad<NP, BK> la{A + kb, xpart, ld, n, bi * ST};
    SLoad<NP, BK> lb{B + kb, xpart, ld, n, bj * ST};
    const int m0 = bi * ST, n0 = bj * ST;
    v16f acc[2][2];
    bfs_tile<NP, 2, 2, BK, 1, false>(la, lb, 0, (int) (ke - kb), acc, n - m0, n - n0);
    float *P = part + (size_t) z * n * n;
    bfs_epilogue<2>(acc, [&](int r, int c, float v) {
        const int m = m0 + r, q = n0 + c;
        if (m < n && q < n) P[(size_t) m * n + q] = v;
    });
}

// G = sum_z part[z] in FP64, one thread per element (every read a whole row of a part)
__global__ void k_chunk_sum(const float *__restrict__ part, int nsplit, long nn, double *__restrict__ G)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nn) return;
    double s0 = 0.0, s1 = 0.0;
    int z = 0;
    for (; z + 1 < nsplit; z += 2) {
        s0 += (double) part[(size_t) z * nn + idx];
        s1 += (double) part[(size_t) (z + 1) * nn + idx];
    }
    if (z < nsplit) s0 += (double) part[(size_t) z * nn + idx];
    G[idx] = s0 + s1;
}

// K = Kref + (G + G^T) / 2, one thread per upper element
__global__ void k_incr_sym(const double *__restrict__ G, int n, const double *__restrict__ Kref,
                           double *__restrict__ K)
{
    const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long) n * n) return;
    const int m = (int) (idx / n), q = (int) (idx % n);
    if (m > q) return;
    const size_t a = (size_t) m * n + q, b = (size_t) q * n + m;
    const double v = Kref[a] + 0.5 * (G[a] + G[b]);
    K[a] = v;
    K[b] = v;
}

__global__ void k_to_f32(const double *__restrict__ in, float *__restrict__ out, long n)
{
    for (long i = (long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long) gridDim.x * blockDim.x)
        out[i] = (float) in[i];
}

}  // namespace

// ===========================================================================
int yg_to_f32(hipStream_t s, const double *in, float *out, long n)
{
    if (n <= 0) return YGPU_OK;
    const long nb = std::min<long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, s, in, out, n);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

template <int NP, int WM, int WN>
static int launch_xbuild_bf(hipStream_t s, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup,
                            const int *xpartner, const float *Cocc32, const float *B32, void *X)
{
    const int ncolt = (naux_pad + 64 * WN - 1) / (64 * WN), nrowt = (n_occ + 64 * WM - 1) / (64 * WM);
    const size_t xpart = (size_t) nbf * n_occ * naux_pad;
    const long total = (long) ncolt * nrowt * nbf;
    if (total > 0x7fffffffL) {
        ygpu_set_error(YGPU_ERR_INVALID, "X build: %ld workgroups exceed the grid limit", total);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL((k_xbuild_bf<NP, WM, WN>), dim3((unsigned) total), dim3(64 * WM * WN), 0, s, xptr, xup,
                       xpartner, Cocc32, B32, (__bf16 *) X, xpart, n_occ, naux_pad, ncolt, nrowt, total);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

template <int NP>
static int xbuild_np(hipStream_t s, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup,
                     const int *xpartner, const float *Cocc32, const float *B32, void *X)
{
    // all occupied rows in one workgroup up to 256 (row tiles of 256 beyond)
    if (n_occ <= 64) return launch_xbuild_bf<NP, 1, 4>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
    if (n_occ <= 128) return launch_xbuild_bf<NP, 2, 4>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
    if (n_occ <= 192) return launch_xbuild_bf<NP, 3, 4>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
    // (a 320 x 128 tile for 256 < n_occ <= 320 wastes no rows but measured slower than two 256-row tiles)
    return launch_xbuild_bf<NP, 4, 4>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
}

int yg_xbuild_bf(hipStream_t s, int nparts, int nbf, int n_occ, int naux_pad, const int *xptr, const int *xup,
                 const int *xpartner, const float *Cocc32, const float *B32, void *X)
{
    if (nbf <= 0 || n_occ <= 0) return YGPU_OK;
    if (nparts == 2) return xbuild_np<2>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
    if (nparts == 3) return xbuild_np<3>(s, nbf, n_occ, naux_pad, xptr, xup, xpartner, Cocc32, B32, X);
    ygpu_set_error(YGPU_ERR_INVALID, "X build: %d BF16 parts (2 or 3)", nparts);
    return YGPU_ERR_INVALID;
}

int yg_gather_split_t(hipStream_t s, int nktiles, const int2 *ktiles, const float *src, long ld, int R,
                      const int *idx, const int *xptr, const int *kofs, void *dst, size_t dpart)
{
    if (nktiles <= 0 || R <= 0) return YGPU_OK;
    // many operand rows (B): four 64-row passes per workgroup (measured 17.6 -> 11.2 ms at C96H24
    // against one pass); few (C_occ): one
    const int subs = R >= 1024 ? 4 : 1, nr = (R + 64 * subs - 1) / (64 * subs);
    if (nr > 65535) {
        ygpu_set_error(YGPU_ERR_INVALID, "operand blocks: %d rows exceed the grid limit", R);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_gather_split_t, dim3(nktiles, nr), dim3(256), 0, s, src, ld, R, idx, xptr, kofs, ktiles, subs,
                       (__bf16 *) dst, dpart);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

template <int BI, int BJ, int WN, int NS>
static int launch_xbuild_ps(hipStream_t s, int nbf, int nrow, int naux_pad, const int *xptr, const int *kofs,
                            const int *order, const void *Cp, size_t cpart, const void *Bt, size_t btpart, void *X)
{
    using P = PsStage<BI, BJ, WN>;
    const int ncolt = (naux_pad + P::BN - 1) / P::BN, nrowt = (nrow + P::BM - 1) / P::BM;
    const size_t xpart = (size_t) nbf * nrow * naux_pad;
    // 8 XCD lanes of ceil(nbf / 8) functions each (functions past nbf: empty workgroups)
    const long total = 8L * ((nbf + 7) / 8) * ncolt * nrowt;
    if (naux_pad % 32) {
        // the tile stores runs of 8 columns, inside one 32-k group of a row
        ygpu_set_error(YGPU_ERR_INVALID, "X build: padded aux dimension %d is not a multiple of 32", naux_pad);
        return YGPU_ERR_INVALID;
    }
    if (total > 0x7fffffffL) {
        ygpu_set_error(YGPU_ERR_INVALID, "X build: %ld workgroups exceed the grid limit", total);
        return YGPU_ERR_INVALID;
    }
    hipLaunchKernelGGL((k_xbuild_ps<BI, BJ, WN, NS>), dim3((unsigned) total), dim3(64 * P::NW), 0, s, xptr, kofs, order,
                       (const __bf16 *) Cp, cpart, (const __bf16 *) Bt, btpart, (__bf16 *) X, xpart, nrow, naux_pad,
                       ncolt, nrowt, nbf);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_xbuild_ps(hipStream_t s, int nbf, int nrow, int naux_pad, const int *xptr, const int *kofs, const int *order,
                 const void *Cp, size_t cpart, const void *Bt, size_t btpart, void *X)
{
    if (nbf <= 0 || nrow <= 0) return YGPU_OK;
#define YG_XPS(BI) launch_xbuild_ps<BI, 2, 4, 2>(s, nbf, nrow, naux_pad, xptr, kofs, order, Cp, cpart, Bt, btpart, X)
    // every row of X in one workgroup up to 320 (row tiles of 320 beyond, on one XCD), 256 aux
    // columns per tile on eight wavefronts; two ring stages of 72 KiB (a third does not fit)
    if (nrow <= 64) return YG_XPS(1);
    if (nrow <= 128) return YG_XPS(2);
    if (nrow <= 192) return YG_XPS(3);
    if (nrow <= 256) return YG_XPS(4);
    return YG_XPS(5);
#undef YG_XPS
}

// K = Kref + (G + G^T) / 2 with G = sum_z part[z]; G (FP64) is kept in part after the nsplit
// chunks, at the next multiple of 256 bytes (the caller sized part for it)
static int incr_reduce(hipStream_t s, DBuf &part, int nsplit, int n, const double *Kref, double *K)
{
    const long nn = (long) n * n;
    const size_t goff = ((size_t) nsplit * nn * sizeof(float) + 255) & ~(size_t) 255;
    double *G = (double *) (part.as<char>() + goff);
    hipLaunchKernelGGL(k_chunk_sum, dim3((nn + 255) / 256), dim3(256), 0, s, part.as<float>(), nsplit, nn, G);
    YG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_incr_sym, dim3((nn + 255) / 256), dim3(256), 0, s, G, n, Kref, K);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

// FP32 partials part[z] (z < *nchunk) of A B^T for the panel-major two-part X (R rows, kdim k,
// parts xpart apart); UPPER: only the tiles holding the upper triangle
template <bool UPPER>
static int pgemm_ps(hipStream_t s, int R, long kdim, const __bf16 *A, const __bf16 *B, size_t xpart, DBuf &part,
                    int *nchunk_out)
{
    constexpr int BI = 5, BJ = 2, WN = 4, NS = 2, BM = PsStage<BI, BJ, WN>::BM, BN = PsStage<BI, BJ, WN>::BN;
    const int nrowt = (R + BM - 1) / BM, ncolt = (R + BN - 1) / BN;
    int ntiles = 0;
    if (UPPER) {
        for (int rt = 0; rt < nrowt; ++rt) ntiles += std::max(0, ncolt - rt * BM / BN);
    } else {
        ntiles = nrowt * ncolt;
    }
    const long npan = kdim / 32;
    // FP32 accumulation over at most 256 panels (8192 k) per chunk, FP64 across chunks: longer
    // FP32 sums raised the error of K measurably (the DIIS error of the BF16 phase levelled off
    // ~25 % higher with 16-36 thousand k per chunk); the chunk count a multiple of 8
    // (fewer, longer chunks where their partials would pass 8 GiB)
    const long cap = std::max<long>(8, (long) ((8ull << 30) / ((size_t) R * R * sizeof(float))) / 8 * 8);
    const int nchunk = (int) std::min<long>(cap, std::max<long>(8, ((npan + 255) / 256 + 7) / 8 * 8));
    const int c = nchunk / 8;
    const long cp = (npan + nchunk - 1) / nchunk;
    // (and an R x R FP64 matrix after the partials for the reduction of the difference build)
    YG_TRY(part.ensure((size_t) nchunk * R * R * sizeof(float) + (size_t) R * R * sizeof(double) + 256));
    const long total = 8L * c * ntiles;
    hipLaunchKernelGGL((k_pgemm_ps<BI, BJ, WN, NS, UPPER>), dim3((unsigned) total), dim3(64 * 2 * WN), 0, s, A, B, xpart, R, npan,
                       cp, ncolt, ntiles, part.as<float>());
    YG_HIP(hipGetLastError());
    *nchunk_out = nchunk;
    return YGPU_OK;
}

template <int NP>
static int syrk_np(hipStream_t s, int nbf, long kdim, const void *X, double *K, DBuf &part)
{
    const int ntb = (nbf + ST - 1) / ST;
    const int ntiles = ntb * (ntb + 1) / 2;
    // chunks of 1024..8192 along k (FP32 accumulation inside a chunk, FP64 across
    // chunks), enough of them for about 1024 workgroups
    long kchunk = kdim * ntiles / 1024;
    kchunk = std::max<long>(1024, std::min<long>(8192, kchunk));
    constexpr int BK = syrk_bk<NP>();
    kchunk = (kchunk + BK - 1) / BK * BK;
    const int nsplit = (int) ((kdim + kchunk - 1) / kchunk);
    YG_TRY(part.ensure((size_t) nsplit * nbf * nbf * sizeof(float)));
    const long total = (long) ntiles * nsplit;
    const size_t xpart = (size_t) nbf * kdim;
    hipLaunchKernelGGL(k_syrk_bf<NP>, dim3((unsigned) total), dim3(NTHR), 0, s, (const __bf16 *) X, xpart, nbf, kdim,
                       ntb, ntiles, kchunk, total, part.as<float>());
    YG_HIP(hipGetLastError());
    const long tot = (long) nbf * nbf;
    hipLaunchKernelGGL(k_syrk_bf_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, part.as<float>(), nsplit, nbf, K);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int yg_syrk_bf(hipStream_t s, int nparts, int nbf, long kdim, const void *X, double *K, DBuf &part)
{
    if (nbf <= 0) return YGPU_OK;
    if (kdim % 8) {
        ygpu_set_error(YGPU_ERR_INVALID, "SYRK: k extent %ld is not a multiple of 8", kdim);
        return YGPU_ERR_INVALID;
    }
    if (nparts == 2) {
        // X panel-major (the two-part X build's layout)
        if (kdim % 32) {
            ygpu_set_error(YGPU_ERR_INVALID, "SYRK: k extent %ld is not a multiple of 32", kdim);
            return YGPU_ERR_INVALID;
        }
        const __bf16 *Xp = (const __bf16 *) X;
        int nsplit = 0;
        YG_TRY(pgemm_ps<true>(s, nbf, kdim, Xp, Xp, (size_t) nbf * kdim, part, &nsplit));
        const long tot = (long) nbf * nbf;
        hipLaunchKernelGGL(k_syrk_bf_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, part.as<float>(), nsplit, nbf, K);
        YG_HIP(hipGetLastError());
        return YGPU_OK;
    }
    if (nparts == 3) return syrk_np<3>(s, nbf, kdim, X, K, part);
    ygpu_set_error(YGPU_ERR_INVALID, "SYRK: %d BF16 parts (2 or 3)", nparts);
    return YGPU_ERR_INVALID;
}

int yg_incr_k_bf(hipStream_t s, int nparts, int nbf, int n_occ, int naux_pad, const void *X, const double *Kref,
                 double *K, DBuf &part)
{
    if (nbf <= 0) return YGPU_OK;
    const long kdim = (long) n_occ * naux_pad, ld = 2 * kdim;
    if (kdim % 8) {
        ygpu_set_error(YGPU_ERR_INVALID, "difference K: k extent %ld is not a multiple of 8", kdim);
        return YGPU_ERR_INVALID;
    }
    if (nparts == 2) {
        // X panel-major over the basis functions (the two-part X build's layout) for the combined
        // [P | M] block: k = m naux_pad + n with X^P_i in k < kdim, X^M_i in kdim <= k < 2 kdim, so
        // X^M starts kdim / 32 panels after X^P
        if (kdim % 32) {
            ygpu_set_error(YGPU_ERR_INVALID, "difference K: k extent %ld is not a multiple of 32", kdim);
            return YGPU_ERR_INVALID;
        }
        const __bf16 *XP = (const __bf16 *) X, *XM = XP + (size_t) kdim * nbf;
        int nsplit = 0;
        YG_TRY(pgemm_ps<false>(s, nbf, kdim, XM, XP, (size_t) nbf * 2 * kdim, part, &nsplit));
        return incr_reduce(s, part, nsplit, nbf, Kref, K);
    }
    const int ntb = (nbf + ST - 1) / ST;
    const int ntiles = ntb * ntb;
    long kchunk = kdim * ntiles / 1024;
    kchunk = std::max<long>(1024, std::min<long>(8192, kchunk));
    const int bk = syrk_bk<3>();
    kchunk = (kchunk + bk - 1) / bk * bk;
    const int nsplit = (int) ((kdim + kchunk - 1) / kchunk);
    YG_TRY(part.ensure((size_t) nsplit * nbf * nbf * sizeof(float) + (size_t) nbf * nbf * sizeof(double) + 256));
    const long total = (long) ntiles * nsplit;
    // X of the combined [P | M] block: rows 0 .. n_occ-1 of basis function i hold X^P_i, rows
    // n_occ .. 2 n_occ-1 X^M_i, so X^M_i starts kdim after X^P_i and rows are 2 kdim apart
    const size_t xpart = (size_t) nbf * ld;
    const __bf16 *XP = (const __bf16 *) X, *XM = XP + kdim;
    if (nparts == 3)
        hipLaunchKernelGGL(k_gemm_bf_nt<3>, dim3((unsigned) total), dim3(NTHR), 0, s, XM, XP, xpart, ld, nbf, kdim, ntb,
                           kchunk, total, part.as<float>());
    else {
        ygpu_set_error(YGPU_ERR_INVALID, "difference K: %d BF16 parts (2 or 3)", nparts);
        return YGPU_ERR_INVALID;
    }
    YG_HIP(hipGetLastError());
    return incr_reduce(s, part, nsplit, nbf, Kref, K);
}
