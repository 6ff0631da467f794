// ygpu_mlp.hip -- the learned-guess network: forward, backward and Adam on the GPU.
//
// Network (FP32): n_in -> nh -> nh -> nh -> n_out, SiLU after each hidden layer,
//   Z1 = X W1 + b1, H1 = silu(Z1), Z2 = H1 W2 + b2, H2 = silu(Z2),
//   Z3 = H2 W3 + b3, H3 = silu(Z3), Y = H3 W4 + b4,
// parameters packed as W1 [n_in][nh], b1, W2 [nh][nh], b2, W3, b3, W4 [nh][n_out], b4.
// Loss on a batch of B samples: L = sum M (Y - T)^2 / (B n_out) with a 0/1 mask M
// marking the coefficients a pair type has.
//
// The products are small (nh = 128) and batched over samples, so one LDS-tiled
// FP32 FMA GEMM (64x64 tiles, 4x4 per thread) with fused epilogues (bias + SiLU
// forward, SiLU' backward) serves all of them; the weight gradients, which
// reduce over the batch, split the batch into chunks with FP32 partials summed
// in a fixed order (deterministic training).  Nothing here is on the SCF hot
// path: inference is one forward pass over the atom pairs of one molecule.
#include "ygpu_internal.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace {

constexpr int TM = 64, TK = 16, NT = 256;

enum { OP_NN = 0, OP_TN = 1, OP_NT = 2 };
enum { EP_STORE = 0, EP_BIAS = 1, EP_BIAS_SILU = 2, EP_SILU_GRAD = 3 };

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + __expf(-z)); }

// C[M][N] (ldc) = op(A) op(B) over k in [k0, k1) with an epilogue; blockIdx.z picks a
// k chunk (EP_STORE into C + z*M*N) when split > 1.
//   OP_NN: A [M][K], B [K][N];  OP_TN: A [K][M];  OP_NT: B [N][K]
template <int OP, int EP>
__global__ __launch_bounds__(NT) void k_sgemm(int M, int N, int K, int kchunk, const float *__restrict__ A, int lda,
                                              const float *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                                              const float *__restrict__ bias, const float *__restrict__ Zin,
                                              float *__restrict__ Hout)
{
    __shared__ float As[TK][TM + 4];
    __shared__ float Bs[TK][TM + 4];
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TM;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    const int t = threadIdx.x, ty = t / 16, tx = t % 16;
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += TK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + NT * q;            // 0..1023 over the 16 x 64 tile
            if (OP == OP_TN) {
                const int kk = e / TM, mm = e % TM, k = k0 + kk, m = m0 + mm;
                As[kk][mm] = (k < ke && m < M) ? A[(size_t) k * lda + m] : 0.f;
            } else {
                const int mm = e / TK, kk = e % TK, k = k0 + kk, m = m0 + mm;
                As[kk][mm] = (k < ke && m < M) ? A[(size_t) m * lda + k] : 0.f;
            }
            if (OP == OP_NT) {
                const int nn = e / TK, kk = e % TK, k = k0 + kk, n = n0 + nn;
                Bs[kk][nn] = (k < ke && n < N) ? B[(size_t) n * ldb + k] : 0.f;
            } else {
                const int kk = e / TM, nn = e % TM, k = k0 + kk, n = n0 + nn;
                Bs[kk][nn] = (k < ke && n < N) ? B[(size_t) k * ldb + n] : 0.f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    float *Cz = C + (size_t) blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
            if (m >= M || n >= N) continue;
            const size_t o = (size_t) m * ldc + n;
            float v = acc[i][j];
            if (EP == EP_STORE) {
                Cz[(size_t) m * N + n] = v;
            } else if (EP == EP_BIAS) {
                C[o] = v + bias[n];
            } else if (EP == EP_BIAS_SILU) {
                v += bias[n];
                C[o] = v;
                Hout[o] = v * sigm(v);
            } else {
                const float z = Zin[o], s = sigm(z);
                C[o] = v * s * (1.f + z * (1.f - s));
            }
        }
}

// out[m][n] = sum_z part[z][m][n] (fixed order)
__global__ void k_sum_split(int MN, int nsplit, const float *__restrict__ part, float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MN) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; ++z) s += part[(size_t) z * MN + i];
    out[i] = s;
}

// db[n] = sum_m dZ[m][n]
__global__ void k_colsum(int M, int N, const float *__restrict__ dZ, float *__restrict__ db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += dZ[(size_t) m * N + n];
    db[n] = s;
}

// dY = 2 M (Y - T) / norm; part[block] = sum M (Y - T)^2 over the block
__global__ __launch_bounds__(256) void k_mse(int total, const float *__restrict__ Y, const float *__restrict__ T,
                                             const float *__restrict__ Mk, float inv_norm, float *__restrict__ dY,
                                             double *__restrict__ part)
{
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < total) {
        const float d = (Y[i] - T[i]) * Mk[i];
        dY[i] = 2.f * d * inv_norm;
        l = (double) d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_gather(int n, int width, const int *__restrict__ idx, const float *__restrict__ src,
                         float *__restrict__ dst)
{
    const long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long) n * width) return;
    const int r = (int) (i / width), c = (int) (i % width);
    dst[i] = src[(size_t) idx[r] * width + c];
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <int OP, int EP>
int sgemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C, int ldc,
          const float *bias = nullptr, const float *Zin = nullptr, float *Hout = nullptr, int nsplit = 1)
{
    const int kchunk = (K + nsplit - 1) / nsplit;
    dim3 grid((N + TM - 1) / TM, (M + TM - 1) / TM, nsplit);
    hipLaunchKernelGGL((k_sgemm<OP, EP>), grid, dim3(NT), 0, s, M, N, K, kchunk, A, lda, B, ldb, C, ldc, bias, Zin,
                       Hout);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

struct Net {
    int n_in, nh, n_out;
    size_t off[8];   // W1 b1 W2 b2 W3 b3 W4 b4
    size_t total;
    Net(int i, int h, int o) : n_in(i), nh(h), n_out(o)
    {
        const size_t sz[8] = {(size_t) i * h, (size_t) h, (size_t) h * h, (size_t) h, (size_t) h * h, (size_t) h,
                              (size_t) h * o, (size_t) o};
        size_t a = 0;
        for (int k = 0; k < 8; ++k) { off[k] = a; a += sz[k]; }
        total = a;
    }
};

// activations of one batch
struct Acts {
    DBuf Z[3], H[3], Y, dY, dH, dZ, wpart;
};

int forward(hipStream_t s, const Net &net, const float *P, int n, const float *X, Acts &a)
{
    const int nh = net.nh;
    for (int l = 0; l < 3; ++l) {
        YG_TRY(a.Z[l].ensure((size_t) n * nh * sizeof(float)));
        YG_TRY(a.H[l].ensure((size_t) n * nh * sizeof(float)));
    }
    YG_TRY(a.Y.ensure((size_t) n * net.n_out * sizeof(float)));
    YG_TRY((sgemm<OP_NN, EP_BIAS_SILU>(s, n, nh, net.n_in, X, net.n_in, P + net.off[0], nh, a.Z[0].as<float>(), nh,
                                        P + net.off[1], nullptr, a.H[0].as<float>())));
    for (int l = 1; l < 3; ++l)
        YG_TRY((sgemm<OP_NN, EP_BIAS_SILU>(s, n, nh, nh, a.H[l - 1].as<float>(), nh, P + net.off[2 * l], nh,
                                            a.Z[l].as<float>(), nh, P + net.off[2 * l + 1], nullptr,
                                            a.H[l].as<float>())));
    YG_TRY((sgemm<OP_NN, EP_BIAS>(s, n, net.n_out, nh, a.H[2].as<float>(), nh, P + net.off[6], net.n_out,
                                   a.Y.as<float>(), net.n_out, P + net.off[7])));
    return YGPU_OK;
}

// weight gradient dW [rows][cols] = In^T dZ over n samples, split over the batch
int weight_grad(hipStream_t s, int n, int rows, int cols, const float *In, const float *dZ, float *dW, DBuf &part)
{
    const int nsplit = std::max(1, std::min(64, n / 256));
    YG_TRY(part.ensure((size_t) nsplit * rows * cols * sizeof(float)));
    YG_TRY((sgemm<OP_TN, EP_STORE>(s, rows, cols, n, In, rows, dZ, cols, part.as<float>(), cols, nullptr, nullptr,
                                    nullptr, nsplit)));
    hipLaunchKernelGGL(k_sum_split, dim3((rows * cols + 255) / 256), dim3(256), 0, s, rows * cols, nsplit,
                       part.as<float>(), dW);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

// gradient of the loss (dY given) into G (packed like the parameters)
int backward(hipStream_t s, const Net &net, const float *P, int n, const float *X, Acts &a, float *G)
{
    const int nh = net.nh;
    YG_TRY(a.dH.ensure((size_t) n * nh * sizeof(float)));
    YG_TRY(a.dZ.ensure((size_t) n * nh * sizeof(float)));
    // output layer
    YG_TRY(weight_grad(s, n, nh, net.n_out, a.H[2].as<float>(), a.dY.as<float>(), G + net.off[6], a.wpart));
    hipLaunchKernelGGL(k_colsum, dim3((net.n_out + 255) / 256), dim3(256), 0, s, n, net.n_out, a.dY.as<float>(),
                       G + net.off[7]);
    YG_HIP(hipGetLastError());
    // dZ3 = (dY W4^T) silu'(Z3)
    YG_TRY((sgemm<OP_NT, EP_SILU_GRAD>(s, n, nh, net.n_out, a.dY.as<float>(), net.n_out, P + net.off[6], net.n_out,
                                        a.dZ.as<float>(), nh, nullptr, a.Z[2].as<float>())));
    for (int l = 2; l >= 0; --l) {
        const float *in = l > 0 ? a.H[l - 1].as<float>() : X;
        const int rows = l > 0 ? nh : net.n_in;
        YG_TRY(weight_grad(s, n, rows, nh, in, a.dZ.as<float>(), G + net.off[2 * l], a.wpart));
        hipLaunchKernelGGL(k_colsum, dim3((nh + 255) / 256), dim3(256), 0, s, n, nh, a.dZ.as<float>(),
                           G + net.off[2 * l + 1]);
        YG_HIP(hipGetLastError());
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l^T) silu'(Z_{l-1}); dH is the staging copy of dZ_l
            YG_HIP(hipMemcpyAsync(a.dH.p, a.dZ.p, (size_t) n * nh * sizeof(float), hipMemcpyDeviceToDevice, s));
            YG_TRY((sgemm<OP_NT, EP_SILU_GRAD>(s, n, nh, nh, a.dH.as<float>(), nh, P + net.off[2 * l], nh,
                                                a.dZ.as<float>(), nh, nullptr, a.Z[l - 1].as<float>())));
        }
    }
    return YGPU_OK;
}

int mse(hipStream_t s, int n, int n_out, const float *Y, const float *T, const float *Mk, float *dY, DBuf &part,
        double *loss_sum)
{
    const int total = n * n_out, nb = (total + 255) / 256;
    YG_TRY(part.ensure((size_t) nb * sizeof(double)));
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, total, Y, T, Mk, 1.f / (float) total, dY,
                       part.as<double>());
    YG_HIP(hipGetLastError());
    std::vector<double> h(nb);
    YG_HIP(hipMemcpyAsync(h.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    YG_HIP(hipStreamSynchronize(s));
    double l = 0.0;
    for (double v : h) l += v;
    *loss_sum = l;
    return YGPU_OK;
}

int gather(hipStream_t s, int n, int width, const int *idx, const float *src, float *dst)
{
    const long tot = (long) n * width;
    hipLaunchKernelGGL(k_gather, dim3((unsigned) ((tot + 255) / 256)), dim3(256), 0, s, n, width, idx, src, dst);
    YG_HIP(hipGetLastError());
    return YGPU_OK;
}

int check_dims(int n_in, int nh, int n_out)
{
    if (n_in <= 0 || nh <= 0 || n_out <= 0 || n_in > 4096 || nh > 4096 || n_out > 4096) {
        ygpu_set_error(YGPU_ERR_INVALID, "MLP dimensions %d/%d/%d out of range", n_in, nh, n_out);
        return YGPU_ERR_INVALID;
    }
    return YGPU_OK;
}

int exception_status_mlp()
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        ygpu_set_error(YGPU_ERR_NOMEM, "host allocation failed");
        return YGPU_ERR_NOMEM;
    } catch (...) {
        ygpu_set_error(YGPU_ERR_INVALID, "unexpected exception");
        return YGPU_ERR_INVALID;
    }
}

}  // namespace

struct ygpu_mlp_trainer {
    ygpu_ctx *ctx;
    Net net{1, 1, 1};
    int n_samples = 0;
    long step = 0;
    DBuf X, Y, M, P, G, Am, Av, Xb, Yb, Mb, idx, lpart;
    Acts a;
};

// ===========================================================================
int ygpu_mlp_forward(ygpu_ctx_t *c, int n_in, int nh, int n_out, const float *params, int n, const float *X,
                     float *Y)
try {
    if (!c || !params || !X || !Y || n < 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_forward: bad arguments");
        return YGPU_ERR_INVALID;
    }
    YG_TRY(check_dims(n_in, nh, n_out));
    if (n == 0) return YGPU_OK;
    YG_HIP(hipSetDevice(c->device));
    Net net(n_in, nh, n_out);
    DBuf dP, dX;
    Acts a;
    YG_TRY(dP.ensure(net.total * sizeof(float)));
    YG_TRY(dX.ensure((size_t) n * n_in * sizeof(float)));
    YG_HIP(hipMemcpyAsync(dP.p, params, net.total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(dX.p, X, (size_t) n * n_in * sizeof(float), hipMemcpyHostToDevice, c->stream));
    YG_TRY(forward(c->stream, net, dP.as<float>(), n, dX.as<float>(), a));
    YG_HIP(hipMemcpyAsync(Y, a.Y.p, (size_t) n * n_out * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    for (int l = 0; l < 3; ++l) { a.Z[l].release(); a.H[l].release(); }
    a.Y.release();
    dP.release();
    dX.release();
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}

int ygpu_mlp_trainer_create(ygpu_ctx_t *c, int n_in, int nh, int n_out, int n_samples, const float *X, const float *Y,
                            const float *M, const float *params_init, ygpu_mlp_trainer_t **out)
try {
    if (!out) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_create: null output");
        return YGPU_ERR_INVALID;
    }
    *out = nullptr;
    if (!c || !X || !Y || !M || !params_init || n_samples <= 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_create: bad arguments");
        return YGPU_ERR_INVALID;
    }
    YG_TRY(check_dims(n_in, nh, n_out));
    YG_HIP(hipSetDevice(c->device));
    ygpu_mlp_trainer *t = new ygpu_mlp_trainer;
    t->ctx = c;
    t->net = Net(n_in, nh, n_out);
    t->n_samples = n_samples;
    const size_t np = t->net.total;
    int st = YGPU_OK;
    auto up = [&](DBuf &b, const void *h, size_t bytes) {
        if (st == YGPU_OK) st = b.ensure(bytes);
        if (st == YGPU_OK && hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            ygpu_set_error(YGPU_ERR_HIP, "ygpu_mlp_trainer_create: upload failed");
            st = YGPU_ERR_HIP;
        }
    };
    up(t->X, X, (size_t) n_samples * n_in * sizeof(float));
    up(t->Y, Y, (size_t) n_samples * n_out * sizeof(float));
    up(t->M, M, (size_t) n_samples * n_out * sizeof(float));
    up(t->P, params_init, np * sizeof(float));
    if (st == YGPU_OK) st = t->G.ensure(np * sizeof(float));
    if (st == YGPU_OK) st = t->Am.ensure(np * sizeof(float));
    if (st == YGPU_OK) st = t->Av.ensure(np * sizeof(float));
    if (st == YGPU_OK && (hipMemsetAsync(t->Am.p, 0, np * sizeof(float), c->stream) != hipSuccess ||
                          hipMemsetAsync(t->Av.p, 0, np * sizeof(float), c->stream) != hipSuccess ||
                          hipStreamSynchronize(c->stream) != hipSuccess)) {
        ygpu_set_error(YGPU_ERR_HIP, "ygpu_mlp_trainer_create: initialisation failed");
        st = YGPU_ERR_HIP;
    }
    if (st != YGPU_OK) {
        delete t;
        return st;
    }
    *out = t;
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}

// loss and gradient on samples idx[0..n) (host indices); optional copy of the gradient
static int trainer_grad(ygpu_mlp_trainer *t, int n, const int *idx_dev, double *loss_mean)
{
    hipStream_t s = t->ctx->stream;
    const Net &net = t->net;
    YG_TRY(t->Xb.ensure((size_t) n * net.n_in * sizeof(float)));
    YG_TRY(t->Yb.ensure((size_t) n * net.n_out * sizeof(float)));
    YG_TRY(t->Mb.ensure((size_t) n * net.n_out * sizeof(float)));
    YG_TRY(gather(s, n, net.n_in, idx_dev, t->X.as<float>(), t->Xb.as<float>()));
    YG_TRY(gather(s, n, net.n_out, idx_dev, t->Y.as<float>(), t->Yb.as<float>()));
    YG_TRY(gather(s, n, net.n_out, idx_dev, t->M.as<float>(), t->Mb.as<float>()));
    YG_TRY(forward(s, net, t->P.as<float>(), n, t->Xb.as<float>(), t->a));
    YG_TRY(t->a.dY.ensure((size_t) n * net.n_out * sizeof(float)));
    double lsum = 0.0;
    YG_TRY(mse(s, n, net.n_out, t->a.Y.as<float>(), t->Yb.as<float>(), t->Mb.as<float>(), t->a.dY.as<float>(),
               t->lpart, &lsum));
    YG_TRY(backward(s, net, t->P.as<float>(), n, t->Xb.as<float>(), t->a, t->G.as<float>()));
    *loss_mean = lsum / ((double) n * net.n_out);
    return YGPU_OK;
}

int ygpu_mlp_trainer_grad(ygpu_mlp_trainer_t *t, int n, const int *idx, float *grad, double *loss)
try {
    if (!t || n <= 0 || !idx || !grad || !loss) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_grad: bad arguments");
        return YGPU_ERR_INVALID;
    }
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= t->n_samples) {
            ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_grad: sample index %d out of range", idx[i]);
            return YGPU_ERR_INVALID;
        }
    hipStream_t s = t->ctx->stream;
    YG_TRY(t->idx.ensure((size_t) n * sizeof(int)));
    YG_HIP(hipMemcpyAsync(t->idx.p, idx, (size_t) n * sizeof(int), hipMemcpyHostToDevice, s));
    YG_TRY(trainer_grad(t, n, t->idx.as<int>(), loss));
    YG_HIP(hipMemcpyAsync(grad, t->G.p, t->net.total * sizeof(float), hipMemcpyDeviceToHost, s));
    YG_HIP(hipStreamSynchronize(s));
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}

int ygpu_mlp_trainer_epoch(ygpu_mlp_trainer_t *t, int batch, float lr, unsigned seed, double *mean_loss)
try {
    if (!t || batch <= 0 || !(lr > 0.f) || !mean_loss) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_epoch: bad arguments");
        return YGPU_ERR_INVALID;
    }
    hipStream_t s = t->ctx->stream;
    std::vector<int> perm(t->n_samples);
    for (int i = 0; i < t->n_samples; ++i) perm[i] = i;
    std::mt19937_64 rng(seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    YG_TRY(t->idx.ensure(perm.size() * sizeof(int)));
    YG_HIP(hipMemcpyAsync(t->idx.p, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, s));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double lw = 0.0;
    long cnt = 0;
    for (int b0 = 0; b0 < t->n_samples; b0 += batch) {
        const int n = std::min(batch, t->n_samples - b0);
        double l = 0.0;
        YG_TRY(trainer_grad(t, n, t->idx.as<int>() + b0, &l));
        ++t->step;
        const float c1 = 1.f - std::pow(b1, (float) t->step), c2 = 1.f - std::pow(b2, (float) t->step);
        const int np = (int) t->net.total;
        hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), dim3(256), 0, s, np, t->P.as<float>(), t->G.as<float>(),
                           t->Am.as<float>(), t->Av.as<float>(), lr, b1, b2, eps, c1, c2);
        YG_HIP(hipGetLastError());
        lw += l * n;
        cnt += n;
    }
    YG_HIP(hipStreamSynchronize(s));
    *mean_loss = lw / (double) cnt;
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}

int ygpu_mlp_trainer_params(ygpu_mlp_trainer_t *t, float *params)
try {
    if (!t || !params) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_mlp_trainer_params: bad arguments");
        return YGPU_ERR_INVALID;
    }
    YG_HIP(hipMemcpyAsync(params, t->P.p, t->net.total * sizeof(float), hipMemcpyDeviceToHost, t->ctx->stream));
    YG_HIP(hipStreamSynchronize(t->ctx->stream));
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}

int ygpu_mlp_trainer_destroy(ygpu_mlp_trainer_t *t)
try {
    if (!t) return YGPU_OK;
    (void) hipStreamSynchronize(t->ctx->stream);
    delete t;
    return YGPU_OK;
} catch (...) {
    return exception_status_mlp();
}
