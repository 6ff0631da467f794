// ygpu_api.cpp -- the C-ABI of libyatscf_gpu.so (declared in include/yatscf_gpu.h).
//
// Owns the device state of one DF-RHF engine: shells, Boys table, the packed
// pair layout, the DF tensor B (rows i <= j only, aux index padded to a
// multiple of 64), J^{-1/2}, and the per-iteration work buffers.  FP64
// eigensolves reduce to tridiagonal form on the device (ygpu_eigen.hip up to
// n = 272, ygpu_sytrd.hip up to 2048, rocSOLVER dsyevd beyond), solve T with the
// device multisection / inverse iteration (density, n <= 272) or rocSOLVER dstedc,
// and apply Q with the register back-transform; every product is one of the MFMA
// kernels of ygpu_gemm.hip.
#include "ygpu_internal.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iterator>
#include <map>
#include <mutex>
#include <new>
#include <vector>

static thread_local std::string g_err;

void ygpu_set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = std::string("[ygpu ") + std::to_string(code) + "] " + buf;
}

namespace {

constexpr size_t CACHE_MIN = 1u << 20;   // blocks below this go straight back to the driver

struct BlockCache {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void *>> free;   // device -> (size, block)
    std::map<int, size_t> held, cap;                      // cached bytes; at most half the device's memory
};

BlockCache &block_cache()
{
    static BlockCache *g = new BlockCache;   // never destroyed: no device call at process exit
    return *g;
}

// a cached block of device dev with n <= size <= 2 n (the smallest such), or nullptr
void *cache_take(int dev, size_t n, size_t *got)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.mu);
    auto &m = bc.free[dev];
    auto it = m.lower_bound(n);
    if (it == m.end() || it->first / 2 > n) return nullptr;
    void *p = it->second;
    *got = it->first;
    bc.held[dev] -= it->first;
    m.erase(it);
    return p;
}

// called with device dev current
void cache_put(int dev, void *p, size_t n)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.mu);
    size_t &cap = bc.cap[dev];
    if (!cap) {
        size_t fr = 0, tot = 0;
        cap = hipMemGetInfo(&fr, &tot) == hipSuccess && tot ? tot / 2 : ((size_t) 64 << 30);
    }
    auto &m = bc.free[dev];
    size_t &held = bc.held[dev];
    // the largest cached blocks go back to the driver first
    while (!m.empty() && held + n > cap) {
        auto last = std::prev(m.end());
        (void) hipFree(last->second);
        held -= last->first;
        m.erase(last);
    }
    if (held + n > cap) {
        (void) hipFree(p);
        return;
    }
    m.emplace(n, p);
    held += n;
}

// hipFree every cached block of device dev (all devices for dev < 0)
void cache_drain(int dev)
{
    BlockCache &bc = block_cache();
    std::lock_guard<std::mutex> lk(bc.mu);
    int cur = 0;
    (void) hipGetDevice(&cur);
    for (auto &kv : bc.free) {
        if (dev >= 0 && kv.first != dev) continue;
        (void) hipSetDevice(kv.first);
        for (auto &b : kv.second) (void) hipFree(b.second);
        kv.second.clear();
        bc.held[kv.first] = 0;
    }
    (void) hipSetDevice(cur);
}

}  // namespace

int DBuf::ensure(size_t n)
{
    if (n <= bytes && p) return YGPU_OK;
    release();
    if (n == 0) n = 8;
    int d = 0;
    YG_HIP(hipGetDevice(&d));
    if (n >= CACHE_MIN) {
        size_t got = 0;
        if (void *q = cache_take(d, n, &got)) {
            p = q;
            bytes = got;
            dev = d;
            return YGPU_OK;
        }
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        (void) hipGetLastError();
        cache_drain(d);
        e = hipMalloc(&p, n);
    }
    if (e != hipSuccess) {
        p = nullptr;
        ygpu_set_error(YGPU_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
        return YGPU_ERR_NOMEM;
    }
    bytes = n;
    dev = d;
    return YGPU_OK;
}

void DBuf::release()
{
    if (p) {
        int cur = 0;
        (void) hipGetDevice(&cur);
        if (cur != dev) (void) hipSetDevice(dev);
        if (bytes >= CACHE_MIN) cache_put(dev, p, bytes);
        else (void) hipFree(p);
        if (cur != dev) (void) hipSetDevice(cur);
    }
    p = nullptr;
    bytes = 0;
}

int ygpu_release_cached_memory(void)
{
    cache_drain(-1);
    return YGPU_OK;
}

namespace {

constexpr int ROUND_AUX = 64;

// F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)), all terms positive
long double boys_series(int n, long double T)
{
    long double term = 1.0L / (2 * n + 1), sum = term;
    for (int k = 1; k < 5000; ++k) {
        term *= 2.0L * T / (long double) (2 * n + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return expl(-T) * sum;
}

int upload(DBuf &b, const void *src, size_t bytes, hipStream_t s)
{
    YG_TRY(b.ensure(bytes));
    if (bytes) YG_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
    return YGPU_OK;
}

int check_ctx(ygpu_ctx *c)
{
    if (!c) {
        ygpu_set_error(YGPU_ERR_INVALID, "null context");
        return YGPU_ERR_INVALID;
    }
    YG_HIP(hipSetDevice(c->device));
    return YGPU_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void) hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Size the eigensolver workspace with a rocBLAS size query and hand the
// handle a library-owned buffer: rocBLAS then never allocates or frees device
// memory itself (its handle destructor has nothing to release).
int ensure_syevd_workspace(ygpu_ctx *c, int n, int ld)
{
    size_t need = 0;
    if (rocblas_start_device_memory_size_query(c->rb) != rocblas_status_success) {
        ygpu_set_error(YGPU_ERR_SOLVER, "rocblas_start_device_memory_size_query failed");
        return YGPU_ERR_SOLVER;
    }
    (void) rocsolver_dsyevd(c->rb, rocblas_evect_original, rocblas_fill_upper, n, nullptr, ld, nullptr, nullptr,
                            nullptr);
    (void) rocsolver_dstedc(c->rb, rocblas_evect_tridiagonal, n, nullptr, nullptr, nullptr, ld, nullptr);
    if (rocblas_stop_device_memory_size_query(c->rb, &need) != rocblas_status_success) {
        ygpu_set_error(YGPU_ERR_SOLVER, "rocblas_stop_device_memory_size_query failed");
        return YGPU_ERR_SOLVER;
    }
    if (need > c->solver_ws_bytes) {
        YG_HIP(hipStreamSynchronize(c->stream));
        YG_TRY(c->d_solver_ws.ensure(need));
        if (rocblas_set_workspace(c->rb, c->d_solver_ws.p, need) != rocblas_status_success) {
            ygpu_set_error(YGPU_ERR_SOLVER, "rocblas_set_workspace(%zu bytes) failed", need);
            return YGPU_ERR_SOLVER;
        }
        c->solver_ws_bytes = need;
    }
    return YGPU_OK;
}

// After yg_sytrd_blocked on A (reflectors in A, d in d_eig, e in d_evE, tau in d_tau): all
// eigenpairs of T by rocSOLVER's divide and conquer into Z (column-major, leading dimension
// ld), then the first ncols eigenvectors of A = Q Z[:, 0:ncols] in place (the reflectors
// applied one wavefront per vector, as LAPACK's dormtr would).
int solver_tridiag_to_full(ygpu_ctx *c, int n, int ld, double *A, double *Z, int ncols)
{
    rocblas_status st = rocsolver_dstedc(c->rb, rocblas_evect_tridiagonal, n, c->d_eig.as<double>(),
                                         c->d_evE.as<double>(), Z, ld, c->d_info.as<int>());
    if (st != rocblas_status_success) {
        ygpu_set_error(YGPU_ERR_SOLVER, "rocsolver_dstedc(n=%d) returned %d", n, (int) st);
        return YGPU_ERR_SOLVER;
    }
    return yg_backtransform_cols(c->stream, n, ncols, A, ld, c->d_tau.as<double>(), Z, ld);
}

// In-place A (n x n, leading dimension ld) -> out = A^{-1/2}; A is destroyed.
int inv_sqrt_device(ygpu_ctx *c, int n, int ld, double *dA, double *dW, double *dOut, double *min_eig)
{
    YG_TRY(c->d_eig.ensure((size_t) n * sizeof(double)));
    YG_TRY(c->d_evE.ensure((size_t) n * sizeof(double)));
    YG_TRY(c->d_info.ensure(sizeof(int)));
    YG_TRY(ensure_syevd_workspace(c, n, ld));
    const bool small = n <= YG_EIG_NMAX_SMALL && ld == n;
    if (small) {
        // A = Q T Q^T on one workgroup, T = Z diag Z^T by divide and conquer, V = Q Z: the rows
        // of dOut receive the eigenvectors, then the rows of dA as the code below expects
        YG_TRY(c->d_tau.ensure((size_t) n * sizeof(double)));
        YG_TRY(yg_sytrd_small(c->stream, n, dA, c->d_eig.as<double>(), c->d_evE.as<double>(), c->d_tau.as<double>()));
        rocblas_status st = rocsolver_dstedc(c->rb, rocblas_evect_tridiagonal, n, c->d_eig.as<double>(),
                                             c->d_evE.as<double>(), dW, n, c->d_info.as<int>());
        if (st != rocblas_status_success) {
            ygpu_set_error(YGPU_ERR_SOLVER, "rocsolver_dstedc(n=%d) returned %d", n, (int) st);
            return YGPU_ERR_SOLVER;
        }
        YG_TRY(yg_backtransform_cols(c->stream, n, n, dA, n, c->d_tau.as<double>(), dW, n));
        YG_HIP(hipMemcpyAsync(dA, dW, (size_t) n * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else if (n <= YG_SYTRD_BLOCKED_NMAX) {
        // blocked reduction (reflectors left in dA), T = Z diag Z^T by divide and conquer into
        // dW, V = Q Z in place, then V back into dA (column-major view)
        YG_TRY(c->d_tau.ensure((size_t) n * sizeof(double)));
        YG_TRY(c->d_tri.ensure(yg_sytrd_blocked_ws(n)));
        YG_TRY(yg_sytrd_blocked(c->stream, n, ld, dA, c->d_eig.as<double>(), c->d_evE.as<double>(),
                                c->d_tau.as<double>(), c->d_tri.as<double>()));
        YG_TRY(solver_tridiag_to_full(c, n, ld, dA, dW, n));
        YG_HIP(hipMemcpyAsync(dA, dW, (size_t) ld * n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else {
        rocblas_status st = rocsolver_dsyevd(c->rb, rocblas_evect_original, rocblas_fill_upper, n, dA, ld,
                                             c->d_eig.as<double>(), c->d_evE.as<double>(), c->d_info.as<int>());
        if (st != rocblas_status_success) {
            ygpu_set_error(YGPU_ERR_SOLVER, "rocsolver_dsyevd(n=%d) returned %d", n, (int) st);
            return YGPU_ERR_SOLVER;
        }
    }
    int info = 0;
    std::vector<double> ev(n);
    YG_HIP(hipMemcpyAsync(&info, c->d_info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipMemcpyAsync(ev.data(), c->d_eig.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    if (info != 0) {
        ygpu_set_error(YGPU_ERR_SOLVER, "symmetric eigensolver (n=%d) info=%d", n, info);
        return YGPU_ERR_SOLVER;
    }
    if (min_eig) *min_eig = ev[0];
    if (!(ev[0] > 0.0)) {
        ygpu_set_error(YGPU_ERR_NUMERIC, "matrix not positive definite: smallest eigenvalue %.3e", ev[0]);
        return YGPU_ERR_NUMERIC;
    }
    // dA holds V column-major: dA[i + j*ld] = V(i,j), i.e. row k of the row-major view is
    // eigenvector k.  W = diag(lambda^{-1/2}) V^T (rows scaled); out = V W = (V^T)^T W.
    YG_HIP(hipMemcpy2DAsync(dW, (size_t) ld * sizeof(double), dA, (size_t) ld * sizeof(double), n * sizeof(double),
                            n, hipMemcpyDeviceToDevice, c->stream));
    YG_TRY(yg_scale_rows(c->stream, n, n, dW, ld, c->d_eig.as<double>(), 1));
    YG_TRY(yg_dgemm(c->stream, 1, 0, n, n, n, 1.0, dA, ld, dW, ld, 0.0, dOut, ld, &c->d_gemm_ws));
    // the two triangles round the scaled products differently: make the result exactly symmetric
    YG_TRY(yg_symmetrize(c->stream, n, dOut, ld));
    return YGPU_OK;
}

// Converts an exception escaping a C-ABI entry point into an error status
// (nothing may unwind through extern "C" frames into the caller).
int exception_status()
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        ygpu_set_error(YGPU_ERR_NOMEM, "host allocation failed");
        return YGPU_ERR_NOMEM;
    } catch (const std::exception &e) {
        ygpu_set_error(YGPU_ERR_INVALID, "exception: %s", e.what());
        return YGPU_ERR_INVALID;
    } catch (...) {
        ygpu_set_error(YGPU_ERR_INVALID, "unknown exception");
        return YGPU_ERR_INVALID;
    }
}

}  // namespace

extern "C" {

const char *ygpu_last_error(void) { return g_err.c_str(); }

static thread_local ygpu_failure_handler_t g_fail_handler = nullptr;

int ygpu_set_failure_handler(ygpu_failure_handler_t handler)
{
    g_fail_handler = handler;
    return YGPU_OK;
}

void ygpu_fail(int status, const char *where)
{
    const char *w = where ? where : "ygpu";
    if (g_fail_handler) g_fail_handler(status, w, g_err.c_str());   // may not return
    std::fprintf(stderr, "%s: %s\n", w, g_err.c_str());
    std::fflush(stderr);
    std::abort();
}

int ygpu_device_count(int *count)
try {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_create(int device, ygpu_ctx_t **out)
try {
    if (!out) {
        ygpu_set_error(YGPU_ERR_INVALID, "null output pointer");
        return YGPU_ERR_INVALID;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        ygpu_set_error(YGPU_ERR_NODEVICE, "no HIP device visible");
        return YGPU_ERR_NODEVICE;
    }
    if (device < 0 || device >= n) {
        ygpu_set_error(YGPU_ERR_INVALID, "device %d out of range (%d visible)", device, n);
        return YGPU_ERR_INVALID;
    }
    YG_HIP(hipSetDevice(device));
    YG_TRY(yg_scratch_init(device));
    ygpu_ctx *c = new ygpu_ctx();
    c->device = device;
    YG_HIP(hipStreamCreate(&c->stream));
    YG_HIP(hipStreamCreate(&c->side));
    // YGPU_J_BESIDE=1: the J build of a Fock build on the side stream beside the K build
    // (default: before it on the main stream, each kernel with the whole GPU)
    if (const char *e = std::getenv("YGPU_J_BESIDE")) c->j_beside = std::atoi(e) != 0;
    YG_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    YG_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
    for (auto &e : c->ev) YG_HIP(hipEventCreate(&e));
    if (rocblas_create_handle(&c->rb) != rocblas_status_success ||
        rocblas_set_stream(c->rb, c->stream) != rocblas_status_success) {
        ygpu_set_error(YGPU_ERR_SOLVER, "rocblas handle creation failed");
        return YGPU_ERR_SOLVER;
    }
    std::vector<double> tab((size_t) YG_BOYS_NPTS * (YG_BOYS_NMAX + 1));
    for (int k = 0; k < YG_BOYS_NPTS; ++k)
        for (int m = 0; m <= YG_BOYS_NMAX; ++m)
            tab[(size_t) k * (YG_BOYS_NMAX + 1) + m] = (double) boys_series(m, (long double) k * (long double) YG_BOYS_H);
    YG_TRY(upload(c->boys, tab.data(), tab.size() * sizeof(double), c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    *out = c;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_host_alloc(size_t bytes, void **ptr)
try {
    if (!ptr) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_host_alloc: null result pointer");
        return YGPU_ERR_INVALID;
    }
    *ptr = nullptr;
    YG_HIP(hipHostMalloc(ptr, bytes ? bytes : 64, hipHostMallocDefault));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_host_free(void *ptr)
try {
    if (ptr) YG_HIP(hipHostFree(ptr));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_destroy(ygpu_ctx_t *c)
try {
    if (!c) return YGPU_OK;
    (void) hipSetDevice(c->device);
    int status = YGPU_OK;
    hipError_t pend = hipStreamSynchronize(c->stream);
    if (pend == hipSuccess && c->side) pend = hipStreamSynchronize(c->side);
    if (pend != hipSuccess) {
        ygpu_set_error(YGPU_ERR_HIP, "ygpu_destroy: pending device error %s", hipGetErrorString(pend));
        status = YGPU_ERR_HIP;
    }
    // the handle first: it only references the library-owned workspace
    if (c->rb) rocblas_destroy_handle(c->rb);
    c->rb = nullptr;
    c->d_solver_ws.release();
    DBuf *bufs[] = {&c->boys, &c->d_charge, &c->d_atom_xyz, &c->d_upmask, &c->d_up_i, &c->d_up_j, &c->d_xptr,
                    &c->d_xup, &c->d_xpartner, &c->d_sp_scr, &c->d_df_scr, &c->d_pqA, &c->d_B, &c->d_Jpq,
                    &c->d_D, &c->d_Cocc, &c->d_H, &c->d_J, &c->d_K, &c->d_F, &c->d_tempJ, &c->d_tempJpart,
                    &c->d_X, &c->d_Kpart, &c->d_mat0, &c->d_mat1, &c->d_mat2, &c->d_eig, &c->d_info, &c->d_evE,
                    &c->d_gA, &c->d_gB, &c->d_gC, &c->d_S, &c->d_Xo, &c->d_Fin, &c->d_Din, &c->d_T1,
                    &c->d_T2, &c->d_diisR, &c->d_diisF, &c->d_dots};
    for (DBuf *b : bufs) b->release();
    for (BasisDev *b : {&c->bas, &c->aux}) {
        DBuf *bb[] = {&b->d_am, &b->d_nprim, &b->d_prim_off, &b->d_bf_start, &b->d_xyz, &b->d_alpha, &b->d_coef};
        for (DBuf *x : bb) x->release();
    }
    for (auto &e : c->ev) if (e) (void) hipEventDestroy(e);
    if (c->fork) (void) hipEventDestroy(c->fork);
    if (c->join) (void) hipEventDestroy(c->join);
    if (c->side) (void) hipStreamDestroy(c->side);
    if (c->stream) (void) hipStreamDestroy(c->stream);
    delete c;
    return status;
} catch (...) {
    return exception_status();
}

// a new basis or pair layout: nothing left on the device belongs to the next call's system
static void invalidate_handoffs(ygpu_ctx *c)
{
    c->fock_valid = c->density_valid = c->fcur_valid = 0;
    c->kref_valid = 0;
    c->jinc_valid = 0;
    c->have_vt = 0;
    c->sp2_last_steps = 0;
    for (int &ok : c->diis_err_ok) ok = 0;
}

int ygpu_set_basis(ygpu_ctx_t *c, int which, int nshells, const int *am, const int *nprim, const int *prim_off,
                   const double *xyz, const double *alpha, const double *coef)
try {
    YG_TRY(check_ctx(c));
    if ((which != 0 && which != 1) || nshells <= 0 || !am || !nprim || !prim_off || !xyz || !alpha || !coef) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_basis: bad arguments");
        return YGPU_ERR_INVALID;
    }
    BasisDev &b = which == 0 ? c->bas : c->aux;
    b.nshells = nshells;
    b.am.assign(am, am + nshells);
    b.nprim.assign(nprim, nprim + nshells);
    b.prim_off.assign(prim_off, prim_off + nshells);
    b.xyz.assign(xyz, xyz + 3 * (size_t) nshells);
    b.bf_start.resize(nshells + 1);
    int nbf = 0, nprim_end = 0;
    b.max_am = 0;
    b.max_nprim = 0;
    for (int i = 0; i < nshells; ++i) {
        if (am[i] < 0 || am[i] > 4 || nprim[i] <= 0 || prim_off[i] < 0) {
            ygpu_set_error(YGPU_ERR_INVALID, "shell %d: am %d nprim %d", i, am[i], nprim[i]);
            return YGPU_ERR_INVALID;
        }
        b.bf_start[i] = nbf;
        nbf += (am[i] + 1) * (am[i] + 2) / 2;
        nprim_end = std::max(nprim_end, prim_off[i] + nprim[i]);
        b.max_am = std::max(b.max_am, am[i]);
        b.max_nprim = std::max(b.max_nprim, nprim[i]);
    }
    b.bf_start[nshells] = nbf;
    b.nbf = nbf;
    b.alpha.assign(alpha, alpha + nprim_end);
    b.coef.assign(coef, coef + nprim_end);
    YG_TRY(upload(b.d_am, b.am.data(), b.am.size() * sizeof(int), c->stream));
    YG_TRY(upload(b.d_nprim, b.nprim.data(), b.nprim.size() * sizeof(int), c->stream));
    YG_TRY(upload(b.d_prim_off, b.prim_off.data(), b.prim_off.size() * sizeof(int), c->stream));
    YG_TRY(upload(b.d_bf_start, b.bf_start.data(), b.bf_start.size() * sizeof(int), c->stream));
    YG_TRY(upload(b.d_xyz, b.xyz.data(), b.xyz.size() * sizeof(double), c->stream));
    YG_TRY(upload(b.d_alpha, b.alpha.data(), b.alpha.size() * sizeof(double), c->stream));
    YG_TRY(upload(b.d_coef, b.coef.data(), b.coef.size() * sizeof(double), c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    if (which == 0) {
        c->nbf = nbf;
        c->n_orth = 0;
    } else {
        c->naux = c->naux_cart = nbf;
        c->aux_sph = 0;
    }
    invalidate_handoffs(c);
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_sph_coefficients(int l, double *T)
{
    if (yg_sph_coefficients(l, T) != YGPU_OK) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_sph_coefficients: l = %d (0..4) / null T", l);
        return YGPU_ERR_INVALID;
    }
    return YGPU_OK;
}

int ygpu_set_aux_spherical(ygpu_ctx_t *c, int on)
try {
    YG_TRY(check_ctx(c));
    if (c->aux.nshells <= 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_aux_spherical: set the auxiliary basis first");
        return YGPU_ERR_INVALID;
    }
    if (on && c->aux.max_am > 4) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_aux_spherical: angular momentum %d beyond g", c->aux.max_am);
        return YGPU_ERR_INVALID;
    }
    if (on) YG_TRY(yg_aux_sph_setup(c));
    else c->naux = c->naux_cart;
    c->aux_sph = on ? 1 : 0;
    // the tensor layout follows from naux: set the sparsity (and build B) again
    c->have_sparsity = 0;
    c->have_df = 0;
    invalidate_handoffs(c);
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_orth_dim(ygpu_ctx_t *c, int n_orth)
try {
    YG_TRY(check_ctx(c));
    if (n_orth < 0 || n_orth > c->nbf) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_orth_dim: %d columns for %d basis functions", n_orth, c->nbf);
        return YGPU_ERR_INVALID;
    }
    c->n_orth = n_orth == c->nbf ? 0 : n_orth;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_atoms(ygpu_ctx_t *c, int natoms, const double *charge, const double *xyz)
try {
    YG_TRY(check_ctx(c));
    if (natoms <= 0 || !charge || !xyz) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_atoms: bad arguments");
        return YGPU_ERR_INVALID;
    }
    c->natoms = natoms;
    YG_TRY(upload(c->d_charge, charge, natoms * sizeof(double), c->stream));
    YG_TRY(upload(c->d_atom_xyz, xyz, 3 * (size_t) natoms * sizeof(double), c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_ovlp_hcore(ygpu_ctx_t *c, double *S, double *H)
try {
    YG_TRY(check_ctx(c));
    if (c->bas.nshells == 0 || c->natoms == 0 || !S || !H) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_ovlp_hcore: basis/atoms not set or null output");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) c->nbf * c->nbf * sizeof(double);
    YG_TRY(c->d_mat0.ensure(ms));
    YG_TRY(c->d_mat1.ensure(ms));
    YG_TRY(c->d_H.ensure(ms));
    YG_TRY(yg_launch_ovlp_hcore(c, c->d_mat0.as<double>(), c->d_H.as<double>()));
    YG_HIP(hipMemcpyAsync(S, c->d_mat0.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipMemcpyAsync(H, c->d_H.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->have_H = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_sym_inv_sqrt(ygpu_ctx_t *c, int n, const double *A, double *X, double *min_eig)
try {
    YG_TRY(check_ctx(c));
    if (n <= 0 || !A || !X) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_sym_inv_sqrt: bad arguments");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) n * n * sizeof(double);
    YG_TRY(c->d_mat0.ensure(ms));
    YG_TRY(c->d_mat1.ensure(ms));
    YG_TRY(c->d_mat2.ensure(ms));
    YG_HIP(hipMemcpyAsync(c->d_mat0.p, A, ms, hipMemcpyHostToDevice, c->stream));
    YG_TRY(inv_sqrt_device(c, n, n, c->d_mat0.as<double>(), c->d_mat1.as<double>(), c->d_mat2.as<double>(), min_eig));
    YG_HIP(hipMemcpyAsync(X, c->d_mat2.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_schwarz(ygpu_ctx_t *c, int quirk_a5, double *sp_scrval, double *bf_pair_scrval, double *df_sp_scrval)
try {
    YG_TRY(check_ctx(c));
    if (c->bas.nshells == 0 || c->aux.nshells == 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_schwarz: basis sets not set");
        return YGPU_ERR_INVALID;
    }
    const int nbf = c->nbf, ns = c->bas.nshells;
    const size_t ms = (size_t) nbf * nbf * sizeof(double);
    YG_TRY(c->d_mat0.ensure(ms));
    YG_TRY(c->d_mat1.ensure(c->aux.nshells * sizeof(double)));
    YG_TRY(yg_launch_schwarz(c, quirk_a5, c->d_mat0.as<double>(), c->d_mat1.as<double>()));
    std::vector<double> bfs((size_t) nbf * nbf), dfs(c->aux.nshells);
    YG_HIP(hipMemcpyAsync(bfs.data(), c->d_mat0.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipMemcpyAsync(dfs.data(), c->d_mat1.p, dfs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    if (bf_pair_scrval) std::memcpy(bf_pair_scrval, bfs.data(), ms);
    if (df_sp_scrval) std::memcpy(df_sp_scrval, dfs.data(), dfs.size() * sizeof(double));
    if (sp_scrval)
        for (int M = 0; M < ns; ++M)
            for (int N = 0; N < ns; ++N) {
                double m = 0.0;
                for (int i = c->bas.bf_start[M]; i < c->bas.bf_start[M + 1]; ++i)
                    for (int j = c->bas.bf_start[N]; j < c->bas.bf_start[N + 1]; ++j)
                        m = std::max(m, bfs[(size_t) i * nbf + j]);
                sp_scrval[(size_t) M * ns + N] = m;
            }
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_sparsity(ygpu_ctx_t *c, const int *mask, int num_uniq_sp, const int *lid, const int *rid,
                      const double *sp_scrval, const double *df_sp_scrval, double scrtol2)
try {
    YG_TRY(check_ctx(c));
    const int nbf = c->nbf, ns = c->bas.nshells;
    if (!mask || num_uniq_sp < 0 || (num_uniq_sp > 0 && (!lid || !rid)) || !sp_scrval || !df_sp_scrval || nbf == 0 ||
        c->naux == 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_sparsity: bad arguments or basis sets not set");
        return YGPU_ERR_INVALID;
    }
    c->naux_pad = (c->naux + ROUND_AUX - 1) / ROUND_AUX * ROUND_AUX;
    c->naux_cart_pad = (c->naux_cart + ROUND_AUX - 1) / ROUND_AUX * ROUND_AUX;
    c->h_mask.assign(mask, mask + (size_t) nbf * nbf);
    c->h_upmask.assign((size_t) nbf * nbf, -1);
    std::vector<int> up_i, up_j;
    long nnz = 0;
    // a pair is kept when either orientation survived the screening: the reference's Schwarz
    // index (quirk A5, TinySCF.c:153) can keep (i, j) and drop (j, i), and it then writes the
    // (j, i) row of pqA at index -1 (build_DF_tensor.c:36-43); the union is the defined reading
    auto kept = [&](int i, int j) { return mask[(size_t) i * nbf + j] >= 0 || mask[(size_t) j * nbf + i] >= 0; };
    for (int i = 0; i < nbf; ++i)
        for (int j = 0; j < nbf; ++j) {
            if (!kept(i, j)) continue;
            ++nnz;
            if (j >= i) {
                c->h_upmask[(size_t) i * nbf + j] = (int) up_i.size();
                up_i.push_back(i);
                up_j.push_back(j);
            }
        }
    for (int i = 0; i < nbf; ++i)
        if (mask[(size_t) i * nbf + i] < 0) {
            ygpu_set_error(YGPU_ERR_INVALID, "diagonal pair (%d,%d) screened out", i, i);
            return YGPU_ERR_INVALID;
        }
    c->n_up = (long) up_i.size();
    c->nnz = nnz;
    std::vector<int> xptr(nbf + 1), xup, xpartner;
    xup.reserve(nnz);
    xpartner.reserve(nnz);
    c->max_row_len = 0;
    for (int i = 0; i < nbf; ++i) {
        xptr[i] = (int) xup.size();
        for (int j = 0; j < nbf; ++j) {
            if (!kept(i, j)) continue;
            const int lo = std::min(i, j), hi = std::max(i, j);
            xup.push_back(c->h_upmask[(size_t) lo * nbf + hi]);
            xpartner.push_back(j);
        }
        c->max_row_len = std::max(c->max_row_len, (int) xup.size() - xptr[i]);
    }
    xptr[nbf] = (int) xup.size();
    std::vector<int> kofs(nbf + 1);
    long ks = 0;
    for (int i = 0; i < nbf; ++i) {
        kofs[i] = (int) ks;
        ks += (xptr[i + 1] - xptr[i] + 31) / 32 * 32;
    }
    kofs[nbf] = (int) ks;
    c->ksum = ks;
    std::vector<int> ktiles;
    for (int i = 0; i < nbf; ++i) {
        const int L = (xptr[i + 1] - xptr[i] + 31) / 32 * 32;
        for (int k0 = 0; k0 < L; k0 += 64) {
            ktiles.push_back(i);
            ktiles.push_back(k0);
        }
    }
    c->nktiles = (int) (ktiles.size() / 2);
    std::vector<int> xorder(nbf);
    for (int i = 0; i < nbf; ++i) xorder[i] = i;
    std::stable_sort(xorder.begin(), xorder.end(),
                     [&](int a, int b) { return xptr[a + 1] - xptr[a] > xptr[b + 1] - xptr[b]; });
    c->num_uniq_sp = num_uniq_sp;
    c->h_uniq_lid.assign(lid, lid + num_uniq_sp);
    c->h_uniq_rid.assign(rid, rid + num_uniq_sp);
    c->h_sp_scrval.assign(sp_scrval, sp_scrval + (size_t) ns * ns);
    c->h_df_sp_scrval.assign(df_sp_scrval, df_sp_scrval + c->aux.nshells);
    c->scrtol2 = scrtol2;
    YG_TRY(upload(c->d_upmask, c->h_upmask.data(), c->h_upmask.size() * sizeof(int), c->stream));
    YG_TRY(upload(c->d_up_i, up_i.data(), std::max<size_t>(1, up_i.size()) * sizeof(int), c->stream));
    YG_TRY(upload(c->d_up_j, up_j.data(), std::max<size_t>(1, up_j.size()) * sizeof(int), c->stream));
    YG_TRY(upload(c->d_xptr, xptr.data(), xptr.size() * sizeof(int), c->stream));
    YG_TRY(upload(c->d_xup, xup.data(), std::max<size_t>(1, xup.size()) * sizeof(int), c->stream));
    YG_TRY(upload(c->d_xpartner, xpartner.data(), std::max<size_t>(1, xpartner.size()) * sizeof(int), c->stream));
    YG_TRY(upload(c->d_kofs, kofs.data(), kofs.size() * sizeof(int), c->stream));
    YG_TRY(upload(c->d_ktiles, ktiles.data(), std::max<size_t>(2, ktiles.size()) * sizeof(int), c->stream));
    YG_TRY(upload(c->d_xorder, xorder.data(), xorder.size() * sizeof(int), c->stream));
    YG_TRY(upload(c->d_sp_scr, c->h_sp_scrval.data(), c->h_sp_scrval.size() * sizeof(double), c->stream));
    YG_TRY(upload(c->d_df_scr, c->h_df_sp_scrval.data(), c->h_df_sp_scrval.size() * sizeof(double), c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->have_sparsity = 1;
    c->have_df = 0;
    invalidate_handoffs(c);
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

static int unpack_to_reference(ygpu_ctx *c, const double *dsrc, double *host_out)
{
    // device packed rows (n_up x naux_pad) -> host reference layout (nnz x naux, both triangles)
    const int nbf = c->nbf, naux = c->naux;
    std::vector<double> up((size_t) c->n_up * c->naux_pad);
    YG_HIP(hipMemcpyAsync(up.data(), dsrc, up.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < nbf; ++i)
        for (int j = 0; j < nbf; ++j) {
            const int r = c->h_mask[(size_t) i * nbf + j];
            if (r < 0) continue;
            const int u = c->h_upmask[(size_t) std::min(i, j) * nbf + std::max(i, j)];
            std::memcpy(host_out + (size_t) r * naux, up.data() + (size_t) u * c->naux_pad, naux * sizeof(double));
        }
    return YGPU_OK;
}

// Work forked onto the side stream is drained when the scope that forked it ends, on the
// error paths too (nothing queued there outlives the call that queued it).
struct SideDrain {
    hipStream_t s;
    ~SideDrain()
    {
        if (s) (void) hipStreamSynchronize(s);
    }
};

// side stream s waits for everything queued on the main stream so far
static int fork_side(ygpu_ctx *c, hipStream_t s)
{
    YG_HIP(hipEventRecord(c->fork, c->stream));
    YG_HIP(hipStreamWaitEvent(s, c->fork, 0));
    return YGPU_OK;
}

// the main stream waits for everything queued on side stream s so far
static int join_side(ygpu_ctx *c, hipStream_t s)
{
    YG_HIP(hipEventRecord(c->join, s));
    YG_HIP(hipStreamWaitEvent(c->stream, c->join, 0));
    return YGPU_OK;
}

// X of the reduced-precision builds into c->d_X, nrow columns of C32 (nbf x nrow, FP32): two
// parts from the k-contiguous blocks (C_occ's laid out here, B's in d_Bt), three parts by the
// gathering X build.  Event ev[8] is recorded right before the X-build kernel, which the "aux K"
// timer then brackets alone.
static int xbuild_mixed(ygpu_ctx *c, int nparts, int nrow, const float *C32, size_t btpart)
{
    const int nbf = c->nbf;
    if (nparts != 2) {
        YG_HIP(hipEventRecord(c->ev[8], c->stream));
        return yg_xbuild_bf(c->stream, nparts, nbf, nrow, (int) c->naux_pad, c->d_xptr.as<int>(), c->d_xup.as<int>(),
                            c->d_xpartner.as<int>(), C32, c->d_B32.as<float>(), c->d_X.p);
    }
    const size_t cpart = (size_t) c->ksum * nrow;
    YG_TRY(c->d_Cp.ensure(2 * cpart * sizeof(__bf16)));
    YG_TRY(yg_gather_split_t(c->stream, c->nktiles, c->d_ktiles.as<int2>(), C32, nrow, nrow, c->d_xpartner.as<int>(),
                             c->d_xptr.as<int>(),
                             c->d_kofs.as<int>(), c->d_Cp.p, cpart));
    YG_HIP(hipEventRecord(c->ev[8], c->stream));
    return yg_xbuild_ps(c->stream, nbf, nrow, (int) c->naux_pad, c->d_xptr.as<int>(), c->d_kofs.as<int>(),
                        c->d_xorder.as<int>(), c->d_Cp.p, cpart, c->d_Bt.p, btpart, c->d_X.p);
}

int ygpu_build_df_tensor_ex(ygpu_ctx_t *c, double *pqA_out, double *Jpq_out, double *df_out, ygpu_df_timing_t *t)
try {
    YG_TRY(check_ctx(c));
    if (!c->have_sparsity) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_df_tensor: call ygpu_set_sparsity first");
        return YGPU_ERR_INVALID;
    }
    const long ld = c->naux_pad;
    // the integrals are made over the Cartesian auxiliary functions (leading dimension ldc); a
    // spherical auxiliary basis transforms them to its own columns before the metric factor
    const long ldc = c->aux_sph ? c->naux_cart_pad : ld;
    const size_t tszc = (size_t) c->n_up * ldc * sizeof(double);
    YG_TRY(c->d_pqA.ensure(tszc));
    YG_TRY(c->d_B.ensure(tszc));
    YG_TRY(c->d_Jpq.ensure((size_t) ld * ld * sizeof(double)));
    YG_TRY(c->d_mat0.ensure((size_t) ldc * ldc * sizeof(double)));
    YG_TRY(c->d_mat1.ensure((size_t) ldc * ldc * sizeof(double)));
    const bool chol = c->metric_mode == YGPU_METRIC_CHOLESKY;
    // in sequence: beside the 3-centre integrals, which fill every CU, the metric's factor (a
    // chain of small launches, or the eigensolver of J^{-1/2}) measured longer by their whole
    // duration (J_PQ on a second stream started only once the integrals were done)
    YG_HIP(hipEventRecord(c->ev[0], c->stream));
    YG_HIP(hipMemsetAsync(c->d_pqA.p, 0, tszc, c->stream));
    YG_TRY(yg_launch_eri3(c, c->d_pqA.as<double>(), c->stream));
    if (c->aux_sph) {
        // (mn|P) T into the other tensor buffer, which then holds the integrals
        YG_TRY(yg_aux_sph_cols(c, c->stream, c->n_up, c->d_pqA.as<double>(), ldc, c->d_B.as<double>(), ld));
        c->d_pqA.swap(c->d_B);
    }
    YG_HIP(hipEventRecord(c->ev[1], c->stream));
    YG_HIP(hipMemsetAsync(c->d_mat0.p, 0, (size_t) ldc * ldc * sizeof(double), c->stream));
    YG_TRY(yg_launch_eri2(c, c->d_mat0.as<double>(), (int) ldc));
    if (c->aux_sph) {
        // T^T J T: the columns of J, then (J symmetric) the columns of (J T)^T = T^T J
        YG_TRY(c->d_mat2.ensure((size_t) ldc * ldc * sizeof(double)));
        YG_TRY(yg_aux_sph_cols(c, c->stream, ldc, c->d_mat0.as<double>(), ldc, c->d_mat1.as<double>(), ld));
        YG_TRY(yg_transpose(c->stream, (int) ldc, (int) ld, c->d_mat1.as<double>(), c->d_mat2.as<double>()));
        YG_TRY(yg_aux_sph_cols(c, c->stream, ld, c->d_mat2.as<double>(), ldc, c->d_mat0.as<double>(), ld));
    }
    YG_HIP(hipEventRecord(c->ev[2], c->stream));
    if (chol) {
        // J = L L^T on the device, W = L^{-1} (lower triangular, identity on the padding), then
        // B = pqA W^T with the zeros of W above its diagonal skipped
        YG_TRY(c->d_info.ensure(sizeof(int)));
        YG_TRY(c->d_mat1.ensure(yg_chol_inv_ws((int) ld) * sizeof(double)));
        YG_HIP(hipMemsetAsync(c->d_info.p, 0, sizeof(int), c->stream));
        YG_TRY(yg_chol_inv_lower(c->stream, c->naux, (int) ld, ld, c->d_mat0.as<double>(), c->d_Jpq.as<double>(), ld,
                                 c->d_info.as<int>(), c->d_mat1.as<double>(), &c->d_gemm_ws));
        YG_HIP(hipEventRecord(c->ev[3], c->stream));
        YG_TRY(yg_dgemm_nt128(c->stream, (int) c->n_up, (int) ld, ld, c->d_pqA.as<double>(), ld,
                              c->d_Jpq.as<double>(), ld, c->d_B.as<double>(), ld, true));
    } else {
        YG_HIP(hipMemsetAsync(c->d_Jpq.p, 0, (size_t) ld * ld * sizeof(double), c->stream));
        double mine = 0.0;
        YG_TRY(inv_sqrt_device(c, c->naux, (int) ld, c->d_mat0.as<double>(), c->d_mat1.as<double>(),
                               c->d_Jpq.as<double>(), &mine));
        YG_HIP(hipEventRecord(c->ev[3], c->stream));
        // B = pqA J^{-1/2}; J^{-1/2} is exactly symmetric (inv_sqrt_device symmetrises it), so its
        // rows serve as the k-contiguous second operand
        YG_TRY(yg_dgemm_nt128(c->stream, (int) c->n_up, (int) ld, ld, c->d_pqA.as<double>(), ld,
                              c->d_Jpq.as<double>(), ld, c->d_B.as<double>(), ld));
    }
    YG_HIP(hipEventRecord(c->ev[4], c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    if (chol) {
        int info = 0;
        YG_HIP(hipMemcpy(&info, c->d_info.p, sizeof(int), hipMemcpyDeviceToHost));
        if (info != 0) {
            ygpu_set_error(YGPU_ERR_NUMERIC, "metric not positive definite: Cholesky pivot %d is not positive", info);
            return YGPU_ERR_NUMERIC;
        }
    }
    c->df_last.t_3c = elapsed(c->ev[0], c->ev[1]) * 1e-3;
    c->df_last.t_2c = elapsed(c->ev[1], c->ev[2]) * 1e-3;
    c->df_last.t_invsqrt = elapsed(c->ev[2], c->ev[3]) * 1e-3;
    c->df_last.t_dfgemm = elapsed(c->ev[3], c->ev[4]) * 1e-3;
    if (t) *t = c->df_last;
    if (Jpq_out) {
        // one contiguous copy: the pitched device-to-host copy runs row by row
        double *packed = c->d_mat1.as<double>();
        YG_HIP(hipMemcpy2DAsync(packed, c->naux * sizeof(double), c->d_Jpq.p, ld * sizeof(double),
                                c->naux * sizeof(double), c->naux, hipMemcpyDeviceToDevice, c->stream));
        YG_HIP(hipMemcpyAsync(Jpq_out, packed, (size_t) c->naux * c->naux * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        YG_HIP(hipStreamSynchronize(c->stream));
    }
    if (pqA_out) YG_TRY(unpack_to_reference(c, c->d_pqA.as<double>(), pqA_out));
    if (df_out) YG_TRY(unpack_to_reference(c, c->d_B.as<double>(), df_out));
    c->d_pqA.release();   // (mn|P) is not needed once B is formed
    c->have_df = 1;
    c->have_B32 = 0;
    c->have_Bt = 0;
    c->kref_valid = 0;
    c->jinc_valid = 0;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_build_df_tensor(ygpu_ctx_t *c, double *Jpq_out, double *df_out, ygpu_df_timing_t *t)
try {
    return ygpu_build_df_tensor_ex(c, nullptr, Jpq_out, df_out, t);
} catch (...) {
    return exception_status();
}

int ygpu_set_df_tensor(ygpu_ctx_t *c, const double *df)
try {
    YG_TRY(check_ctx(c));
    if (!c->have_sparsity || !df) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_df_tensor: call ygpu_set_sparsity first / null tensor");
        return YGPU_ERR_INVALID;
    }
    const int nbf = c->nbf, naux = c->naux;
    const long ld = c->naux_pad;
    std::vector<double> up((size_t) c->n_up * ld, 0.0);
    for (int i = 0; i < nbf; ++i)
        for (int j = i; j < nbf; ++j) {
            const int u = c->h_upmask[(size_t) i * nbf + j];
            if (u < 0) continue;
            // the row of either orientation (one of them may have been screened out)
            int r = c->h_mask[(size_t) i * nbf + j];
            if (r < 0) r = c->h_mask[(size_t) j * nbf + i];
            std::memcpy(up.data() + (size_t) u * ld, df + (size_t) r * naux, naux * sizeof(double));
        }
    YG_TRY(upload(c->d_B, up.data(), up.size() * sizeof(double), c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->have_df = 1;
    c->have_B32 = 0;
    c->have_Bt = 0;
    c->kref_valid = 0;
    c->jinc_valid = 0;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_build_fock(ygpu_ctx_t *c, int n_occ, const double *D, const double *Cocc, const double *H, double *J,
                    double *K, double *F, ygpu_fock_timing_t *t)
try {
    YG_TRY(check_ctx(c));
    // D and C_occ of the last density build, already on the device: both NULL, or in the SCF
    // hand-off mode whenever that build's outputs are there
    const bool resident = (!D && !Cocc) || (c->resident && c->density_valid);
    // Hcore: NULL, or in the hand-off mode, the device copy once one exists
    const bool h_cached = !H || (c->resident && c->have_H);
    if (!c->have_df || n_occ <= 0 || (!resident && (!D || !Cocc)) || !F || (h_cached && !c->have_H) ||
        (resident && !c->density_valid)) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_fock: DF tensor not built, no device density or bad arguments");
        return YGPU_ERR_INVALID;
    }
    const int nbf = c->nbf;
    const size_t ms = (size_t) nbf * nbf * sizeof(double);
    YG_TRY(c->d_D.ensure(ms));
    YG_TRY(c->d_H.ensure(ms));
    YG_TRY(c->d_J.ensure(ms));
    YG_TRY(c->d_K.ensure(ms));
    YG_TRY(c->d_F.ensure(ms));
    YG_TRY(c->d_Cocc.ensure((size_t) nbf * n_occ * sizeof(double)));
    YG_TRY(c->d_X.ensure((size_t) nbf * n_occ * c->naux_pad * sizeof(double)));
    if (!resident) YG_HIP(hipMemcpyAsync(c->d_D.p, D, ms, hipMemcpyHostToDevice, c->stream));
    if (!h_cached) {
        YG_HIP(hipMemcpyAsync(c->d_H.p, H, ms, hipMemcpyHostToDevice, c->stream));
        c->have_H = 1;
    }
    if (!resident)
        YG_HIP(hipMemcpyAsync(c->d_Cocc.p, Cocc, (size_t) nbf * n_occ * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    c->fock_valid = 0;
    float tj1 = 0.f, tj2 = 0.f;
    const long nn = (long) nbf * nbf;
    int mode = c->fock_mode;
    c->incr_fallback = 0;
    if (mode == YGPU_FOCK_INCR) {
        // the difference build needs the C_occ and K of an FP64 build with the same n_occ
        if (!c->kref_valid || c->kref_nocc != n_occ || !c->have_S) {
            mode = YGPU_FOCK_FP64;
            c->incr_fallback = 1;
        } else {
            // orthogonal Q minimising |C Q - C_ref| in the S metric (maximising tr(Q^T A)): the
            // polar factor of A = C^T S C_ref by Newton-Schulz, Q <- Q (3 I - Q^T Q) / 2 (the
            // singular values of A are the cosines of the principal angles between the two
            // occupied spaces)
            const size_t qs = (size_t) n_occ * n_occ * sizeof(double);
            YG_TRY(c->d_SC.ensure((size_t) nbf * n_occ * sizeof(double)));
            YG_TRY(c->d_C1.ensure((size_t) nbf * n_occ * sizeof(double)));
            YG_TRY(c->d_Qa.ensure(qs));
            YG_TRY(c->d_Qb.ensure(qs));
            YG_TRY(c->d_Qt.ensure(qs));
            YG_TRY(c->d_nserr.ensure(sizeof(double)));
            double *Qa = c->d_Qa.as<double>(), *Qb = c->d_Qb.as<double>(), *Qt = c->d_Qt.as<double>();
            YG_TRY(yg_dgemm(c->stream, 0, 0, nbf, n_occ, nbf, 1.0, c->d_S.as<double>(), nbf, c->d_Cocc.as<double>(),
                            n_occ, 0.0, c->d_SC.as<double>(), n_occ, &c->d_gemm_ws));
            YG_TRY(yg_dgemm(c->stream, 1, 0, n_occ, n_occ, nbf, 1.0, c->d_SC.as<double>(), n_occ,
                            c->d_Cref.as<double>(), n_occ, 0.0, Qa, n_occ, &c->d_gemm_ws));
            for (int it = 0; it < YG_NS_ITERS; ++it) {
                YG_TRY(yg_dgemm(c->stream, 1, 0, n_occ, n_occ, n_occ, 1.0, Qa, n_occ, Qa, n_occ, 0.0, Qt, n_occ,
                                &c->d_gemm_ws));
                YG_TRY(yg_ns_shift(c->stream, n_occ, Qt));
                YG_TRY(yg_dgemm(c->stream, 0, 0, n_occ, n_occ, n_occ, 0.5, Qa, n_occ, Qt, n_occ, 0.0, Qb, n_occ,
                                &c->d_gemm_ws));
                std::swap(Qa, Qb);
            }
            YG_TRY(yg_dgemm(c->stream, 1, 0, n_occ, n_occ, n_occ, 1.0, Qa, n_occ, Qa, n_occ, 0.0, Qt, n_occ,
                            &c->d_gemm_ws));
            YG_TRY(yg_dev_identity(c->stream, n_occ, Qt, c->d_nserr.as<double>()));
            YG_HIP(hipMemcpyAsync(&c->incr_align_err, c->d_nserr.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            YG_HIP(hipStreamSynchronize(c->stream));
            if (!(c->incr_align_err < 1e-12)) {
                // the occupied spaces are too far apart for the iteration (a principal angle near
                // 90 degrees): a full FP64 build, which becomes the new reference
                mode = YGPU_FOCK_FP64;
                c->incr_fallback = 1;
            } else {
                // C_occ aligned to the reference: C1 = C Q
                YG_TRY(yg_dgemm(c->stream, 0, 0, nbf, n_occ, n_occ, 1.0, c->d_Cocc.as<double>(), n_occ, Qa, n_occ,
                                0.0, c->d_C1.as<double>(), n_occ, &c->d_gemm_ws));
            }
        }
    }
    if (mode == YGPU_FOCK_MIXED || mode == YGPU_FOCK_MIXED6 || mode == YGPU_FOCK_INCR) {
        const int nparts = mode == YGPU_FOCK_MIXED6 ? 3 : 2;
        if (!c->have_B32) {
            YG_TRY(c->d_B32.ensure((size_t) c->n_up * c->naux_pad * sizeof(float)));
            YG_TRY(yg_to_f32(c->stream, c->d_B.as<double>(), c->d_B32.as<float>(), c->n_up * c->naux_pad));
            c->have_B32 = 1;
        }
        // two-part products: B's rows of every basis function transposed into k-contiguous
        // blocks (both triangles, two BF16 parts) once per DF tensor
        const size_t btpart = (size_t) c->ksum * c->naux_pad;
        if (nparts == 2 && !c->have_Bt) {
            YG_TRY(c->d_Bt.ensure(2 * btpart * sizeof(__bf16)));
            YG_TRY(yg_gather_split_t(c->stream, c->nktiles, c->d_ktiles.as<int2>(), c->d_B32.as<float>(), c->naux_pad,
                                     (int) c->naux_pad,
                                     c->d_xup.as<int>(), c->d_xptr.as<int>(), c->d_kofs.as<int>(), c->d_Bt.p, btpart));
            c->have_Bt = 1;
        }
        YG_TRY(c->d_Dprev.ensure(ms));
        YG_TRY(c->d_Jprev.ensure(ms));
        YG_TRY(c->d_dD.ensure(ms));
        YG_TRY(c->d_Cocc32.ensure((size_t) nbf * n_occ * sizeof(float)));
        // J first on the main stream, or beside the K build on the side stream (the two are
        // independent until F is assembled)
        hipStream_t js = c->j_beside ? c->side : c->stream;
        if (c->j_beside) YG_TRY(fork_side(c, c->side));
        SideDrain drain{c->j_beside ? c->side : nullptr};
        YG_HIP(hipEventRecord(c->ev[5], js));
        if (!c->jinc_valid) {
            // first build of a mixed run: full FP64 J, kept with its D as the base of the increments
            YG_TRY(yg_jbuild(c, js, c->d_D.as<double>(), c->d_J.as<double>(), nullptr, nullptr, nullptr));
            YG_HIP(hipMemcpyAsync(c->d_Jprev.p, c->d_J.p, ms, hipMemcpyDeviceToDevice, js));
            YG_HIP(hipMemcpyAsync(c->d_Dprev.p, c->d_D.p, ms, hipMemcpyDeviceToDevice, js));
            c->jinc_valid = 1;
        } else {
            // J = J_prev + J(D - D_prev) from the FP32 tensor with FP64 arithmetic: the
            // rounding of the tensor enters only through the small difference density
            YG_TRY(yg_delta(js, nn, c->d_D.as<double>(), c->d_Dprev.as<double>(), c->d_dD.as<double>()));
            YG_TRY(yg_jbuild(c, js, c->d_dD.as<double>(), c->d_J.as<double>(), c->d_B32.as<float>(), nullptr,
                             nullptr));
            YG_TRY(yg_accum(js, nn, c->d_J.as<double>(), c->d_Jprev.as<double>()));
        }
        YG_HIP(hipEventRecord(c->ev[3], js));
        if (mode == YGPU_FOCK_INCR) {
            // D - D_ref = (P M^T + M P^T) / 2 with P = C1 + C_ref, M = C1 - C_ref: K(D - D_ref) =
            // (X^M (X^P)^T + X^P (X^M)^T) / 2, BF16-split products whose error scales with |M|;
            // K = K_ref (FP64) + that difference
            YG_TRY(c->d_PM32.ensure((size_t) nbf * 2 * n_occ * sizeof(float)));
            YG_TRY(yg_pm32(c->stream, nbf, n_occ, c->d_C1.as<double>(), c->d_Cref.as<double>(), c->d_PM32.as<float>()));
            YG_TRY(xbuild_mixed(c, nparts, 2 * n_occ, c->d_PM32.as<float>(), btpart));
            YG_HIP(hipEventRecord(c->ev[4], c->stream));
            YG_TRY(yg_incr_k_bf(c->stream, nparts, nbf, n_occ, (int) c->naux_pad, c->d_X.as<float>(),
                                c->d_Kref.as<double>(), c->d_K.as<double>(), c->d_Kpart32));
        } else {
            // K from BF16-split products; X stored pre-split in the X buffer
            YG_TRY(yg_to_f32(c->stream, c->d_Cocc.as<double>(), c->d_Cocc32.as<float>(), (long) nbf * n_occ));
            YG_TRY(xbuild_mixed(c, nparts, n_occ, c->d_Cocc32.as<float>(), btpart));
            YG_HIP(hipEventRecord(c->ev[4], c->stream));
            YG_TRY(yg_syrk_bf(c->stream, nparts, nbf, (long) n_occ * c->naux_pad, c->d_X.as<float>(),
                              c->d_K.as<double>(), c->d_Kpart32));
        }
        YG_HIP(hipEventRecord(c->ev[6], c->stream));
        if (c->j_beside) YG_TRY(join_side(c, c->side));
    } else {
        // J first on the main stream, or beside the K build on the side stream
        YG_TRY(c->d_Dprev.ensure(ms));
        YG_TRY(c->d_Jprev.ensure(ms));
        hipStream_t js = c->j_beside ? c->side : c->stream;
        if (c->j_beside) YG_TRY(fork_side(c, c->side));
        SideDrain drain{c->j_beside ? c->side : nullptr};
        YG_HIP(hipEventRecord(c->ev[5], js));
        YG_TRY(yg_jbuild(c, js, c->d_D.as<double>(), c->d_J.as<double>(), nullptr, nullptr, nullptr));
        YG_HIP(hipEventRecord(c->ev[3], js));
        // this FP64 J and its D are the base of the increments of later FP32-tensor builds
        YG_HIP(hipMemcpyAsync(c->d_Jprev.p, c->d_J.p, ms, hipMemcpyDeviceToDevice, js));
        YG_HIP(hipMemcpyAsync(c->d_Dprev.p, c->d_D.p, ms, hipMemcpyDeviceToDevice, js));
        c->jinc_valid = 1;
        YG_HIP(hipEventRecord(c->ev[8], c->stream));
        YG_TRY(yg_xbuild(c->stream, nbf, n_occ, (int) c->naux_pad, c->d_xptr.as<int>(), c->d_xup.as<int>(),
                         c->d_xpartner.as<int>(), c->d_Cocc.as<double>(), c->d_B.as<double>(), c->d_X.as<double>(),
                         c->max_row_len));
        YG_HIP(hipEventRecord(c->ev[4], c->stream));
        YG_TRY(yg_syrk_upper(c->stream, nbf, (long) n_occ * c->naux_pad, c->d_X.as<double>(), c->d_K.as<double>(),
                             c->d_Kpart));
        YG_HIP(hipEventRecord(c->ev[6], c->stream));
        // this FP64 build is the reference of later difference builds (YGPU_FOCK_INCR)
        YG_TRY(c->d_Cref.ensure((size_t) nbf * n_occ * sizeof(double)));
        YG_TRY(c->d_Kref.ensure(ms));
        YG_HIP(hipMemcpyAsync(c->d_Cref.p, c->d_Cocc.p, (size_t) nbf * n_occ * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
        YG_HIP(hipMemcpyAsync(c->d_Kref.p, c->d_K.p, ms, hipMemcpyDeviceToDevice, c->stream));
        c->kref_valid = 1;
        c->kref_nocc = n_occ;
        if (c->j_beside) YG_TRY(join_side(c, c->side));
    }
    YG_TRY(yg_fock_assemble(c->stream, nbf, c->d_H.as<double>(), c->d_J.as<double>(), c->d_K.as<double>(),
                            c->d_F.as<double>()));
    YG_HIP(hipEventRecord(c->ev[7], c->stream));
    // energy terms: sum D (H + F), and after a difference build the reference term against the F
    // and D of the last full FP64 build, which an FP64 build then records
    const bool eref = mode == YGPU_FOCK_INCR;
    const int neb = yg_energy_blocks();
    YG_TRY(c->d_epart.ensure((size_t) 2 * neb * sizeof(double)));
    YG_TRY(yg_energy_terms(c->stream, nbf, c->d_D.as<double>(), c->d_H.as<double>(), c->d_F.as<double>(),
                           eref ? c->d_Dref.as<double>() : nullptr, eref ? c->d_Fref.as<double>() : nullptr,
                           c->d_epart.as<double>()));
    if (mode == YGPU_FOCK_FP64) {
        YG_TRY(c->d_Fref.ensure(ms));
        YG_TRY(c->d_Dref.ensure(ms));
        YG_HIP(hipMemcpyAsync(c->d_Fref.p, c->d_F.p, ms, hipMemcpyDeviceToDevice, c->stream));
        YG_HIP(hipMemcpyAsync(c->d_Dref.p, c->d_D.p, ms, hipMemcpyDeviceToDevice, c->stream));
    }
    std::vector<double> epart((size_t) 2 * neb);
    YG_HIP(hipMemcpyAsync(epart.data(), c->d_epart.p, epart.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    YG_HIP(hipMemcpyAsync(F, c->d_F.p, ms, hipMemcpyDeviceToHost, c->stream));
    if (J) YG_HIP(hipMemcpyAsync(J, c->d_J.p, ms, hipMemcpyDeviceToHost, c->stream));
    if (K) YG_HIP(hipMemcpyAsync(K, c->d_K.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    // block partials added pairwise in a fixed order
    for (int w = 1; w < neb; w <<= 1)
        for (int b = 0; b + w < neb; b += 2 * w) {
            epart[2 * b] += epart[2 * (b + w)];
            epart[2 * b + 1] += epart[2 * (b + w) + 1];
        }
    c->energy_dhf = epart[0];
    c->energy_ref = eref ? epart[1] : 0.0;
    c->fock_valid = 1;
    ygpu_fock_timing_t tm;
    tj1 = elapsed(c->ev[0], c->ev[1]);
    tj2 = elapsed(c->ev[1], c->ev[2]);
    tm.t_tempJ = tj1 * 1e-3;
    tm.t_Jmat = tj2 * 1e-3;
    tm.t_tempK = elapsed(c->ev[8], c->ev[4]) * 1e-3;   // J: ev 5 .. 3, before or beside the K build
    tm.t_Kmat = elapsed(c->ev[4], c->ev[6]) * 1e-3;
    tm.t_symm = elapsed(c->ev[6], c->ev[7]) * 1e-3;
    tm.t_total = elapsed(c->ev[5], c->ev[7]) * 1e-3;
    if (c->fock_hist.size() < YG_MAX_FOCK_HIST) c->fock_hist.push_back(tm);
    if (t) *t = tm;
    c->last_mode = mode;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_fock_precision(ygpu_ctx_t *c, int mode)
try {
    YG_TRY(check_ctx(c));
    if (mode != YGPU_FOCK_FP64 && mode != YGPU_FOCK_MIXED && mode != YGPU_FOCK_MIXED6 && mode != YGPU_FOCK_INCR) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_fock_precision: unknown mode %d", mode);
        return YGPU_ERR_INVALID;
    }
    // the increments of J continue across a change of arithmetic: every FP64 build leaves its
    // exact J and D as their base
    c->fock_mode = mode;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_metric_factor(ygpu_ctx_t *c, int mode)
try {
    YG_TRY(check_ctx(c));
    if (mode != YGPU_METRIC_CHOLESKY && mode != YGPU_METRIC_INVSQRT) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_metric_factor: unknown mode %d", mode);
        return YGPU_ERR_INVALID;
    }
    c->metric_mode = mode;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

// The density from the projector P onto the n_occ lowest eigenvectors of F' (in d_mat0), by SP2
// purification (ygpu_purify.hip), and the occupied vectors from the previous build's V_prev:
// Y = P V_prev spans the range of P while V_prev^T P V_prev = L L^T is well conditioned (its
// eigenvalues are the squared cosines of the principal angles between the two occupied spaces),
// and V = Y L^{-T} is an orthonormal basis of it, V V^T = P.  d_Cocc = X V, d_VT = V^T.  *ok = 0
// (nothing written to d_Cocc or d_VT) when SP2 did not converge within YG_SP2_KMAX steps or the
// least squared pivot of L is below YG_SP2_MINPIV: the caller then runs the eigensolver.
static int density_purify(ygpu_ctx *c, int n_occ, const double *dX, int *ok)
{
    *ok = 0;
    const int n = c->nbf;
    hipStream_t s = c->stream;
    const long ldx = (n + 15) / 16 * 16;
    // k split of the squares: about two resident workgroups per CU over the upper tiles
    // (128 x 128 from n = 256, 64 x 64 below), chunks of at least 64
    const int tsz = n >= 256 ? 128 : 64, ntb = (n + tsz - 1) / tsz, ntiles = ntb * (ntb + 1) / 2;
    const int want = std::max(1, std::min<int>((n >= 256 ? 512 : 1024) / ntiles, (int) (ldx / 64)));
    long kchunk = (ldx + want - 1) / want;
    kchunk = (kchunk + 15) / 16 * 16;
    const int nsplit = (int) ((ldx + kchunk - 1) / kchunk);
    const int nop = (n_occ + 63) / 64 * 64;
    YG_TRY(c->d_P.ensure((size_t) n * ldx * sizeof(double)));
    YG_TRY(c->d_sp2part.ensure((size_t) nsplit * n * n * sizeof(double)));
    YG_TRY(c->d_sp2sum.ensure(yg_sp2_wsum_doubles(n) * sizeof(double)));
    YG_TRY(c->d_sp2log.ensure(YG_SP2_KMAX * sizeof(double)));
    YG_TRY(c->d_sp2slog.ensure(YG_SP2_KMAX * sizeof(int)));
    YG_TRY(c->d_sp2misc.ensure(4 * sizeof(double)));
    YG_TRY(c->d_Y.ensure((size_t) n * n_occ * sizeof(double)));
    YG_TRY(c->d_VT2.ensure((size_t) n_occ * n * sizeof(double)));
    YG_TRY(c->d_Mo.ensure((size_t) nop * nop * sizeof(double)));
    YG_TRY(c->d_Mw.ensure((size_t) nop * nop * sizeof(double)));
    YG_TRY(c->d_MT.ensure(yg_chol_inv_ws(nop) * sizeof(double)));
    double *P = c->d_P.as<double>(), *misc = c->d_sp2misc.as<double>();
    YG_TRY(yg_sp2_begin(s, n, c->n_orth, c->d_mat0.as<double>(), n, P, ldx, c->d_sp2sum.as<double>(), misc));
    double elog[YG_SP2_KMAX];
    int done = 0, first_small = -1;
    int batch = c->sp2_last_steps > 0 ? c->sp2_last_steps : 24;
    bool conv = false;
    while (done < YG_SP2_KMAX) {
        batch = std::max(1, std::min(batch, YG_SP2_KMAX - done));
        YG_TRY(yg_sp2_steps(s, n, ldx, P, n_occ, c->d_sp2part.as<double>(), kchunk, c->d_sp2sum.as<double>(), done,
                            batch, c->d_sp2log.as<double>(), c->d_sp2slog.as<int>()));
        done += batch;
        YG_HIP(hipMemcpyAsync(elog, c->d_sp2log.p, done * sizeof(double), hipMemcpyDeviceToHost, s));
        YG_HIP(hipStreamSynchronize(s));
        first_small = -1;
        for (int k = 0; k < done; ++k)
            if (std::fabs(elog[k]) < YG_SP2_TOL) {
                first_small = k;
                break;
            }
        // X_k is within about sqrt(|e_k|) of a projector; two more steps square that
        if (first_small >= 0 && done >= first_small + 2) {
            conv = true;
            break;
        }
        batch = first_small >= 0 ? first_small + 2 - done : 2;
    }
    c->density_last_steps = done;
    if (!conv) {
        c->sp2_last_steps = 0;
        return YGPU_OK;
    }
    c->sp2_last_steps = first_small + 2;
    // Y = P V_prev; M = Y^T Y (zero padding to nop) = L L^T; V^T = L^{-1} Y^T
    YG_TRY(yg_dgemm(s, 0, 1, n, n_occ, n, 1.0, P, (int) ldx, c->d_VT.as<double>(), n, 0.0, c->d_Y.as<double>(), n_occ,
                    &c->d_gemm_ws));
    YG_HIP(hipMemsetAsync(c->d_Mo.p, 0, (size_t) nop * nop * sizeof(double), s));
    YG_TRY(yg_dgemm(s, 1, 0, n_occ, n_occ, n, 1.0, c->d_Y.as<double>(), n_occ, c->d_Y.as<double>(), n_occ, 0.0,
                    c->d_Mo.as<double>(), nop, &c->d_gemm_ws));
    YG_TRY(c->d_info.ensure(sizeof(int)));
    YG_HIP(hipMemsetAsync(c->d_info.p, 0, sizeof(int), s));
    YG_TRY(yg_chol_inv_lower(s, n_occ, nop, nop, c->d_Mo.as<double>(), c->d_Mw.as<double>(), nop,
                             c->d_info.as<int>(), c->d_MT.as<double>(), &c->d_gemm_ws));
    YG_TRY(yg_min_diag_sq(s, n_occ, nop, c->d_Mo.as<double>(), misc + 2));
    YG_TRY(yg_dgemm(s, 0, 1, n_occ, n, n_occ, 1.0, c->d_Mw.as<double>(), nop, c->d_Y.as<double>(), n_occ, 0.0,
                    c->d_VT2.as<double>(), n, &c->d_gemm_ws));
    int info = 0;
    double minpiv = 0.0;
    YG_HIP(hipMemcpyAsync(&info, c->d_info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    YG_HIP(hipMemcpyAsync(&minpiv, misc + 2, sizeof(double), hipMemcpyDeviceToHost, s));
    YG_HIP(hipStreamSynchronize(s));
    if (info != 0 || !(minpiv > YG_SP2_MINPIV)) return YGPU_OK;
    c->d_VT.swap(c->d_VT2);
    YG_TRY(yg_dgemm(s, 0, 1, n, n_occ, n, 1.0, dX, n, c->d_VT.as<double>(), n, 0.0, c->d_Cocc.as<double>(), n_occ,
                    &c->d_gemm_ws));
    *ok = 1;
    return YGPU_OK;
}

int ygpu_build_density(ygpu_ctx_t *c, int n_occ, const double *Fo, const double *X, double *D, double *Cocc,
                       double *eigval, int *ev_idx)
try {
    YG_TRY(check_ctx(c));
    const int nbf = c->nbf;
    // F_orth of the last DIIS step and the cached X: NULL, or in the SCF hand-off mode whenever
    // the device holds them
    const bool fo_dev = !Fo || (c->resident && c->fcur_valid);
    const bool x_dev = !X || (c->resident && c->have_S);
    if (nbf <= 0 || n_occ <= 0 || n_occ > nbf || (fo_dev && !c->fcur_valid) || (x_dev && !c->have_S) || !D || !Cocc ||
        !eigval) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_density: bad arguments / no device F_orth");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) nbf * nbf * sizeof(double);
    YG_TRY(c->d_mat0.ensure(ms));
    YG_TRY(c->d_mat1.ensure(ms));
    YG_TRY(c->d_mat2.ensure(ms));
    YG_TRY(c->d_D.ensure(ms));
    YG_TRY(c->d_Cocc.ensure((size_t) nbf * n_occ * sizeof(double)));
    YG_TRY(c->d_eig.ensure(nbf * sizeof(double)));
    YG_TRY(c->d_evE.ensure(nbf * sizeof(double)));
    YG_TRY(c->d_info.ensure(sizeof(int)));
    // Fo NULL: the last DIIS transform or extrapolation, already on the device
    if (!fo_dev) YG_HIP(hipMemcpyAsync(c->d_mat0.p, Fo, ms, hipMemcpyHostToDevice, c->stream));
    else YG_HIP(hipMemcpyAsync(c->d_mat0.p, c->d_Fcur.p, ms, hipMemcpyDeviceToDevice, c->stream));
    // directions outside the orthogonaliser's columns (zero rows and columns of F') moved far
    // above the spectrum, so that they are never occupied
    YG_TRY(yg_shift_tail_diag(c->stream, nbf, c->n_orth, c->d_mat0.as<double>(), YGPU_ORTH_SHIFT));
    c->fock_valid = 0;                          // d_D and d_Cocc are overwritten below
    c->density_valid = 0;
    const double *dX = c->d_Xo.as<double>();
    if (!x_dev) {
        YG_HIP(hipMemcpyAsync(c->d_mat1.p, X, ms, hipMemcpyHostToDevice, c->stream));
        dX = c->d_mat1.as<double>();
    }
    int used = YGPU_DENSITY_EIGEN;
    c->density_last_steps = 0;
    if (c->density_method == YGPU_DENSITY_PURIFY && c->have_vt && c->vt_nocc == n_occ) {
        int ok = 0;
        YG_TRY(density_purify(c, n_occ, dX, &ok));
        if (ok) used = YGPU_DENSITY_PURIFY;
    }
    // the eigensolver leaves the occupied vectors' V^T (n_occ x nbf) at the start of vsrc
    const double *vsrc = nullptr;
    if (used == YGPU_DENSITY_EIGEN) YG_TRY(ensure_syevd_workspace(c, nbf, nbf));
    // eigensolver route: one workgroup up to YG_EIG_NMAX_SMALL, the blocked reduction up to
    // YG_SYTRD_BLOCKED_NMAX, rocSOLVER dsyevd beyond (and when the small route's inverse
    // iteration leaves a vector above its residual bound)
    int route = used == YGPU_DENSITY_PURIFY ? -1 : nbf <= YG_EIG_NMAX_SMALL ? 0 : nbf <= YG_SYTRD_BLOCKED_NMAX ? 1 : 2;
    if (route == 0) {
        // F' = Q T Q^T on one workgroup; all eigenvalues of T by multisection and its n_occ
        // lowest eigenvectors by inverse iteration; only those are carried back:
        // V_occ = Q Z_occ in place of Z_occ (columns of d_mat2, column-major)
        const size_t nvp = (size_t) (n_occ + 63) / 64 * 64;
        YG_TRY(c->d_tau.ensure((size_t) nbf * sizeof(double)));
        YG_TRY(c->d_tri.ensure((nvp + 2) * nbf * sizeof(double)));
        YG_TRY(c->d_Fsave.ensure(ms));
        YG_HIP(hipMemcpyAsync(c->d_Fsave.p, c->d_mat0.p, ms, hipMemcpyDeviceToDevice, c->stream));
        double *tri_ws = c->d_tri.as<double>();
        double *dT = tri_ws + nvp * nbf, *eT = dT + nbf;
        YG_HIP(hipMemsetAsync(c->d_info.p, 0, sizeof(int), c->stream));
        YG_TRY(yg_sytrd_small(c->stream, nbf, c->d_mat0.as<double>(), dT, eT, c->d_tau.as<double>()));
        YG_TRY(yg_tridiag_eig(c->stream, nbf, n_occ, dT, eT, c->d_eig.as<double>(), c->d_mat2.as<double>(), tri_ws,
                              c->d_info.as<int>()));
        int resid = 0;
        YG_HIP(hipMemcpyAsync(&resid, c->d_info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        YG_HIP(hipStreamSynchronize(c->stream));
        if (resid != 0) {
            // F' again into d_mat0 for the general eigensolver
            YG_HIP(hipMemcpyAsync(c->d_mat0.p, c->d_Fsave.p, ms, hipMemcpyDeviceToDevice, c->stream));
            YG_HIP(hipMemsetAsync(c->d_info.p, 0, sizeof(int), c->stream));
            c->eig_resid_fallbacks++;
            route = 2;
        } else {
            YG_TRY(yg_backtransform_cols(c->stream, nbf, n_occ, c->d_mat0.as<double>(), nbf, c->d_tau.as<double>(),
                                         c->d_mat2.as<double>(), nbf));
            // C_occ = X V_occ
            YG_TRY(yg_dgemm(c->stream, 0, 1, nbf, n_occ, nbf, 1.0, dX, nbf, c->d_mat2.as<double>(), nbf, 0.0,
                            c->d_Cocc.as<double>(), n_occ, &c->d_gemm_ws));
            vsrc = c->d_mat2.as<double>();
        }
    }
    if (route == 1) {
        // blocked reduction, all eigenpairs of T, Q applied to the n_occ lowest only: their
        // columns of d_mat2 (column-major) are the rows of V_occ^T
        YG_TRY(c->d_tau.ensure((size_t) nbf * sizeof(double)));
        YG_TRY(c->d_tri.ensure(yg_sytrd_blocked_ws(nbf)));
        YG_TRY(yg_sytrd_blocked(c->stream, nbf, nbf, c->d_mat0.as<double>(), c->d_eig.as<double>(),
                                c->d_evE.as<double>(), c->d_tau.as<double>(), c->d_tri.as<double>()));
        YG_TRY(solver_tridiag_to_full(c, nbf, nbf, c->d_mat0.as<double>(), c->d_mat2.as<double>(), n_occ));
        YG_TRY(yg_dgemm(c->stream, 0, 1, nbf, n_occ, nbf, 1.0, dX, nbf, c->d_mat2.as<double>(), nbf, 0.0,
                        c->d_Cocc.as<double>(), n_occ, &c->d_gemm_ws));
        vsrc = c->d_mat2.as<double>();
    }
    if (route == 2) {
        rocblas_status st = rocsolver_dsyevd(c->rb, rocblas_evect_original, rocblas_fill_upper, nbf,
                                             c->d_mat0.as<double>(), nbf, c->d_eig.as<double>(),
                                             c->d_evE.as<double>(), c->d_info.as<int>());
        if (st != rocblas_status_success) {
            ygpu_set_error(YGPU_ERR_SOLVER, "rocsolver_dsyevd(n=%d) returned %d", nbf, (int) st);
            return YGPU_ERR_SOLVER;
        }
        // C = X V: with V column-major in d_mat0, V(k,n) = d_mat0[n*nbf+k] = op_T access
        YG_TRY(yg_dgemm(c->stream, 0, 1, nbf, nbf, nbf, 1.0, dX, nbf, c->d_mat0.as<double>(), nbf, 0.0,
                        c->d_mat2.as<double>(), nbf, &c->d_gemm_ws));
        YG_TRY(yg_gather_cols(c->stream, nbf, n_occ, c->d_mat2.as<double>(), nullptr, c->d_Cocc.as<double>()));
        vsrc = c->d_mat0.as<double>();
    }
    if (vsrc) {
        // the occupied vectors of this build start the purification of the next
        YG_TRY(c->d_VT.ensure((size_t) n_occ * nbf * sizeof(double)));
        YG_HIP(hipMemcpyAsync(c->d_VT.p, vsrc, (size_t) n_occ * nbf * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
        c->have_vt = 1;
        c->vt_nocc = n_occ;
    }
    c->density_last_method = used;
    YG_TRY(yg_dgemm(c->stream, 0, 1, nbf, nbf, n_occ, 1.0, c->d_Cocc.as<double>(), n_occ, c->d_Cocc.as<double>(), n_occ,
                    0.0, c->d_D.as<double>(), nbf, &c->d_gemm_ws));
    int info = 0;
    if (used == YGPU_DENSITY_EIGEN) {
        YG_HIP(hipMemcpyAsync(&info, c->d_info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        YG_HIP(hipMemcpyAsync(eigval, c->d_eig.p, nbf * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    YG_HIP(hipMemcpyAsync(D, c->d_D.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipMemcpyAsync(Cocc, c->d_Cocc.p, (size_t) nbf * n_occ * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    if (info != 0) {
        ygpu_set_error(YGPU_ERR_SOLVER, "symmetric eigensolver (n=%d) info=%d", nbf, info);
        return YGPU_ERR_SOLVER;
    }
    if (ev_idx)
        for (int i = 0; i < nbf; ++i) ev_idx[i] = i;
    c->density_valid = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_overlap(ygpu_ctx_t *c, const double *S, const double *X)
try {
    YG_TRY(check_ctx(c));
    if (!S || !X || c->nbf <= 0) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_overlap: bad arguments or basis not set");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) c->nbf * c->nbf * sizeof(double);
    YG_TRY(upload(c->d_S, S, ms, c->stream));
    YG_TRY(upload(c->d_Xo, X, ms, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->have_S = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

// F_orth = X^T F X into dOut (device), F already on the device in dF
static int orth_transform(ygpu_ctx *c, const double *dF, double *dOut)
{
    const int n = c->nbf;
    YG_TRY(yg_dgemm(c->stream, 1, 0, n, n, n, 1.0, c->d_Xo.as<double>(), n, dF, n, 0.0, c->d_T1.as<double>(), n,
                    &c->d_gemm_ws));
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, c->d_T1.as<double>(), n, c->d_Xo.as<double>(), n, 0.0, dOut, n,
                    &c->d_gemm_ws));
    return YGPU_OK;
}

static int diis_buffers(ygpu_ctx *c)
{
    const size_t ms = (size_t) c->nbf * c->nbf * sizeof(double);
    YG_TRY(c->d_Fin.ensure(ms));
    YG_TRY(c->d_Din.ensure(ms));
    YG_TRY(c->d_T1.ensure(ms));
    YG_TRY(c->d_T2.ensure(ms));
    YG_TRY(c->d_Fcur.ensure(ms));
    YG_TRY(c->d_diisR.ensure(ms * YGPU_MAX_DIIS));
    YG_TRY(c->d_diisF.ensure(ms * YGPU_MAX_DIIS));
    return YGPU_OK;
}

int ygpu_diis_transform(ygpu_ctx_t *c, const double *F, double *F_orth)
try {
    YG_TRY(check_ctx(c));
    // F NULL, or in the SCF hand-off mode whenever it is there: the last Fock build's F
    const bool f_dev = !F || (c->resident && c->fock_valid);
    if (!c->have_S || (f_dev && !c->fock_valid) || !F_orth) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_diis_transform: call ygpu_set_overlap first / no device Fock / null argument");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) c->nbf * c->nbf * sizeof(double);
    YG_TRY(diis_buffers(c));
    const double *dF = c->d_F.as<double>();     // F NULL: the last Fock build's
    if (!f_dev) {
        YG_HIP(hipMemcpyAsync(c->d_Fin.p, F, ms, hipMemcpyHostToDevice, c->stream));
        dF = c->d_Fin.as<double>();
    }
    YG_TRY(orth_transform(c, dF, c->d_Fcur.as<double>()));
    YG_HIP(hipMemcpyAsync(F_orth, c->d_Fcur.p, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->fcur_valid = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_diis_update(ygpu_ctx_t *c, int slot, int len, const double *F, const double *D, double *dots,
                     double *F_orth)
try {
    YG_TRY(check_ctx(c));
    // F / D NULL, or in the SCF hand-off mode whenever they are there: the last Fock build's F
    // and the density it was built from
    const bool fd_dev = !F || !D || (c->resident && c->fock_valid);
    if (!c->have_S || (fd_dev && !c->fock_valid) || !dots || slot < 0 || slot >= YGPU_MAX_DIIS ||
        len <= slot || len > YGPU_MAX_DIIS) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_diis_update: bad arguments (slot %d, len %d)", slot, len);
        return YGPU_ERR_INVALID;
    }
    const int n = c->nbf;
    const long nn = (long) n * n;
    const size_t ms = (size_t) nn * sizeof(double);
    YG_TRY(diis_buffers(c));
    const int nchunk = 64;
    YG_TRY(c->d_dots.ensure((size_t) (YGPU_MAX_DIIS + 1) * nchunk * sizeof(double)));
    const double *dF = c->d_F.as<double>(), *dD = c->d_D.as<double>();
    if (!fd_dev) {
        YG_HIP(hipMemcpyAsync(c->d_Fin.p, F, ms, hipMemcpyHostToDevice, c->stream));
        dF = c->d_Fin.as<double>();
        YG_HIP(hipMemcpyAsync(c->d_Din.p, D, ms, hipMemcpyHostToDevice, c->stream));
        dD = c->d_Din.as<double>();
    }
    double *T1 = c->d_T1.as<double>(), *T2 = c->d_T2.as<double>();
    double *R = c->d_diisR.as<double>() + (size_t) slot * nn;
    double *F0 = c->d_diisF.as<double>() + (size_t) slot * nn;
    // FDS = F D S (DIIS.c:51-54); R = X^T (FDS - FDS^T) X (DIIS.c:57-65)
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, dF, n, dD, n, 0.0, T1, n, &c->d_gemm_ws));
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, T1, n, c->d_S.as<double>(), n, 0.0, T2, n, &c->d_gemm_ws));
    YG_TRY(yg_antisym(c->stream, n, T2, T1));
    YG_TRY(yg_dgemm(c->stream, 1, 0, n, n, n, 1.0, c->d_Xo.as<double>(), n, T1, n, 0.0, T2, n,
                    &c->d_gemm_ws));
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, T2, n, c->d_Xo.as<double>(), n, 0.0, R, n,
                    &c->d_gemm_ws));
    // with the chunk maxima of |R| (this slot's max |R|, for ygpu_diis_error) from the same pass
    double *dmax = c->d_dots.as<double>() + (size_t) YGPU_MAX_DIIS * nchunk;
    YG_TRY(yg_dots(c->stream, nn, R, c->d_diisR.as<double>(), nn, len, c->d_dots.as<double>(), nchunk, dmax));
    YG_TRY(orth_transform(c, dF, F0));
    std::vector<double> part((size_t) (len + 1) * nchunk);
    YG_HIP(hipMemcpyAsync(part.data(), c->d_dots.p, (size_t) len * nchunk * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    YG_HIP(hipMemcpyAsync(part.data() + (size_t) len * nchunk, dmax, nchunk * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    if (F_orth) YG_HIP(hipMemcpyAsync(F_orth, F0, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    for (int j = 0; j < len; ++j) {
        double sacc = 0.0;
        for (int k = 0; k < nchunk; ++k) sacc += part[(size_t) j * nchunk + k];
        dots[j] = sacc;
    }
    double emax = 0.0;
    for (int k = 0; k < nchunk; ++k) emax = std::max(emax, part[(size_t) len * nchunk + k]);
    c->diis_err[slot] = emax;
    c->diis_err_ok[slot] = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_diis_error(ygpu_ctx_t *c, int slot, double *max_abs)
try {
    YG_TRY(check_ctx(c));
    if (slot < 0 || slot >= YGPU_MAX_DIIS || !max_abs || !c->d_diisR.p) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_diis_error: bad arguments (slot %d)", slot);
        return YGPU_ERR_INVALID;
    }
    if (c->diis_err_ok[slot]) {               // formed by the slot's update
        *max_abs = c->diis_err[slot];
        return YGPU_OK;
    }
    const long nn = (long) c->nbf * c->nbf;
    YG_TRY(c->d_dots.ensure(sizeof(double)));
    YG_TRY(yg_maxabs(c->stream, nn, c->d_diisR.as<double>() + (size_t) slot * nn, c->d_dots.as<double>()));
    YG_HIP(hipMemcpyAsync(max_abs, c->d_dots.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_diis_extrapolate(ygpu_ctx_t *c, int len, const double *coef, double *F_out)
try {
    YG_TRY(check_ctx(c));
    if (len <= 0 || len > YGPU_MAX_DIIS || !coef || !F_out || !c->d_diisF.p) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_diis_extrapolate: bad arguments (len %d)", len);
        return YGPU_ERR_INVALID;
    }
    const long nn = (long) c->nbf * c->nbf;
    YG_TRY(c->d_Fcur.ensure(nn * sizeof(double)));
    YG_TRY(yg_lincomb(c->stream, nn, c->d_diisF.as<double>(), nn, len, coef, c->d_Fcur.as<double>()));
    YG_HIP(hipMemcpyAsync(F_out, c->d_Fcur.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    c->fcur_valid = 1;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_dgemm(ygpu_ctx_t *c, int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc)
try {
    YG_TRY(check_ctx(c));
    if (m < 0 || n < 0 || k < 0 || !A || !B || !C) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_dgemm: bad arguments");
        return YGPU_ERR_INVALID;
    }
    if (m == 0 || n == 0) return YGPU_OK;
    const size_t arows = ta ? k : m, brows = tb ? n : k;
    const size_t asz = arows * lda * sizeof(double), bsz = brows * ldb * sizeof(double), csz = (size_t) m * ldc * sizeof(double);
    YG_TRY(c->d_gA.ensure(asz));
    YG_TRY(c->d_gB.ensure(bsz));
    YG_TRY(c->d_gC.ensure(csz));
    YG_HIP(hipMemcpyAsync(c->d_gA.p, A, asz, hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(c->d_gB.p, B, bsz, hipMemcpyHostToDevice, c->stream));
    if (beta != 0.0) YG_HIP(hipMemcpyAsync(c->d_gC.p, C, csz, hipMemcpyHostToDevice, c->stream));
    YG_TRY(yg_dgemm(c->stream, ta, tb, m, n, k, alpha, c->d_gA.as<double>(), lda, c->d_gB.as<double>(), ldb, beta,
                    c->d_gC.as<double>(), ldc, &c->d_gemm_ws));
    YG_HIP(hipMemcpyAsync(C, c->d_gC.p, csz, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_project_guess(ygpu_ctx_t *c, int n_occ, const double *Dg, double *D, double *Cocc, double *eigval)
try {
    YG_TRY(check_ctx(c));
    const int n = c->nbf;
    if (!c->have_S || !Dg || !D || !Cocc || !eigval || n_occ <= 0 || n_occ > n) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_project_guess: bad arguments or no cached S / X");
        return YGPU_ERR_INVALID;
    }
    const size_t ms = (size_t) n * n * sizeof(double);
    YG_TRY(c->d_gA.ensure(ms));
    YG_TRY(c->d_gB.ensure(ms));
    YG_TRY(c->d_gC.ensure(ms));
    YG_TRY(c->d_Fcur.ensure(ms));
    YG_HIP(hipMemcpyAsync(c->d_gA.p, Dg, ms, hipMemcpyHostToDevice, c->stream));
    // S^{1/2} = S X; P = -(S X)^T D (S X): its lowest eigenvalues are the largest occupations
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, c->d_S.as<double>(), n, c->d_Xo.as<double>(), n, 0.0,
                    c->d_gB.as<double>(), n, &c->d_gemm_ws));
    YG_TRY(yg_dgemm(c->stream, 0, 0, n, n, n, 1.0, c->d_gA.as<double>(), n, c->d_gB.as<double>(), n, 0.0,
                    c->d_gC.as<double>(), n, &c->d_gemm_ws));
    YG_TRY(yg_dgemm(c->stream, 1, 0, n, n, n, -1.0, c->d_gB.as<double>(), n, c->d_gC.as<double>(), n, 0.0,
                    c->d_Fcur.as<double>(), n, &c->d_gemm_ws));
    const int method = c->density_method;
    c->density_method = YGPU_DENSITY_EIGEN;
    c->fcur_valid = 1;
    const int st = ygpu_build_density(c, n_occ, nullptr, nullptr, D, Cocc, eigval, nullptr);
    c->fcur_valid = 0;
    c->density_method = method;
    return st;
} catch (...) {
    return exception_status();
}

int ygpu_fock_energy(ygpu_ctx_t *c, double *e_dhf, double *e_ref)
try {
    YG_TRY(check_ctx(c));
    if (!c->fock_valid) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_fock_energy: no Fock build since the last density build");
        return YGPU_ERR_INVALID;
    }
    if (e_dhf) *e_dhf = c->energy_dhf;
    if (e_ref) *e_ref = c->energy_ref;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_density_method(ygpu_ctx_t *c, int method)
try {
    YG_TRY(check_ctx(c));
    if (method != YGPU_DENSITY_EIGEN && method != YGPU_DENSITY_PURIFY) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_set_density_method: unknown method %d", method);
        return YGPU_ERR_INVALID;
    }
    c->density_method = method;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_density_last_info(ygpu_ctx_t *c, int *method_used, int *sp2_steps)
try {
    YG_TRY(check_ctx(c));
    if (method_used) *method_used = c->density_last_method;
    if (sp2_steps) *sp2_steps = c->density_last_steps;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_fock_last_info(ygpu_ctx_t *c, int *mode_used, double *align_err)
try {
    YG_TRY(check_ctx(c));
    if (mode_used) *mode_used = c->last_mode;
    if (align_err) *align_err = c->last_mode == YGPU_FOCK_INCR ? c->incr_align_err : 0.0;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_set_scf_resident(ygpu_ctx_t *c, int on)
try {
    YG_TRY(check_ctx(c));
    c->resident = on ? 1 : 0;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_get_fock_timings(ygpu_ctx_t *c, int max, ygpu_fock_timing_t *out, int *count)
try {
    YG_TRY(check_ctx(c));
    if (max < 0 || (max > 0 && !out)) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_get_fock_timings: bad arguments");
        return YGPU_ERR_INVALID;
    }
    const int n = (int) c->fock_hist.size();
    for (int k = 0; k < n && k < max; ++k) out[k] = c->fock_hist[k];
    if (count) *count = n;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_get_df_timing(ygpu_ctx_t *c, ygpu_df_timing_t *out)
try {
    YG_TRY(check_ctx(c));
    if (!out) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_get_df_timing: null output");
        return YGPU_ERR_INVALID;
    }
    *out = c->df_last;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_clear_timings(ygpu_ctx_t *c)
try {
    YG_TRY(check_ctx(c));
    c->fock_hist.clear();
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

int ygpu_layout_info(ygpu_ctx_t *c, long *n_up, long *nnz, int *naux_pad)
try {
    YG_TRY(check_ctx(c));
    if (n_up) *n_up = c->n_up;
    if (nnz) *nnz = c->nnz;
    if (naux_pad) *naux_pad = c->naux_pad;
    return YGPU_OK;
} catch (...) {
    return exception_status();
}

}  // extern "C"
