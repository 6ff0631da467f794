// ygpu_noalloc.hip -- the FP64 Fock build without device allocation.
//
// ygpu_build_fock grows per-engine device buffers on first use (device copies
// of D, C_occ, Hcore, J, K, F; the X = C_occ^T B intermediate; the split-K and
// J-pass partials).  Callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
// call ygpu_build_fock_noalloc instead: the same kernels, with every one of
// those buffers carved from one statically sized device array, `scratch`,
// which each GPU holds from the library's load until the process ends.
//
// The array's name is a host-side stand-in, not a device address: the
// address is taken once per device with hipGetSymbolAddress (with that device
// current, from ygpu_create) and only that address is handed to kernels and
// copies.  Every engine on a GPU shares that GPU's array, which keeps the last
// call's contents; calls that overlap in time on one GPU race.
#include "ygpu_internal.h"

#include <mutex>

namespace {

__device__ __align__(256) unsigned char scratch[YGPU_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[MAX_DEVICES] = {};

constexpr size_t ALIGN = 256;

size_t align_up(size_t n) { return (n + ALIGN - 1) / ALIGN * ALIGN; }

}  // namespace

int yg_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) {
        ygpu_set_error(YGPU_ERR_INVALID, "scratch: device %d beyond the %d supported", device, MAX_DEVICES);
        return YGPU_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return YGPU_OK;
    int cur = -1;
    YG_HIP(hipGetDevice(&cur));
    if (cur != device) {
        ygpu_set_error(YGPU_ERR_INVALID, "scratch: device %d is not current (%d is)", device, cur);
        return YGPU_ERR_INVALID;
    }
    void *p = nullptr;
    YG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        ygpu_set_error(YGPU_ERR_HIP, "scratch: no device address for device %d", device);
        return YGPU_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
    return YGPU_OK;
}

extern "C" int ygpu_build_fock_noalloc(ygpu_ctx_t *c, int n_occ, const double *D, const double *Cocc,
                                       const double *H, double *J, double *K, double *F)
try {
    if (!c) {
        ygpu_set_error(YGPU_ERR_INVALID, "null context");
        return YGPU_ERR_INVALID;
    }
    YG_HIP(hipSetDevice(c->device));
    if (!c->have_df || n_occ <= 0 || n_occ > c->nbf || !D || !Cocc || !H || !F) {
        ygpu_set_error(YGPU_ERR_INVALID,
                       "ygpu_build_fock_noalloc: DF tensor not built or bad arguments (D, Cocc, Hcore and F are "
                       "required host matrices)");
        return YGPU_ERR_INVALID;
    }
    if (c->fock_mode != YGPU_FOCK_FP64) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_fock_noalloc: FP64 arithmetic only (set YGPU_FOCK_FP64)");
        return YGPU_ERR_INVALID;
    }
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        base = (unsigned char *) g_scratch_ptr[c->device];
    }
    if (!base) {
        ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_fock_noalloc: scratch of device %d not set up", c->device);
        return YGPU_ERR_INVALID;
    }
    const int nbf = c->nbf;
    const long ld = c->naux_pad;
    const size_t ms = (size_t) nbf * nbf * sizeof(double);
    const long kdim = (long) n_occ * ld;
    long kchunk = 0;
    const int nsplit = yg_syrk_upper_split(nbf, kdim, &kchunk);
    const long ngroup = yg_jbuild_groups(c->n_up, ld, false);
    // the layout of the scratch array for this call
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += align_up(bytes); return o; };
    const size_t oD = take(ms), oH = take(ms), oJ = take(ms), oK = take(ms), oF = take(ms);
    const size_t oC = take((size_t) nbf * n_occ * sizeof(double));
    const size_t oTj = take((size_t) ld * sizeof(double));
    const size_t oTp = take((size_t) ngroup * ld * sizeof(double));
    const size_t oW = take((size_t) c->n_up * sizeof(double));
    const size_t oX = take((size_t) nbf * kdim * sizeof(double));
    const size_t oKp = take((size_t) nsplit * ms);
    if (off > (size_t) YGPU_NOALLOC_SCRATCH_BYTES) {
        ygpu_set_error(YGPU_ERR_INVALID,
                       "ygpu_build_fock_noalloc: nbf %d, n_occ %d, naux %d need %zu bytes of scratch, more than the "
                       "%zu available (use ygpu_build_fock)",
                       nbf, n_occ, c->naux, off, (size_t) YGPU_NOALLOC_SCRATCH_BYTES);
        return YGPU_ERR_INVALID;
    }
    double *dD = (double *) (base + oD), *dH = (double *) (base + oH), *dJ = (double *) (base + oJ);
    double *dK = (double *) (base + oK), *dF = (double *) (base + oF), *dC = (double *) (base + oC);
    double *dTj = (double *) (base + oTj), *dTp = (double *) (base + oTp), *dX = (double *) (base + oX);
    double *dKp = (double *) (base + oKp), *dW = (double *) (base + oW);
    // this call's inputs overwrite the device hand-off buffers' role: nothing resident stays valid
    c->fock_valid = 0;
    c->density_valid = 0;
    YG_HIP(hipMemcpyAsync(dD, D, ms, hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(dH, H, ms, hipMemcpyHostToDevice, c->stream));
    YG_HIP(hipMemcpyAsync(dC, Cocc, (size_t) nbf * n_occ * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // J on the side stream (after the uploads), K on the main stream, as in ygpu_build_fock
    YG_HIP(hipEventRecord(c->fork, c->stream));
    YG_HIP(hipStreamWaitEvent(c->side, c->fork, 0));
    int st = yg_jbuild_ws(c, c->side, dD, dJ, nullptr, dTp, dTj, dW, ngroup);
    if (st == YGPU_OK)
        st = yg_xbuild(c->stream, nbf, n_occ, (int) ld, c->d_xptr.as<int>(), c->d_xup.as<int>(),
                       c->d_xpartner.as<int>(), dC, c->d_B.as<double>(), dX, c->max_row_len);
    if (st == YGPU_OK) st = yg_syrk_upper_ws(c->stream, nbf, kdim, dX, dK, dKp, kchunk);
    if (st != YGPU_OK) {
        (void) hipStreamSynchronize(c->side);
        return st;
    }
    YG_HIP(hipEventRecord(c->join, c->side));
    YG_HIP(hipStreamWaitEvent(c->stream, c->join, 0));
    YG_TRY(yg_fock_assemble(c->stream, nbf, dH, dJ, dK, dF));
    YG_HIP(hipMemcpyAsync(F, dF, ms, hipMemcpyDeviceToHost, c->stream));
    if (J) YG_HIP(hipMemcpyAsync(J, dJ, ms, hipMemcpyDeviceToHost, c->stream));
    if (K) YG_HIP(hipMemcpyAsync(K, dK, ms, hipMemcpyDeviceToHost, c->stream));
    YG_HIP(hipStreamSynchronize(c->stream));
    return YGPU_OK;
} catch (...) {
    ygpu_set_error(YGPU_ERR_INVALID, "ygpu_build_fock_noalloc: exception");
    return YGPU_ERR_INVALID;
}
