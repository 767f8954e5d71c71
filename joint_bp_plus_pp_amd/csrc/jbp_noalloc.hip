// Allocation-free batched decode: jbp_decode_batch_noalloc.
//
// For callers that cannot allocate device memory at decode time (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// It is the FP32 decode of jbp_decode_batch (same kernels, same per-trial
// results) with inputs and outputs in the caller's device memory and every
// other buffer the kernels use in one file-scope device array, `scratch`:
//   [0, 256)            work counter of the persistent decoder
//   [256, 512)          channel table of the call (FP32 constants, FP64 prior)
//   [512, 512 + 4N)     decode-order keys (syndrome weights)
//   [512 + 4N, 512 + 8N) decode order (heaviest syndrome first)
// with N = kNoAllocMaxTrials.  Each GPU holds its own copy of the array until
// the process ends; its device address is looked up once per device when a
// context is created on it (noalloc_setup) and that address, never the
// array's host-side name, is what host code passes to kernels.
//
// Calls on one GPU share its array and keep the last call's contents; the
// counter is cleared by a fill kernel on every call.  Calls that overlap on
// one GPU (two streams, two threads) race on the array and must not be
// issued.
#include <algorithm>
#include <mutex>

#include "jbp_ctx.h"

namespace jbp {

constexpr int kNoAllocMaxTrials = 65536;
constexpr size_t kNoAllocHeader = 512;
constexpr size_t kNoAllocBytes = kNoAllocHeader + 2 * sizeof(int) * static_cast<size_t>(kNoAllocMaxTrials);
constexpr int kNoAllocMaxDevices = 64;

namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_mu;
void *g_scratch[kNoAllocMaxDevices] = {};

struct PriorArg {
    float p32[kPrior32];
    double p[4];
};

// Writes the call's channel constants into the scratch header (a kernel, not
// a copy, so nothing is read from host memory when the work runs).
__global__ void set_prior_kernel(float *p32, double *pd, PriorArg a) {
    const int i = threadIdx.x;
    if (i < kPrior32) p32[i] = a.p32[i];
    if (i < 4) pd[i] = a.p[i];
}

}  // namespace

int noalloc_setup(int device, char *err, size_t errlen) {
    if (device < 0 || device >= kNoAllocMaxDevices) {
        set_err(err, errlen, "device %d beyond the allocation-free table (%d devices)", device,
                kNoAllocMaxDevices);
        return JBP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return JBP_OK;
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) {
        set_err(err, errlen, "scratch lookup needs device %d current (have %d)", device, cur);
        return JBP_ERR_ARG;
    }
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
        set_err(err, errlen, "hipGetSymbolAddress returned no address for the scratch array");
        return JBP_ERR_DEVICE;
    }
    g_scratch[device] = ptr;
    return JBP_OK;
}

}  // namespace jbp

using namespace jbp;

extern "C" int jbp_decode_batch_noalloc(jbp_ctx *ctx, int count, const uint32_t *d_sx_bits,
                                        const uint32_t *d_sz_bits, const double prior[4],
                                        int max_iter, int *d_iterations, uint8_t *d_syn_match,
                                        uint32_t *d_est_x, uint32_t *d_est_z, void *stream,
                                        char *err, size_t errlen) {
    if (!ctx || count < 0 || !prior || max_iter < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (count > kNoAllocMaxTrials) {
        set_err(err, errlen, "%d trials exceed the allocation-free maximum of %d per call", count,
                kNoAllocMaxTrials);
        return JBP_ERR_ARG;
    }
    if (count == 0) return JBP_OK;
    if (!d_sx_bits || !d_sz_bits || !d_iterations || !d_syn_match || !d_est_x || !d_est_z) {
        set_err(err, errlen, "every input and output buffer is required");
        return JBP_ERR_ARG;
    }
    if (!ctx->qc) {
        set_err(err, errlen, "the allocation-free decode runs the LDS-resident FP32 decoder, which "
                             "needs a 768 x (3 x 12) permutation-block code");
        return JBP_ERR_UNSUPPORTED;
    }
    HIPCHK(hipSetDevice(ctx->device));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (ctx->device < kNoAllocMaxDevices) base = g_scratch[ctx->device];
    }
    if (!base) {
        set_err(err, errlen, "scratch of device %d was not set up", ctx->device);
        return JBP_ERR_DEVICE;
    }
    unsigned char *sp = static_cast<unsigned char *>(base);
    int *counter = reinterpret_cast<int *>(sp);
    float *p32 = reinterpret_cast<float *>(sp + 256);
    double *pd = reinterpret_cast<double *>(sp + 256 + 64);
    int *keys = reinterpret_cast<int *>(sp + kNoAllocHeader);
    int *order = keys + kNoAllocMaxTrials;
    hipStream_t st = static_cast<hipStream_t>(stream);

    PriorArg pa;
    const Prior pri = make_prior(prior);
    prior32_from(pri, pa.p32);
    for (int i = 0; i < 4; ++i) pa.p[i] = pri.p[i];
    hipLaunchKernelGGL(set_prior_kernel, dim3(1), dim3(64), 0, st, p32, pd, pa);
    NoiseTab nt;
    nt.prior32 = p32;
    nt.prior = pd;
    nt.n = 1;
    nt.phase = 0;

    TrialPlanes tp;
    tp.syn_x = const_cast<uint32_t *>(d_sx_bits);
    tp.syn_z = const_cast<uint32_t *>(d_sz_bits);
    tp.est_x = d_est_x;
    tp.est_z = d_est_z;
    tp.iters = d_iterations;
    tp.syn_ok = d_syn_match;
    launch_order_by_weight(ctx->dg, tp, count, keys, order, st);
    tp.order = order;
    launch_fill_i32(counter, 0, 1, st);
    launch_decode_qc_f32(ctx->dg, ctx->qcp, ctx->d_qc_tables, tp, count, nt, max_iter, 0, 1,
                         std::max(1, ctx->cus), counter, st);
    HIPCHK(hipGetLastError());
    return JBP_OK;
}
