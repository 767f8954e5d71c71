// Post-processing stage of a simulate chunk: belief propagation for the
// trials that are still unconverged when the reference's PP triggers first
// become possible, on the device, with the PP decisions and repairs
// (jbp_pp.cpp) on host threads.
//
// Why the split works: the reference only considers PP from iteration 100
// on (or at the last iteration, jointbp_ets.cpp:2563), so a trial that BP
// finishes in H = min(99, max_iter - 1) iterations never reaches it.  Those
// trials keep the ordinary decoders' results.  The rest are re-decoded here
// from the first iteration in FP64 with the reference's operation order, so
// their estimates, iteration counts and marginals are the reference's own:
//   first phase: iterations 0..H in one launch per batch of up to 1,024
//            trials, recording per iteration the unsatisfied-check counts and
//            whether the estimate equals the previous / the one before (all
//            the trigger logic reads), and iteration H's marginals, which go
//            to the host with the records;
//   later iterations (max_iter above 100): one launch each; the host reads
//            the records and both latest estimates, runs the trigger state
//            machine and, where it fires, ETS / flip / OSD (OSD's |LLR| come
//            from the marginals); trials that end at max_iter get the
//            post-decode pass.
// Batches alternate between two buffer sets and streams, so that the host
// repairs of one overlap the next one's launch.  Final estimates, iteration
// counts and syndrome flags go back into the chunk's planes, so the residual
// test and counters that follow treat these trials like any other.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "jbp_ctx.h"
#include "jbp_f64.h"

namespace jbp {
namespace {

using namespace f64;

// Per-slot state, in bytes from the slot's base: v2c[2] (4 doubles per
// message row), c2v[2] (2 per row), marginals (4 doubles per qubit), then
// the estimate and the previous estimate (one byte per qubit each).
struct StepSlot {
    double *v2c[2];
    double *c2v[2];
    double *tot;
    uint8_t *est;
    uint8_t *prev;
};

// Per-block output of a step launch, in bytes: the estimate and the previous
// estimate (one byte per qubit each), then the launch's last syndrome
// differences, syndrome(est) XOR the trial's syndrome, as bit words (X
// checks, then Z checks; bit c of word c / 32).
__host__ __device__ inline size_t step_out_bytes(const DevGraph &g) {
    return ((2 * static_cast<size_t>(g.n) + 3) & ~static_cast<size_t>(3)) + 4 * static_cast<size_t>(g.wm[0] + g.wm[1]);
}
__host__ __device__ inline size_t step_out_diff(const DevGraph &g) {
    return (2 * static_cast<size_t>(g.n) + 3) & ~static_cast<size_t>(3);
}

// Bits of one wavefront's 64 consecutive checks c - lane .. into words
// (lane 0 is active whenever any lane is: c grows with the lane).
__device__ __forceinline__ void put_diff_bits(uint32_t *w, int c, int bad, int wm) {
    const uint64_t bits = __ballot(bad);
    if ((threadIdx.x & 63) == 0) {
        w[c >> 5] = static_cast<uint32_t>(bits);
        if ((c >> 5) + 1 < wm) w[(c >> 5) + 1] = static_cast<uint32_t>(bits >> 32);
    }
}

__host__ __device__ inline size_t step_slot_bytes(int E0, int E1, int n) {
    size_t b = (6 * static_cast<size_t>(E0 + E1) + 4 * static_cast<size_t>(n)) * sizeof(double) +
               2 * static_cast<size_t>(n);
    return (b + 255) & ~static_cast<size_t>(255);
}

__device__ inline StepSlot step_slot(const DevGraph &g, char *state, size_t stride, int slot) {
    double *base = reinterpret_cast<double *>(state + static_cast<size_t>(slot) * stride);
    StepSlot s;
    s.v2c[0] = base;
    s.v2c[1] = s.v2c[0] + 4 * static_cast<size_t>(g.side[0].cap);
    s.c2v[0] = s.v2c[1] + 4 * static_cast<size_t>(g.side[1].cap);
    s.c2v[1] = s.c2v[0] + 2 * static_cast<size_t>(g.side[0].cap);
    s.tot = s.c2v[1] + 2 * static_cast<size_t>(g.side[1].cap);
    s.est = reinterpret_cast<uint8_t *>(s.tot + 4 * static_cast<size_t>(g.n));
    s.prev = s.est + g.n;
    return s;
}

// Record of one iteration: {unsat X, unsat Z, flags, 0}; flags bit 0 = the
// estimate equals the previous one, bit 1 = equals the one before that.
constexpr int kRecInts = 4;

// Block b advances the trial of slot list[b] by up to `steps` iterations
// (stopping when both syndromes match).  init: start from the reference's
// initial messages.  want_tot: 1 writes every iteration's marginals (the slot
// then holds the last iteration's), 2 only the launch's last iteration's (a
// trial that stops earlier leaves none).  At the end the block copies its
// estimate and previous estimate to out[b].
// Waves per SIMD the register budget allows: 4 for degree 12 / 3 with pair
// sums (128 VGPRs, every block of a 1024-trial batch resident at once), 2
// for the larger recursions and for the quasi-cyclic pair path, whose
// pipelined check pass holds two checks' messages and whose variable pass
// stages a column block's rows in LDS (72 KB: two blocks per compute unit).
template <int MAXC, int MAXV, bool PAIR, bool QC>
__global__ __launch_bounds__(256, (MAXC <= 12 && PAIR && !QC) ? 4 : 2) void pp_step_kernel(DevGraph g, TrialPlanes tp, NoiseTab nt, const int *list,
                               const int *slot_trial, int steps, int init, int want_tot,
                               double damping, char *state, size_t stride, int *rec, uint8_t *out) {
    __shared__ int s_cnt[2];
    __shared__ int s_qc[QC ? 6 * kQcMaxBlocks : 1];  // the block maps (QC codes)
    extern __shared__ double2 s_msg[];               // QC pair sums: a column block's rows
    const int tid = threadIdx.x, bs = blockDim.x;
    if (QC) {
        for (int i = tid; i < 6 * g.qc_J * g.qc_L; i += bs) s_qc[i] = g.qc_tab[i];
        __syncthreads();
    }
    const int b = blockIdx.x;
    const int slot = list[b];
    const int trial = slot_trial[slot];
    const StepSlot sp = step_slot(g, state, stride, slot);
    const int mx = g.side[0].m, mz = g.side[1].m, n = g.n;
    const double *pq = nt.prior + 4 * nt.entry(trial);
    const M4 prior{{pq[0], pq[1], pq[2], pq[3]}};
    const uint32_t *tsx = tp.syn_x + static_cast<size_t>(trial) * g.wm[0];
    const uint32_t *tsz = tp.syn_z + static_cast<size_t>(trial) * g.wm[1];
    if (init) {
        for (int s = 0; s < 2; ++s)
            for (int e = tid; e < g.side[s].cap; e += bs) {
                store_v2c<PAIR>(sp.v2c[s], static_cast<size_t>(e), s, prior);
                sp.c2v[s][2 * static_cast<size_t>(e)] = 0.25;
                sp.c2v[s][2 * static_cast<size_t>(e) + 1] = 0.25;
            }
        for (int v = tid; v < n; v += bs) sp.est[v] = sp.prev[v] = 0;
        __syncthreads();
    }
    uint8_t *o = out + static_cast<size_t>(b) * step_out_bytes(g);
    uint32_t *dw = reinterpret_cast<uint32_t *>(o + step_out_diff(g));
    int res = -1;  // QC pair path: the column block resident in LDS
    for (int j = 0; j < steps; ++j) {
        if (tid == 0) s_cnt[0] = s_cnt[1] = 0;
        // one loop per side: a single loop branching between the sides
        // keeps both inlined recursions' registers live
        if (QC && PAIR) {
            check_passes_qc<MAXC>(g, res, tsx, tsz, sp.v2c, sp.c2v, s_msg);
        } else {
#pragma unroll 1
            for (int c = tid; c < mx; c += bs)
                check_update<MAXC, PAIR, QC>(g.side[0], 0, c, (tsx[c >> 5] >> (c & 31)) & 1, sp.v2c[0], sp.c2v[0]);
#pragma unroll 1
            for (int c = tid; c < mz; c += bs)
                check_update<MAXC, PAIR, QC>(g.side[1], 1, c, (tsz[c >> 5] >> (c & 31)) & 1, sp.v2c[1], sp.c2v[1]);
        }
        __syncthreads();
        double *tot = (want_tot == 1 || (want_tot == 2 && j == steps - 1)) ? sp.tot : nullptr;
        int changed1 = 0, changed2 = 0;
        auto record = [&](int v, int best) {
            const int e1 = sp.est[v], e2 = sp.prev[v];
            changed1 |= best != e1;
            changed2 |= best != e2;
            sp.prev[v] = static_cast<uint8_t>(e1);
            sp.est[v] = static_cast<uint8_t>(best);
        };
        if (QC && PAIR) {
            var_pass_qc_lds<MAXV>(g, prior, sp.v2c, sp.c2v, s_qc, s_msg, res, [&](int v, int best, const M4 &total) {
                if (tot) {
                    double *t = tot + 4 * static_cast<size_t>(v);
                    t[0] = total.v[0]; t[1] = total.v[1]; t[2] = total.v[2]; t[3] = total.v[3];
                }
                record(v, best);
            });
        } else {
            for (int v = tid; v < n; v += bs)
                record(v, var_update<MAXV, PAIR, QC>(g, v, prior, damping, sp.v2c, sp.c2v, tot, s_qc));
        }
        __syncthreads();
        int ux = 0, uz = 0;
        for (int c = tid; c < mx; c += bs) {
            const int bad = check_parity<MAXC>(g.side[0], 0, c, sp.est) != static_cast<int>((tsx[c >> 5] >> (c & 31)) & 1u);
            put_diff_bits(dw, c, bad, g.wm[0]);
            ux += bad;
        }
        for (int c = tid; c < mz; c += bs) {
            const int bad = check_parity<MAXC>(g.side[1], 1, c, sp.est) != static_cast<int>((tsz[c >> 5] >> (c & 31)) & 1u);
            put_diff_bits(dw + g.wm[0], c, bad, g.wm[1]);
            uz += bad;
        }
        if (ux) atomicAdd(&s_cnt[0], ux);
        if (uz) atomicAdd(&s_cnt[1], uz);
        const int any1 = __syncthreads_or(changed1);
        const int any2 = __syncthreads_or(changed2);
        const int usx = s_cnt[0], usz = s_cnt[1];
        if (tid == 0) {
            int *r = rec + (static_cast<size_t>(b) * steps + j) * kRecInts;
            r[0] = usx;
            r[1] = usz;
            r[2] = (any1 ? 0 : 1) | (any2 ? 0 : 2);
            r[3] = 0;
        }
        __syncthreads();
        if (usx + usz == 0) break;
    }
    // the slot's messages complete in HBM for the next launch
    if (QC && PAIR) var_pass_flush(g, res, sp.v2c, s_msg);
    for (int v = tid; v < n; v += bs) {
        o[v] = sp.est[v];
        o[n + v] = sp.prev[v];
    }
}

void launch_step(const DevGraph &g, const TrialPlanes &tp, const NoiseTab &nt, const int *list,
                 int count, const int *slot_trial, int steps, bool init, int want_tot,
                 double damping, char *state, size_t stride, int *rec, uint8_t *out,
                 hipStream_t st) {
    const bool exact12 = g.max_cdeg <= 12 && g.max_vdeg <= 3;
    const bool small = g.max_cdeg <= 16 && g.max_vdeg <= 4, pair = damping <= 0.0;
    const bool qc = g.qc_tab && g.qc_J == 3 && g.qc_L == 12 && exact12;
    const size_t lds = f64::var_pass_lds_bytes(g.side[0].m);
    const bool qc_lds = qc && pair && lds <= kF64MaxLds;
    const int ini = init ? 1 : 0, wt = want_tot;
#define JBP_STEP(C, V, PR, Q)                                                                                    \
    hipLaunchKernelGGL((pp_step_kernel<C, V, PR, Q>), dim3(count), dim3(256), (PR && Q) ? lds : 0, st, g, tp, nt, \
                       list, slot_trial, steps, ini, wt, damping, state, stride, rec, out)
    if (qc_lds) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(pp_step_kernel<12, 3, true, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        JBP_STEP(12, 3, true, true);
    } else if (qc && !pair) JBP_STEP(12, 3, false, true);
    else if (exact12 && pair) JBP_STEP(12, 3, true, false);
    else if (exact12) JBP_STEP(12, 3, false, false);
    else if (small && pair) JBP_STEP(16, 4, true, false);
    else if (small) JBP_STEP(16, 4, false, false);
    else if (pair) JBP_STEP(32, 16, true, false);
    else JBP_STEP(32, 16, false, false);
#undef JBP_STEP
}

// dst[k][0..len) = src[idx[k] * stride + off + 0..len), in 8-byte units
// (the chunk syndromes of listed trials, the marginals of listed slots).
__global__ void gather_kernel(const uint64_t *src, size_t stride, size_t off, int len,
                              const int *idx, int count, uint64_t *dst) {
    const int k = blockIdx.y;
    if (k >= count) return;
    const uint64_t *s = src + static_cast<size_t>(idx[k]) * stride + off;
    uint64_t *d = dst + static_cast<size_t>(k) * len;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) d[i] = s[i];
}

// Writes the stage's outcome of listed trial k back into the chunk planes:
// est words (X then Z, wn each), iterations and syndrome flag.
__global__ void scatter_outcome_kernel(const uint32_t *planes, const int *iters, const uint8_t *ok,
                                       const int *idx, int count, int wn, TrialPlanes tp) {
    const int k = blockIdx.y;
    if (k >= count) return;
    const size_t t = static_cast<size_t>(idx[k]);
    const uint32_t *px = planes + static_cast<size_t>(k) * 2 * wn;
    const int stride = gridDim.x * blockDim.x;
    if ((wn & 3) == 0) {
        // 16 bytes per load: the planes may sit in host memory, where every
        // load is a bus round trip (rows 16-byte aligned: 8 wn bytes apart,
        // from aligned bases)
        const int w4 = wn >> 2;
        const uint4 *sx = reinterpret_cast<const uint4 *>(px), *sz = sx + w4;
        uint4 *dx = reinterpret_cast<uint4 *>(tp.est_x + t * wn), *dz = reinterpret_cast<uint4 *>(tp.est_z + t * wn);
        for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < w4; w += stride) {
            dx[w] = sx[w];
            dz[w] = sz[w];
        }
    } else {
        for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < wn; w += stride) {
            tp.est_x[t * wn + w] = px[w];
            tp.est_z[t * wn + w] = px[wn + w];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        tp.iters[t] = iters[k];
        tp.syn_ok[t] = ok[k];
    }
}

void launch_gather(const void *src, size_t stride, size_t off, int len, const int *idx, int count,
                   void *dst, hipStream_t st) {
    if (count <= 0 || len <= 0) return;
    const int bx = std::min(64, (len + 255) / 256);
    hipLaunchKernelGGL(gather_kernel, dim3(bx, count), dim3(256), 0, st,
                       static_cast<const uint64_t *>(src), stride, off, len, idx, count,
                       static_cast<uint64_t *>(dst));
}

// Host worker pool: run(count, fn) calls fn(0..count-1) across the workers
// and the calling thread and returns when all calls are done.
class Pool {
public:
    explicit Pool(int threads) {
        for (int i = 1; i < threads; ++i) th_.emplace_back([this] { work(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int count, const std::function<void(int)> &fn) {
        if (count <= 0) return;
        if (th_.empty() || count == 1) {
            for (int i = 0; i < count; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            count_ = count;
            next_.store(0);
            pending_ = static_cast<int>(th_.size());
            ++round_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

private:
    void drain() {
        for (int i; (i = next_.fetch_add(1)) < count_;) (*fn_)(i);
    }
    void work() {
        long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || round_ != seen; });
                if (stop_) return;
                seen = round_;
            }
            drain();
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    int count_ = 0;
    std::atomic<int> next_{0};
    int pending_ = 0;
    long round_ = 0;
    bool stop_ = false;
};

int pool_threads() {
    unsigned hc = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(hc ? hc : 1u, 16u)));
}

// Trials per batch of the stage (each holds about 3 MB of FP64 state for
// the included code, 22 MB for HGP(150x200)), and the state cap per batch.
constexpr int kPPSlots = 1024;
constexpr size_t kPPStateBytes = size_t(8) << 30;

void free_pp_buf(PPBuf &pb) {
    for (hipEvent_t &e : pb.tot_ev)
        if (e) (void)hipEventDestroy(e);
    if (pb.ready_ev) (void)hipEventDestroy(pb.ready_ev);
    for (void *p : {static_cast<void *>(pb.state), static_cast<void *>(pb.list),
                    static_cast<void *>(pb.slot_trial), static_cast<void *>(pb.rec),
                    static_cast<void *>(pb.out), static_cast<void *>(pb.idx), pb.stage})
        if (p) (void)hipFree(p);
    for (void *p : {static_cast<void *>(pb.h_rec), static_cast<void *>(pb.h_out),
                    static_cast<void *>(pb.h_tot), pb.h_stage, static_cast<void *>(pb.h_list)})
        if (p) (void)hipHostFree(p);
    pb = PPBuf();
}

int ensure_pp_buffers(jbp_ctx *ctx, PPBuf &pb, int slots, int rec_per_slot, char *err, size_t errlen) {
    const DevGraph &g = ctx->dg;
    const size_t stride = step_slot_bytes(g.side[0].cap, g.side[1].cap, g.n);
    if (pb.cap >= slots && pb.rec_cap >= rec_per_slot && pb.stride == stride) return JBP_OK;
    // Grow only (slots rounded up to 256, within the per-batch cap): a
    // fresh pinned buffer's first copies run at a fraction of the warm
    // rate, so calls of varying sizes keep the buffers they already have.
    const bool same = pb.stride == stride;
    const int batch_cap = static_cast<int>(std::max<size_t>(1, std::min<size_t>(kPPSlots, kPPStateBytes / stride)));
    slots = std::max(std::min((slots + 255) / 256 * 256, std::max(slots, batch_cap)), same ? pb.cap : 0);
    rec_per_slot = std::max(rec_per_slot, same ? pb.rec_cap : 0);
    free_pp_buf(pb);
    pb.cap = slots;
    pb.rec_cap = std::max(rec_per_slot, 1);
    pb.stride = stride;
    HIPCHK(hipMalloc(&pb.state, stride * slots));
    HIPCHK(hipMalloc(&pb.list, sizeof(int) * slots));
    HIPCHK(hipMalloc(&pb.slot_trial, sizeof(int) * slots));
    HIPCHK(hipMalloc(&pb.rec, sizeof(int) * kRecInts * static_cast<size_t>(slots) * pb.rec_cap));
    HIPCHK(hipMalloc(&pb.out, step_out_bytes(g) * slots));
    HIPCHK(hipMalloc(&pb.idx, sizeof(int) * slots));
    // staging: syndromes (as 8-byte units), marginals of every slot, or the
    // write-back planes + iterations + flags, whichever is largest
    const size_t syn = 8 * static_cast<size_t>(slots) * ((g.wm[0] + 1) / 2 + (g.wm[1] + 1) / 2);
    const size_t tot = 32 * static_cast<size_t>(slots) * g.n;
    const size_t back = static_cast<size_t>(slots) * (8 * static_cast<size_t>(g.wn) + 8);
    pb.stage_bytes = std::max(syn, std::max(tot, back));
    HIPCHK(hipMalloc(&pb.stage, pb.stage_bytes));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&pb.h_rec),
                         sizeof(int) * kRecInts * static_cast<size_t>(slots) * pb.rec_cap));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&pb.h_out), step_out_bytes(g) * slots));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&pb.h_tot), tot));
    HIPCHK(hipHostMalloc(&pb.h_stage, std::max(syn, back)));
    HIPCHK(hipHostGetDevicePointer(&pb.d_hstage, pb.h_stage, 0));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&pb.h_list), 3 * sizeof(int) * static_cast<size_t>(slots)));
    for (hipEvent_t &e : pb.tot_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&pb.ready_ev, hipEventDisableTiming));
    return JBP_OK;
}

// Copies list[0..count) to dev through region `which` (0..2) of the buffer
// set's pinned list array.  The caller must not refill that region before
// the stream has passed this copy (each region is reused only after a
// synchronisation that follows it).
int copy_list(PPBuf &pb, int which, const int *list, int count, int *dev, hipStream_t st,
              char *err, size_t errlen) {
    if (count <= 0) return JBP_OK;
    if (count > pb.cap || which < 0 || which > 2 || !pb.h_list) {
        set_err(err, errlen, "post-processing list of %d entries exceeds the buffer set (%d)", count,
                pb.cap);
        return JBP_ERR_ARG;
    }
    int *h = pb.h_list + static_cast<size_t>(which) * pb.cap;
    std::copy(list, list + count, h);
    HIPCHK(hipMemcpyAsync(dev, h, sizeof(int) * static_cast<size_t>(count), hipMemcpyHostToDevice, st));
    return JBP_OK;
}

void unpack_bits(const uint32_t *w, int count, uint8_t *out) {
    for (int i = 0; i < count; ++i) out[i] = static_cast<uint8_t>((w[i >> 5] >> (i & 31)) & 1u);
}

}  // namespace

void free_pp(jbp_ctx *ctx) {
    for (PPBuf &pb : ctx->ppb) free_pp_buf(pb);
    if (ctx->pp_idx) (void)hipFree(ctx->pp_idx);
    ctx->pp_idx = nullptr;
    ctx->pp_idx_cap = 0;
    if (ctx->pp_stream) (void)hipStreamDestroy(ctx->pp_stream);
    ctx->pp_stream = nullptr;
}

// One batch of the stage: its trials' host state and the buffer set and
// stream it uses.
struct PPBatch {
    size_t b0 = 0;
    int K = 0;
    const int *slot_trial = nullptr;  // device: trial index of each slot
    PPBuf *pb = nullptr;
    hipStream_t st = nullptr;
    int per_all = 1;        // rows per chunk of the first phase's marginal copies
    std::vector<int> live, live_row;  // slots running after iteration H; slot -> marginal row or -1
    bool syn_wait = false;  // syndromes queued for the host, not yet unpacked
    bool phase_a = false;   // phase A launched, results not yet read
    std::vector<uint8_t> sx, sz;
    std::vector<pp::Trial> tr;
    std::vector<std::vector<uint8_t>> est;
    std::vector<PPOutcome> o;
    std::vector<int> active;
    std::vector<pp::Attempt> att;
    std::vector<uint8_t> need_llr, at_end, attempted;
};

int run_pp(jbp_ctx *ctx, const NoiseTab &nt, int max_iter, double damping, int hist_window,
           const std::vector<int> &trials, std::vector<PPOutcome> &outs, char *err, size_t errlen) {
    const DevGraph &g = ctx->dg;
    const pp::Graph &hg = ctx->pp_graph;
    const pp::Tables &tabs = ctx->pp_tables;
    const int n = g.n, mx = g.side[0].m, mz = g.side[1].m;
    const size_t ostride = step_out_bytes(g), odiff = step_out_diff(g);  // step output per block
    const int H = std::min(99, max_iter - 1);  // iterations before any trigger can fire
    pp::Params prm;
    prm.max_iter = max_iter;
    prm.hist_window = hist_window;
    outs.assign(trials.size(), PPOutcome());
    if (trials.empty()) return JBP_OK;  // the usual case with long decodes: nothing to start
    if (!ctx->pp_stream) HIPCHK(hipStreamCreate(&ctx->pp_stream));
    Pool pool(pool_threads());
    const size_t stride = step_slot_bytes(g.side[0].cap, g.side[1].cap, n);
    const size_t batch = std::max<size_t>(1, std::min<size_t>(kPPSlots, kPPStateBytes / stride));
    const size_t nbatch = (trials.size() + batch - 1) / batch;
    hipStream_t streams[2] = {ctx->stream, ctx->pp_stream};
    // Every return, an error's included, leaves both streams drained: queued
    // copies and kernels use the context's buffers, which a later call may
    // grow or free.
    auto body = [&]() -> int {

    // The slot list 0..batch-1 and every trial index of the stage, uploaded
    // once: a batch's first launches then have no copy ahead of them in
    // their stream (a copy waits behind the previous batch's marginals on
    // the copy engine, and the launches would wait with it).
    {
        const size_t nidx = batch + trials.size();
        if (ctx->pp_idx_cap < nidx) {
            if (ctx->pp_idx) (void)hipFree(ctx->pp_idx);
            ctx->pp_idx = nullptr;
            ctx->pp_idx_cap = 0;
            HIPCHK(hipMalloc(&ctx->pp_idx, sizeof(int) * nidx));
            ctx->pp_idx_cap = nidx;
        }
        std::vector<int> idx(nidx);
        for (size_t i = 0; i < batch; ++i) idx[i] = static_cast<int>(i);
        std::copy(trials.begin(), trials.end(), idx.begin() + batch);
        HIPCHK(hipMemcpy(ctx->pp_idx, idx.data(), sizeof(int) * nidx, hipMemcpyHostToDevice));
    }
    const int *d_slots = ctx->pp_idx, *d_trials = ctx->pp_idx + batch;
    // where each side's syndrome words of a batch of K start in the staging
    // buffers (32-bit words; the Z side 8-byte aligned for the gather)
    auto syn_off = [&](int K, int s) -> size_t {
        return s == 0 ? 0 : ((static_cast<size_t>(K) * g.wm[0] + 1) & ~static_cast<size_t>(1));
    };
    // Iterations of the first device phase: the H iterations 0..H-1 before
    // any trigger can fire, plus iteration H, the first at which one can
    // (nothing the host decides before it changes it: a trial that
    // converges earlier stops on the device).
    const int HA = H + 1;
    const size_t row = 4 * static_cast<size_t>(n);
    const size_t tot_off = 6 * static_cast<size_t>(g.side[0].cap + g.side[1].cap);
    // Marginals of the listed slots gathered, then copied to the host in
    // chunks, one event each, so that the repairs of a chunk start as soon
    // as it has arrived.
    auto queue_marginals = [&](PPBuf &pb, hipStream_t st, const int *idx, int count, int &per) -> int {
        per = std::max(1, (count + PPBuf::kTotChunks - 1) / PPBuf::kTotChunks);
        if (count <= 0) return JBP_OK;
        launch_gather(pb.state, pb.stride / 8, tot_off, 4 * n, idx, count, pb.stage, st);
        HIPCHK(hipGetLastError());
        return JBP_OK;
    };
    auto queue_marginal_copies = [&](PPBuf &pb, hipStream_t st, int count, int per) -> int {
        for (int c = 0; c * per < count; ++c) {
            const size_t w0 = static_cast<size_t>(c) * per, w1 = std::min<size_t>(count, w0 + per);
            HIPCHK(hipMemcpyAsync(pb.h_tot + w0 * row, static_cast<const double *>(pb.stage) + w0 * row,
                                  sizeof(double) * (w1 - w0) * row, hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(pb.tot_ev[c], st));
        }
        return JBP_OK;
    };

    // Queues a batch without waiting: the syndromes' gather, the first
    // device phase (iterations 0..H), the copies of the syndromes, records
    // and estimates (ready event).  No copy goes ahead of the launches in the
    // stream (a copy waits behind the previous batch's marginals on the copy
    // engine, and the launches would wait with it).
    auto start = [&](PPBatch &B, size_t bi) -> int {
        B.b0 = bi * batch;
        B.K = static_cast<int>(std::min<size_t>(batch, trials.size() - B.b0));
        B.pb = &ctx->ppb[bi & 1];
        B.st = streams[bi & 1];
        const int K = B.K;
        // The batch that used this set before may still be writing back: its
        // copies and kernel precede everything queued here in the stream,
        // and the host writes nothing they read, so only growing the buffers
        // (freeing the old ones) waits for it.
        if (B.pb->cap < K || B.pb->rec_cap < HA || B.pb->stride != stride) HIPCHK(hipStreamSynchronize(B.st));
        int rc = ensure_pp_buffers(ctx, *B.pb, K, HA, err, errlen);
        if (rc) return rc;
        PPBuf &pb = *B.pb;
        hipStream_t st = B.st;
        B.slot_trial = d_trials + B.b0;
        for (int s = 0; s < 2; ++s)
            if (g.wm[s] % 2 == 0) {
                launch_gather(s == 0 ? ctx->tp.syn_x : ctx->tp.syn_z, g.wm[s] / 2, 0, g.wm[s] / 2,
                              B.slot_trial, K, static_cast<uint32_t *>(pb.stage) + syn_off(K, s), st);
                HIPCHK(hipGetLastError());
            }
        B.phase_a = max_iter > 0;
        if (B.phase_a) {
            launch_step(g, ctx->tp, nt, d_slots, K, B.slot_trial, HA, true, 2, damping, pb.state,
                        pb.stride, pb.rec, pb.out, st);
            HIPCHK(hipGetLastError());
        }
        for (int s = 0; s < 2; ++s) {
            const int wm = g.wm[s];
            uint32_t *w = static_cast<uint32_t *>(pb.h_stage) + syn_off(K, s);
            if (wm % 2 == 0) {
                HIPCHK(hipMemcpyAsync(w, static_cast<const uint32_t *>(pb.stage) + syn_off(K, s),
                                      4 * static_cast<size_t>(K) * wm, hipMemcpyDeviceToHost, st));
            } else {
                const uint32_t *src = s == 0 ? ctx->tp.syn_x : ctx->tp.syn_z;
                for (int i = 0; i < K; ++i)
                    HIPCHK(hipMemcpyAsync(w + static_cast<size_t>(i) * wm,
                                          src + static_cast<size_t>(trials[B.b0 + i]) * wm, 4 * wm,
                                          hipMemcpyDeviceToHost, st));
            }
        }
        B.syn_wait = true;
        B.per_all = 1;
        if (B.phase_a) {
            HIPCHK(hipMemcpyAsync(pb.h_rec, pb.rec, sizeof(int) * static_cast<size_t>(K) * HA * kRecInts,
                                  hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(pb.h_out, pb.out, static_cast<size_t>(K) * ostride, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipEventRecord(pb.ready_ev, st));
        return JBP_OK;
    };

    // The first phase's records and iteration H, further iterations one
    // launch each, the repairs and the write-back.  kick() starts the next
    // batch, whose first phase then holds every slot a step kernel could
    // use: it is called as soon as this batch has no step left, i.e. right
    // after the first phase when iteration H is the last, else after the
    // last step (with that iteration's marginals gathered first).
    auto finish_batch = [&](PPBatch &B, const std::function<int()> &kick) -> int {
        bool kicked = false;
        auto kick_once = [&]() -> int {
            if (kicked) return JBP_OK;
            kicked = true;
            return kick();
        };
        const int K = B.K;
        // host state (while the device runs phase A); vectors keep their
        // allocations from the batch that used this object before
        B.sx.resize(static_cast<size_t>(K) * mx);
        B.sz.resize(static_cast<size_t>(K) * mz);
        B.tr.assign(K, pp::Trial());
        B.est.resize(K);
        for (auto &e : B.est) e.assign(n, 0);
        B.o.assign(K, PPOutcome());
        B.active.resize(K);
        for (int i = 0; i < K; ++i) B.active[i] = i;
        B.att.resize(K);  // (every attempt resets its fields when it starts)
        B.need_llr.assign(K, 0);
        B.at_end.assign(K, 0);
        B.attempted.assign(K, 0);
        PPBuf &pb = *B.pb;
        hipStream_t st = B.st;
        auto &tr = B.tr;
        auto &est = B.est;
        auto &o = B.o;
        auto &active = B.active;
        auto &att = B.att;
        auto &need_llr = B.need_llr;
        auto &at_end = B.at_end;
        auto &attempted = B.attempted;
        const uint8_t *sx = B.sx.data(), *sz = B.sz.data();
        auto finish = [&](int i) {
            o[i].bp_ok = tr[i].bp_ok;
            o[i].ok = tr[i].ok;
            o[i].iterations = tr[i].iterations;
            o[i].pp_success = tr[i].pp_success;
            o[i].pp_ets = tr[i].pp_ets;
            o[i].pp_flip = tr[i].pp_flip;
            o[i].used_any = tr[i].used_any;
            o[i].used6 = tr[i].used6;
            o[i].used12 = tr[i].used12;
        };
        // Trial i ends unconverged at max_iter after an in-loop attempt at
        // that same iteration failed.  The reference's post-decode pass then
        // repeats exactly that attempt (ETS, the same flip sets, OSD with the
        // same |LLR|) on the decoder's estimate, in place: it fails the same
        // way and leaves the estimate the attempt built.  Take that result
        // instead of computing it twice.
        auto end_trial = [&](int i) {
            tr[i].done = true;
            tr[i].iterations = max_iter;
            finish(i);
            o[i].post_run = true;
            o[i].post = pp::PostOut();
            o[i].post.ets_applied = att[i].ets_used;
            o[i].post.used6 = att[i].ets6;
            o[i].post.used12 = att[i].ets12;
            o[i].ok = false;
            est[i].swap(att[i].trial);
        };
        int rc = JBP_OK;
        // the first phase, its copies and the marginals' gather (queued by start)
        HIPCHK(hipEventSynchronize(pb.ready_ev));
        if (B.phase_a) {
            // Iteration H's marginals of the slots still running after it
            // (the converged ones need none), gathered and their copies
            // queued before the next batch takes the device and the copy
            // engine.
            B.live.clear();
            B.live_row.assign(K, -1);
            for (int i = 0; i < K; ++i) {
                bool conv = false;
                for (int t = 0; t <= H && !conv; ++t) {
                    const int *r = &pb.h_rec[(static_cast<size_t>(i) * HA + t) * kRecInts];
                    conv = r[0] == 0 && r[1] == 0;
                }
                if (!conv) {
                    B.live_row[i] = static_cast<int>(B.live.size());
                    B.live.push_back(i);
                }
            }
            const int L = static_cast<int>(B.live.size());
            if (L > 0) {
                if ((rc = copy_list(pb, 0, B.live.data(), L, pb.idx, st, err, errlen))) return rc;
                if ((rc = queue_marginals(pb, st, pb.idx, L, B.per_all))) return rc;
                if ((rc = queue_marginal_copies(pb, st, L, B.per_all))) return rc;
            }
        }
        if (HA >= max_iter && (rc = kick_once())) return rc;
        if (B.syn_wait) {
            const uint32_t *w = static_cast<const uint32_t *>(pb.h_stage);
            const uint32_t *wz = w + syn_off(K, 1);
            pool.run(K, [&](int i) {
                unpack_bits(w + static_cast<size_t>(i) * g.wm[0], mx, B.sx.data() + static_cast<size_t>(i) * mx);
                unpack_bits(wz + static_cast<size_t>(i) * g.wm[1], mz, B.sz.data() + static_cast<size_t>(i) * mz);
            });
            B.syn_wait = false;
        }
        if (max_iter <= 0) {
            // no iteration at all: the post-decode pass on the zero estimate
            pool.run(K, [&](int i) {
                o[i].iterations = max_iter < 0 ? 0 : max_iter;
                o[i].post_run = true;
                o[i].post = pp::post_decode(hg, tabs, prm, tr[i], est[i], sx + static_cast<size_t>(i) * mx,
                                            sz + static_cast<size_t>(i) * mz, nullptr, nullptr);
                o[i].ok = o[i].post.ok;
            });
            active.clear();
        }
        int *rec = pb.h_rec;
        uint8_t *out = pb.h_out;
        if (B.phase_a) {
            // iterations before H: bookkeeping only (no trigger can fire)
            pool.run(K, [&](int i) {
                for (int t = 0; t < H; ++t) {
                    const int *r = &rec[(static_cast<size_t>(i) * HA + t) * kRecInts];
                    pp::IterRec ir;
                    ir.us[0] = r[0];
                    ir.us[1] = r[1];
                    ir.syn_all = r[0] == 0 && r[1] == 0;
                    ir.same = r[2] & 1;
                    ir.osc = (r[2] >> 1) & 1;
                    if (pp::observe(hg, prm, tr[i], t, ir, nullptr, nullptr) == pp::Step::kConverged) {
                        std::copy(&out[static_cast<size_t>(i) * ostride], &out[static_cast<size_t>(i) * ostride] + n,
                                  est[i].begin());
                        finish(i);
                        break;
                    }
                }
            });
            active.erase(std::remove_if(active.begin(), active.end(), [&](int i) { return tr[i].done; }),
                         active.end());
        }
        double *tot = pb.h_tot;
        std::vector<int> pos(K);  // trial -> row of its marginals (a step's last iteration)
        for (int t = H; B.phase_a && t < max_iter && !active.empty(); ++t) {
            const int A = static_cast<int>(active.size());
            const bool first = t == H;  // iteration H: computed by the first phase
            const bool last = t + 1 == max_iter;
            int per = B.per_all;
            if (!first) {
                if ((rc = copy_list(pb, 1, active.data(), A, pb.list, st, err, errlen))) return rc;
                launch_step(g, ctx->tp, nt, pb.list, A, B.slot_trial, 1, false, 1, damping, pb.state,
                            pb.stride, pb.rec, pb.out, st);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(rec, pb.rec, sizeof(int) * static_cast<size_t>(A) * kRecInts,
                                      hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(out, pb.out, static_cast<size_t>(A) * ostride, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                // In the last iteration every active trial's marginals are
                // gathered now, before the next batch takes the device (most
                // of them are needed); before that only those of the trials
                // that reach OSD, below.
                if (last) {
                    if ((rc = queue_marginals(pb, st, pb.list, A, per))) return rc;
                    if ((rc = kick_once())) return rc;
                    if ((rc = queue_marginal_copies(pb, st, A, per))) return rc;
                    for (int a = 0; a < A; ++a) pos[active[a]] = a;
                }
            }
            // iteration t's records and estimates: by slot after the first
            // phase, by position in the list after a step
            auto rec_of = [&](int a) -> const int * {
                return first ? &rec[(static_cast<size_t>(active[a]) * HA + H) * kRecInts]
                             : &rec[static_cast<size_t>(a) * kRecInts];
            };
            auto out_of = [&](int a) -> const uint8_t * {
                return &out[static_cast<size_t>(first ? active[a] : a) * ostride];
            };
            pool.run(A, [&](int a) {
                const int i = active[a];
                const int *r = rec_of(a);
                const uint8_t *e_now = out_of(a);
                const uint8_t *e_prev = e_now + n;
                pp::IterRec ir;
                ir.us[0] = r[0];
                ir.us[1] = r[1];
                ir.syn_all = r[0] == 0 && r[1] == 0;
                ir.same = r[2] & 1;
                ir.osc = (r[2] >> 1) & 1;
                need_llr[i] = at_end[i] = attempted[i] = 0;
                att[i].want_osd = false;
                std::copy(e_now, e_now + n, est[i].begin());
                const pp::Step step = pp::observe(hg, prm, tr[i], t, ir, e_now, t > 0 ? e_prev : nullptr);
                if (step == pp::Step::kConverged) {
                    finish(i);
                    return;
                }
                const uint8_t *tsx = sx + static_cast<size_t>(i) * mx, *tsz = sz + static_cast<size_t>(i) * mz;
                // the device's syndrome differences of this estimate
                const uint32_t *dbits[2] = {reinterpret_cast<const uint32_t *>(e_now + odiff),
                                            reinterpret_cast<const uint32_t *>(e_now + odiff) + g.wm[0]};
                if (step == pp::Step::kAttempt) {
                    attempted[i] = 1;
                    pp::attempt_begin(hg, tabs, prm, tr[i], est[i], tsx, tsz, att[i], dbits);
                    if (pp::attempt_end(tr[i], att[i], t, est[i])) {
                        finish(i);
                        return;
                    }
                    need_llr[i] = att[i].want_osd;
                }
                if (last) {
                    at_end[i] = 1;
                    if (!attempted[i]) {
                        // the post-decode pass runs itself (no attempt to reuse):
                        // ETS and flip sets now, |LLR| only if OSD follows
                        tr[i].iterations = max_iter;
                        o[i].post_run = true;
                        o[i].post = pp::post_begin(hg, tabs, prm, tr[i], est[i], tsx, tsz, att[i].diff, dbits);
                        need_llr[i] = o[i].post.want_osd;
                        if (!need_llr[i]) {
                            tr[i].done = true;
                            finish(i);
                            o[i].ok = o[i].post.ok;
                            return;
                        }
                    }
                }
                if (at_end[i] && !need_llr[i]) end_trial(i);
            });
            // marginals of the trials that need |LLR| (OSD now or after the decode)
            std::vector<int> want;
            for (int i : active)
                if (!tr[i].done && need_llr[i]) want.push_back(i);
            const int W = static_cast<int>(want.size());
            if (!first && !last && W > 0) {
                if ((rc = copy_list(pb, 2, want.data(), W, pb.idx, st, err, errlen))) return rc;
                if ((rc = queue_marginals(pb, st, pb.idx, W, per))) return rc;
                if ((rc = queue_marginal_copies(pb, st, W, per))) return rc;
            }
            std::atomic<int> ev_err{0};
            pool.run(W, [&](int w) {
                const int i = want[w];
                // row of the trial's marginals
                const int r = first ? B.live_row[i] : last ? pos[i] : w;
                if (r < 0) {  // (a trial that needs them is running: never)
                    ev_err.store(2);
                    return;
                }
                if (hipEventSynchronize(pb.tot_ev[r / per]) != hipSuccess) {
                    ev_err.store(1);
                    return;
                }
                // OSD takes the marginals and computes the |LLR| it needs
                const pp::LlrSide lx = pp::LlrSide::marginals(tot + static_cast<size_t>(r) * row), lz = lx;
                if (att[i].want_osd) {
                    pp::attempt_osd(hg, att[i], lx, lz);
                    if (pp::attempt_end(tr[i], att[i], t, est[i])) {
                        finish(i);
                        return;
                    }
                }
                if (at_end[i]) {
                    if (attempted[i]) {
                        end_trial(i);
                    } else {
                        tr[i].done = true;
                        finish(i);
                        pp::post_osd(hg, o[i].post, est[i], att[i].diff, lx, lz);
                        o[i].ok = o[i].post.ok;
                    }
                }
            });
            if (ev_err.load()) {
                set_err(err, errlen, ev_err.load() == 2 ? "marginals missing for a trial that needs them"
                                                        : "waiting for the marginals' copy to the host failed");
                return JBP_ERR_DEVICE;
            }
            active.erase(std::remove_if(active.begin(), active.end(), [&](int i) { return tr[i].done; }),
                         active.end());
        }
        if ((rc = kick_once())) return rc;
        // write back: estimate planes, iterations, syndrome flag.  Queued
        // without a wait: the host reads nothing it writes before the final
        // synchronisation, and this buffer set's next use is ordered after
        // it on the same stream.  The kernel reads the pinned buffer
        // directly: a copy would wait behind the marginal copies on the
        // copy engine, and the next batch queued behind it with it.
        const size_t nplane = static_cast<size_t>(K) * 2 * g.wn;
        uint32_t *planes = static_cast<uint32_t *>(pb.h_stage);
        int *it = reinterpret_cast<int *>(planes + nplane);
        uint8_t *okb = reinterpret_cast<uint8_t *>(it + K);
        pool.run(K, [&](int i) {
            uint32_t *px = planes + static_cast<size_t>(i) * 2 * g.wn, *pz = px + g.wn;
            std::fill(px, px + 2 * g.wn, 0u);
            const uint8_t *e = est[i].data();
            int v = 0;
            for (; v + 8 <= n; v += 8) {
                // bit 0 (X) and bit 1 (Z) of eight states, gathered into a byte each
                constexpr uint64_t kLow = 0x0101010101010101ull, kGather = 0x0102040810204080ull;
                uint64_t x;
                std::memcpy(&x, e + v, 8);
                px[v >> 5] |= static_cast<uint32_t>(((x & kLow) * kGather) >> 56) << (v & 31);
                pz[v >> 5] |= static_cast<uint32_t>((((x >> 1) & kLow) * kGather) >> 56) << (v & 31);
            }
            for (; v < n; ++v) {
                px[v >> 5] |= static_cast<uint32_t>(e[v] & 1) << (v & 31);
                pz[v >> 5] |= static_cast<uint32_t>((e[v] >> 1) & 1) << (v & 31);
            }
            it[i] = o[i].iterations;
            okb[i] = o[i].ok ? 1 : 0;
        });
        for (int i = 0; i < K; ++i) outs[B.b0 + i] = o[i];
        const uint32_t *d_planes = static_cast<const uint32_t *>(pb.d_hstage);
        const int *d_it = reinterpret_cast<const int *>(d_planes + nplane);
        const uint8_t *d_ok = reinterpret_cast<const uint8_t *>(d_it + K);
        const int bx = std::min(16, (g.wn + 255) / 256);
        hipLaunchKernelGGL(scatter_outcome_kernel, dim3(bx, K), dim3(256), 0, st, d_planes, d_it, d_ok,
                           B.slot_trial, K, g.wn, ctx->tp);
        HIPCHK(hipGetLastError());
        return JBP_OK;
    };

    // Batches alternate between the two buffer sets and streams; batch
    // b + 1 starts from inside batch b's finish (kick).
    PPBatch cur, next;
    int rc = start(cur, 0);
    if (rc) return rc;
    for (size_t bi = 0; bi < nbatch; ++bi) {
        auto kick = [&]() -> int { return bi + 1 < nbatch ? start(next, bi + 1) : JBP_OK; };
        if ((rc = finish_batch(cur, kick))) return rc;
        std::swap(cur, next);
    }
    HIPCHK(hipStreamSynchronize(streams[0]));
    HIPCHK(hipStreamSynchronize(streams[1]));
    return JBP_OK;
    };
    const int status = body();
    if (status != JBP_OK) {
        (void)hipStreamSynchronize(streams[0]);
        (void)hipStreamSynchronize(streams[1]);
    }
    return status;
}

}  // namespace jbp

extern "C" int jbp_set_pp_tables(jbp_ctx *ctx, const jbp_pp_tables *t, char *err, size_t errlen) {
    if (!ctx) {
        jbp::set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    ctx->pp_tables = t ? t->t : jbp::pp::Tables();
    return JBP_OK;
}

extern "C" int jbp_decode_batch_pp(jbp_ctx *ctx, int count, const uint32_t *sx_bits,
                                   const uint32_t *sz_bits, const double prior[4], int max_iter,
                                   double damping, int precision, int flip_hist,
                                   int *iterations_out, uint8_t *syn_match_out, uint8_t *pp_out,
                                   int *est_out, char *err, size_t errlen) {
    using namespace jbp;
    if (!ctx || count < 0 || !sx_bits || !sz_bits || !prior || max_iter < 0 || flip_hist < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (count == 0) return JBP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, count, err, errlen);
    if (rc) return rc;
    const DevGraph &g = ctx->dg;
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_x, sx_bits, static_cast<size_t>(count) * g.wm[0] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_z, sz_bits, static_cast<size_t>(count) * g.wm[1] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    {
        const Prior pri = make_prior(prior);
        if ((rc = set_noise(ctx, &pri, nullptr, 1, err, errlen))) return rc;
    }
    const int h = std::min(99, max_iter - 1);
    std::vector<uint8_t> ok(count, 0);
    if (h > 0) {
        rc = run_decode(ctx, count, ctx->noise, h, 0, 1, damping, precision, err, errlen);
        if (rc) return rc;
        HIPCHK(hipMemcpy(ok.data(), ctx->tp.syn_ok, count, hipMemcpyDeviceToHost));
    }
    std::vector<int> list;
    for (int i = 0; i < count; ++i)
        if (h <= 0 || !ok[i]) list.push_back(i);
    std::vector<PPOutcome> outs;
    rc = run_pp(ctx, ctx->noise, max_iter, damping, flip_hist, list, outs, err, errlen);
    if (rc) return rc;
    if (pp_out) {
        for (int i = 0; i < count; ++i) pp_out[i] = JBP_PPF_BP_OK;
        for (size_t k = 0; k < list.size(); ++k) {
            const PPOutcome &o = outs[k];
            const bool ets = o.pp_ets || (o.post_run && o.post.ets_ok);
            const bool flip = o.pp_flip || (o.post_run && o.post.flip_ok);
            const bool osd = (o.post_run && o.post.osd_ok) || (o.pp_success && !o.pp_ets && !o.pp_flip);
            const bool any = o.pp_success || ets || flip || osd;
            const bool u6 = o.used6 || (o.post_run && o.post.used6);
            const bool u12 = o.used12 || (o.post_run && o.post.used12);
            const bool used = o.used_any || u6 || u12 || (o.post_run && o.post.ets_applied);
            pp_out[list[k]] = static_cast<uint8_t>(
                (o.bp_ok ? JBP_PPF_BP_OK : 0) | (any && o.ok ? JBP_PPF_PP_SUCCESS : 0) |
                (ets && o.ok ? JBP_PPF_ETS_SAVED : 0) | (flip && o.ok ? JBP_PPF_FLIP_SAVED : 0) |
                (osd && o.ok ? JBP_PPF_OSD_SAVED : 0) | (used ? JBP_PPF_ETS_USED : 0) |
                (u6 ? JBP_PPF_ETS6 : 0) | (u12 ? JBP_PPF_ETS12 : 0));
        }
    }
    if (iterations_out)
        HIPCHK(hipMemcpy(iterations_out, ctx->tp.iters, static_cast<size_t>(count) * 4,
                         hipMemcpyDeviceToHost));
    if (syn_match_out)
        HIPCHK(hipMemcpy(syn_match_out, ctx->tp.syn_ok, static_cast<size_t>(count),
                         hipMemcpyDeviceToHost));
    if (est_out) {
        if ((rc = ensure_states(ctx, static_cast<size_t>(count) * g.n, err, errlen))) return rc;
        launch_unpack_states(ctx->tp.est_x, ctx->tp.est_z, g.wn, g.n, count, ctx->d_states,
                             ctx->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(est_out, ctx->d_states, static_cast<size_t>(count) * g.n * 4,
                              hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return JBP_OK;
}

extern "C" int jbp_decode_llr(jbp_ctx *ctx, int count, const uint32_t *sx_bits, const uint32_t *sz_bits,
                              const double prior[4], int max_iter, double damping,
                              int *iterations_out, uint8_t *syn_match_out, int *est_out,
                              double *llr_out, char *err, size_t errlen) {
    using namespace jbp;
    if (!ctx || count < 0 || !sx_bits || !sz_bits || !prior || max_iter < 1 || damping < 0.0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (count == 0) return JBP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, count, err, errlen);
    if (rc) return rc;
    const DevGraph &g = ctx->dg;
    const int n = g.n;
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_x, sx_bits, static_cast<size_t>(count) * g.wm[0] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_z, sz_bits, static_cast<size_t>(count) * g.wm[1] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    {
        const Prior pri = make_prior(prior);
        if ((rc = set_noise(ctx, &pri, nullptr, 1, err, errlen))) return rc;
    }
    hipStream_t st = ctx->stream;
    const size_t batch = std::max<size_t>(
        1, std::min<size_t>(kPPSlots, kPPStateBytes / step_slot_bytes(g.side[0].cap, g.side[1].cap, n)));
    for (int b0 = 0; b0 < count; b0 += static_cast<int>(batch)) {
        const int K = static_cast<int>(std::min<size_t>(batch, static_cast<size_t>(count - b0)));
        if ((rc = ensure_pp_buffers(ctx, ctx->ppb[0], K, max_iter, err, errlen))) return rc;
        PPBuf &pb = ctx->ppb[0];
        std::vector<int> ids(K);
        for (int i = 0; i < K; ++i) ids[i] = i;
        std::vector<int> trials(K);
        for (int i = 0; i < K; ++i) trials[i] = b0 + i;
        if ((rc = copy_list(pb, 0, ids.data(), K, pb.list, st, err, errlen))) return rc;
        if ((rc = copy_list(pb, 1, trials.data(), K, pb.slot_trial, st, err, errlen))) return rc;
        // every iteration in one launch, marginals kept (the last one stays)
        launch_step(g, ctx->tp, ctx->noise, pb.list, K, pb.slot_trial, max_iter, true, 1, damping,
                    pb.state, pb.stride, pb.rec, pb.out, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(pb.h_rec, pb.rec, sizeof(int) * kRecInts * static_cast<size_t>(K) * max_iter,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(pb.h_out, pb.out, static_cast<size_t>(K) * step_out_bytes(g), hipMemcpyDeviceToHost, st));
        launch_gather(pb.state, pb.stride / 8, 6 * static_cast<size_t>(g.side[0].cap + g.side[1].cap),
                      4 * n, pb.list, K, pb.stage, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(pb.h_tot, pb.stage, sizeof(double) * static_cast<size_t>(K) * 4 * n,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < K; ++i) {
            int iters = max_iter, match = 0;
            for (int t = 0; t < max_iter; ++t) {
                const int *r = pb.h_rec + (static_cast<size_t>(i) * max_iter + t) * kRecInts;
                if (r[0] == 0 && r[1] == 0) {
                    iters = t + 1;
                    match = 1;
                    break;
                }
            }
            const size_t k = static_cast<size_t>(b0 + i);
            if (iterations_out) iterations_out[k] = iters;
            if (syn_match_out) syn_match_out[k] = static_cast<uint8_t>(match);
            if (est_out)
                for (int v = 0; v < n; ++v) est_out[k * n + v] = pb.h_out[static_cast<size_t>(i) * step_out_bytes(g) + v];
            if (llr_out)
                pp::abs_llr_from_marginals(n, pb.h_tot + static_cast<size_t>(i) * 4 * n, llr_out + k * 2 * n,
                                           llr_out + k * 2 * n + n);
        }
    }
    return JBP_OK;
}
