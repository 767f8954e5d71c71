// FP64 parity decoder: the reference's joint BP (jointbp_ets.cpp:1943-2535,
// 2733-2735, 3078) in the reference's exact operation order, so that the
// per-trial estimate and iteration count are identical to its CPU decoder.
// This translation unit is compiled with -ffp-contract=off (no FMA), and
// normalisation divides by the left-to-right sum as normalize_msg does
// (jointbp_ets.cpp:650-657).
//
// Layout: one workgroup decodes one trial at a time and pulls the next trial
// index from a device counter when it finishes (persistent, refill on
// completion).  Its messages live in a private slice of a global scratch
// buffer, position-major (row i*m + c: position i of check c, so a check
// pass over consecutive checks reads and writes contiguous rows):
//   v2c[side][rows][2 or 4]  variable-to-check: the two pair sums the check
//                            update reads, or all four Pauli components
//                            when damping mixes old and new messages
//   c2v[side][rows][2]       check-to-variable; the reference's X-side
//                            message is {a, b, a, b} and Z-side
//                            {a, a, b, b}, so two values carry it
//   est[n]                   hard decision, one byte per qubit
// Any code is served through its CSR tables; quasi-cyclic codes with pair
// sums compute their rows from the block maps instead and stage the
// variable pass through LDS one column block at a time.
#include "jbp_common.h"
#include "jbp_f64.h"

namespace jbp {
namespace {

using namespace f64;

struct SlotPtrs {
    double *v2c[2];
    double *c2v[2];
    uint8_t *est;
};

__device__ __forceinline__ SlotPtrs slot_ptrs(const DevGraph &g, double *scratch, size_t stride) {
    double *base = scratch + static_cast<size_t>(blockIdx.x) * stride;
    SlotPtrs s;
    s.v2c[0] = base;
    s.v2c[1] = s.v2c[0] + 4 * static_cast<size_t>(g.side[0].cap);
    s.c2v[0] = s.v2c[1] + 4 * static_cast<size_t>(g.side[1].cap);
    s.c2v[1] = s.c2v[0] + 2 * static_cast<size_t>(g.side[0].cap);
    s.est = reinterpret_cast<uint8_t *>(s.c2v[1] + 2 * static_cast<size_t>(g.side[1].cap));
    return s;
}

// Blocks of up to 512 threads (the tuner's largest); the second bound is
// waves per SIMD (2: up to 256 VGPRs).  Degree 12 / 3 with pair sums needs
// about 143, so three 256-thread blocks share a compute unit; the
// quasi-cyclic pair path (pipelined check pass, 72 KB of LDS) runs two.
// FREEZE: the reference's --freeze-syn (jointbp_ets.cpp:1884-1901,
// 2526-2535): once a side's syndrome matches, its check pass stops and its
// check-to-variable messages become the deterministic messages of the
// estimate's bits at that iteration (generic rows only, QC = false).
template <int MAXC, int MAXV, bool PAIR, bool QC, bool FREEZE = false>
__global__ __launch_bounds__(512, 2) void decode_f64_kernel(DevGraph g, TrialPlanes tp, int count, NoiseTab nt, int max_iter,
                                  double damping, double *scratch, size_t stride, int *counter) {
    static_assert(!(FREEZE && QC), "freeze runs on the generic rows");
    __shared__ int s_trial;
    __shared__ int s_qc[QC ? 6 * kQcMaxBlocks : 1];  // the block maps (QC codes)
    extern __shared__ double2 s_msg[];               // QC pair sums: a column block's rows
    const int tid = threadIdx.x;
    const int bs = blockDim.x;
    if (QC) {
        for (int i = tid; i < 6 * g.qc_J * g.qc_L; i += bs) s_qc[i] = g.qc_tab[i];
        __syncthreads();
    }
    const SlotPtrs sp = slot_ptrs(g, scratch, stride);
    const int mx = g.side[0].m, mz = g.side[1].m;
    for (;;) {
        if (tid == 0) {
            const int k = atomicAdd(counter, 1);
            s_trial = (tp.order && k < count) ? tp.order[k] : k;
        }
        __syncthreads();
        const int trial = s_trial;
        __syncthreads();
        if (trial >= count) return;
        const double *pq = nt.prior + 4 * nt.entry(trial);
        const M4 prior{{pq[0], pq[1], pq[2], pq[3]}};
        const uint32_t *tsx = tp.syn_x + static_cast<size_t>(trial) * g.wm[0];
        const uint32_t *tsz = tp.syn_z + static_cast<size_t>(trial) * g.wm[1];

        // init (jointbp_ets.cpp:1970-1982): v2c = prior, c2v = uniform
        for (int s = 0; s < 2; ++s) {
            const int E = g.side[s].cap;
            for (int e = tid; e < E; e += bs) {
                store_v2c<PAIR>(sp.v2c[s], static_cast<size_t>(e), s, prior);
                sp.c2v[s][2 * static_cast<size_t>(e)] = 0.25;
                sp.c2v[s][2 * static_cast<size_t>(e) + 1] = 0.25;
            }
        }
        __syncthreads();

        int iterations = max_iter;
        int matched = 0;
        int res = -1;  // QC pair path: the column block resident in LDS
        bool frozen[2] = {false, false};
        for (int iter = 0; iter < max_iter; ++iter) {
            // one loop per side: a single loop branching between the sides
            // keeps both inlined recursions' registers live
            if (QC && PAIR) {
                check_passes_qc<MAXC>(g, res, tsx, tsz, sp.v2c, sp.c2v, s_msg);
            } else {
#pragma unroll 1
                for (int c = tid; c < mx && !frozen[0]; c += bs)
                    check_update<MAXC, PAIR, QC>(g.side[0], 0, c, (tsx[c >> 5] >> (c & 31)) & 1, sp.v2c[0], sp.c2v[0]);
#pragma unroll 1
                for (int c = tid; c < mz && !frozen[1]; c += bs)
                    check_update<MAXC, PAIR, QC>(g.side[1], 1, c, (tsz[c >> 5] >> (c & 31)) & 1, sp.v2c[1], sp.c2v[1]);
            }
            __syncthreads();
            if (QC && PAIR) {
                var_pass_qc_lds<MAXV>(g, prior, sp.v2c, sp.c2v, s_qc, s_msg, res,
                                      [&](int v, int best, const M4 &) { sp.est[v] = static_cast<uint8_t>(best); });
            } else {
                for (int v = tid; v < g.n; v += bs)
                    sp.est[v] = static_cast<uint8_t>(
                        var_update<MAXV, PAIR, QC>(g, v, prior, damping, sp.v2c, sp.c2v, nullptr, s_qc));
            }
            __syncthreads();
            // both syndromes of est (jointbp_ets.cpp:2521-2525)
            if (FREEZE) {
                int badx = 0, badz = 0;
                for (int c = tid; c < mx && !badx; c += bs)
                    badx |= check_parity<MAXC>(g.side[0], 0, c, sp.est) != static_cast<int>((tsx[c >> 5] >> (c & 31)) & 1u);
                for (int c = tid; c < mz && !badz; c += bs)
                    badz |= check_parity<MAXC>(g.side[1], 1, c, sp.est) != static_cast<int>((tsz[c >> 5] >> (c & 31)) & 1u);
                const bool okx = !__syncthreads_or(badx), okz = !__syncthreads_or(badz);
                // freeze_side_msgs: every check-to-variable message of a side
                // whose syndrome now matches becomes the estimate bit's
                // deterministic message ({.5,0,.5,0} / {0,.5,0,.5} on the X
                // side, {.5,.5,0,0} / {0,0,.5,.5} on the Z side)
                for (int s = 0; s < 2; ++s) {
                    if (!(s == 0 ? okx : okz) || frozen[s]) continue;
                    frozen[s] = true;
                    const DevSide &sd = g.side[s];
                    const size_t m = static_cast<size_t>(sd.m);
                    for (int c = tid; c < sd.m; c += bs) {
                        const int b = sd.chk_off[c], deg = sd.chk_off[c + 1] - b;
                        for (int i = 0; i < deg; ++i) {
                            const int st = sp.est[sd.edge_var[b + i]];
                            const int bit = s == 0 ? (st & 1) : (st >> 1);
                            double2 o;
                            o.x = bit ? 0.0 : 0.5;
                            o.y = bit ? 0.5 : 0.0;
                            *reinterpret_cast<double2 *>(sp.c2v[s] + 2 * i * m + 2 * c) = o;
                        }
                    }
                }
                __syncthreads();
                if (okx && okz) {
                    iterations = iter + 1;
                    matched = 1;
                    break;
                }
                continue;
            }
            int bad = 0;
            for (int c = tid; c < mx && !bad; c += bs)
                bad |= check_parity<MAXC>(g.side[0], 0, c, sp.est) != static_cast<int>((tsx[c >> 5] >> (c & 31)) & 1u);
            for (int c = tid; c < mz && !bad; c += bs)
                bad |= check_parity<MAXC>(g.side[1], 1, c, sp.est) != static_cast<int>((tsz[c >> 5] >> (c & 31)) & 1u);
            if (!__syncthreads_or(bad)) {
                iterations = iter + 1;
                matched = 1;
                break;
            }
        }
        // est bit planes, one ballot per 64 qubits
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int nwaves = (bs + 63) >> 6;
        uint32_t *ox = tp.est_x + static_cast<size_t>(trial) * g.wn;
        uint32_t *oz = tp.est_z + static_cast<size_t>(trial) * g.wn;
        for (int v0 = wave * 64; v0 < g.n; v0 += nwaves * 64) {
            int v = v0 + lane;
            int st = v < g.n ? sp.est[v] : 0;
            uint64_t bx = __ballot(st & 1);
            uint64_t bz = __ballot(st >> 1);
            if (lane == 0) {
                ox[v0 >> 5] = static_cast<uint32_t>(bx);
                oz[v0 >> 5] = static_cast<uint32_t>(bz);
                if ((v0 >> 5) + 1 < g.wn) {
                    ox[(v0 >> 5) + 1] = static_cast<uint32_t>(bx >> 32);
                    oz[(v0 >> 5) + 1] = static_cast<uint32_t>(bz >> 32);
                }
            }
        }
        if (tid == 0) {
            tp.iters[trial] = iterations;
            tp.syn_ok[trial] = static_cast<uint8_t>(matched);
        }
        __syncthreads();
    }
}

}  // namespace

size_t decode_f64_scratch_bytes(const DevGraph &g, int slots) {
    size_t per = (6 * static_cast<size_t>(g.side[0].cap + g.side[1].cap)) * sizeof(double) +
                 ((static_cast<size_t>(g.n) + 7) / 8) * sizeof(double);
    per = (per + 255) / 256 * 256;  // every slot 256-byte aligned (16-byte message pairs)
    return per * static_cast<size_t>(slots);
}

void launch_decode_f64(const DevGraph &g, TrialPlanes tp, int count, const NoiseTab &nt, int max_iter,
                       double damping, int slots, int block, double *scratch, int *counter,
                       hipStream_t st, bool freeze) {
    if (count <= 0) return;
    size_t stride = decode_f64_scratch_bytes(g, 1) / sizeof(double);
    int grid = slots < count ? slots : count;
    // recursion arrays sized to the code (registers): degree 12 / 3 for the
    // permutation-block codes, 16 / 4, or 32 / 16
    const bool exact12 = g.max_cdeg <= 12 && g.max_vdeg <= 3;
    const bool small = g.max_cdeg <= 16 && g.max_vdeg <= 4, pair = damping <= 0.0;
    // quasi-cyclic with 3 x 12 blocks: rows computed from the block maps;
    // with pair sums the variable pass stages a column block's rows in LDS
    const bool qc = g.qc_tab && g.qc_J == 3 && g.qc_L == 12 && exact12;
    const size_t lds = f64::var_pass_lds_bytes(g.side[0].m);
    const bool qc_lds = qc && pair && lds <= kF64MaxLds;
#define JBP_F64(C, V, PR, Q)                                                                                \
    hipLaunchKernelGGL((decode_f64_kernel<C, V, PR, Q>), dim3(grid), dim3(block), (PR && Q) ? lds : 0, st, g, \
                       tp, count, nt, max_iter, damping, scratch, stride, counter)
    if (freeze) {
#define JBP_F64F(C, V, PR)                                                                             \
    hipLaunchKernelGGL((decode_f64_kernel<C, V, PR, false, true>), dim3(grid), dim3(block), 0, st, g, tp, \
                       count, nt, max_iter, damping, scratch, stride, counter)
        if (exact12 && pair) JBP_F64F(12, 3, true);
        else if (exact12) JBP_F64F(12, 3, false);
        else if (small && pair) JBP_F64F(16, 4, true);
        else if (small) JBP_F64F(16, 4, false);
        else if (pair) JBP_F64F(32, 16, true);
        else JBP_F64F(32, 16, false);
#undef JBP_F64F
        return;
    }
    if (qc_lds) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(decode_f64_kernel<12, 3, true, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        JBP_F64(12, 3, true, true);
    } else if (qc && !pair) JBP_F64(12, 3, false, true);
    else if (exact12 && pair) JBP_F64(12, 3, true, false);
    else if (exact12) JBP_F64(12, 3, false, false);
    else if (small && pair) JBP_F64(16, 4, true, false);
    else if (small) JBP_F64(16, 4, false, false);
    else if (pair) JBP_F64(32, 16, true, false);
    else JBP_F64(32, 16, false, false);
#undef JBP_F64
}

}  // namespace jbp
