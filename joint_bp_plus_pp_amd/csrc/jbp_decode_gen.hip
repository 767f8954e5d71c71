// FP32 joint BP for any CSS code given as check-major CSR (the hypergraph-
// product codes, permutation-block codes of other sizes): the same binary-LLR
// arithmetic as the LDS-resident decoder (jbp_bp32.h), with the message state
// in HBM.
//
// Slot-interleaved layout.  The engine keeps S = 64*groups trials resident,
// one per slot; slot j of group g is lane j of every wavefront that works on
// group g.  Every per-edge and per-variable array is stored [group][item][64
// lanes], so a wavefront that handles one edge (or variable) of one group
// reads or writes 256 contiguous bytes, and the graph indices it follows are
// wave-uniform (scalar loads).  Hard decisions and syndromes are 64-bit
// words per (group, variable) and (group, check): bit j belongs to slot j, so
// the syndrome test of a check for 64 trials is the XOR of its variables'
// words.
//
// One engine step runs one flooding iteration for every resident trial:
//   var     lambda -> (est, G_x, G_z) per variable and slot
//   check   syndrome test of est (OR of mismatches per group) and the new
//           lambda of every edge from G - lambda_old
//   book    per slot: stop (converged / max_iter), or count the iteration;
//           stopped and empty slots take the next trial from a counter
//   emit    hard decisions of the stopped trials to their bit planes
//   load    syndrome words of the refilled slots (a 64x32 bit transpose by
//           wave ballots)
// A refilled slot is marked fresh: the var and check kernels read its old
// messages as 0 for one step, so no per-slot clearing pass is needed.
// Trials are independent of the slot they run in, so results do not depend
// on S or on the order in which trials finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "jbp_bp32.h"
#include "jbp_gen.h"

namespace jbp {
namespace {

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
    uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ PriorF slot_prior(const GenState &s, int group, int lane) {
    PriorF pr;
    const float *q = s.prior32 + static_cast<size_t>(group) * kPrior32 * 64 + lane;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pr.lp[i] = q[i * 64];
        pr.fx[i] = q[(4 + i) * 64];
        pr.fz[i] = q[(8 + i) * 64];
    }
    return pr;
}

// grid (ceil(n / (4*VPW)), groups), block 256: wave w handles VPW variables.
template <int VPW>
__global__ void __launch_bounds__(256) gen_var_kernel(DevGraph g, GenState s) {
    const int group = blockIdx.y;
    const uint64_t act = uni64(s.active[group]);
    if (!act) return;
    const int lane = threadIdx.x & 63;
    const bool fresh = (uni64(s.fresh[group]) >> lane) & 1u;
    const PriorF pr = slot_prior(s, group, lane);
    const int v0 = (blockIdx.x * 4 + wave_id()) * VPW;
    const int v1 = min(v0 + VPW, g.n);
    const size_t E0 = g.side[0].E, E1 = g.side[1].E;
    const float *lam0 = s.lam[0] + static_cast<size_t>(group) * E0 * 64 + lane;
    const float *lam1 = s.lam[1] + static_cast<size_t>(group) * E1 * 64 + lane;
    float *gb0 = s.gb[0] + static_cast<size_t>(group) * g.n * 64 + lane;
    float *gb1 = s.gb[1] + static_cast<size_t>(group) * g.n * 64 + lane;
    uint64_t *ew0 = s.estw[0] + static_cast<size_t>(group) * g.n;
    uint64_t *ew1 = s.estw[1] + static_cast<size_t>(group) * g.n;
    for (int v = v0; v < v1; ++v) {
        float lx = 0.0f, lz = 0.0f;
        const int a0 = g.side[0].var_off[v], b0 = g.side[0].var_off[v + 1];
        for (int i = a0; i < b0; ++i) lx += lam0[static_cast<size_t>(g.side[0].var_edge[i]) * 64];
        const int a1 = g.side[1].var_off[v], b1 = g.side[1].var_off[v + 1];
        for (int i = a1; i < b1; ++i) lz += lam1[static_cast<size_t>(g.side[1].var_edge[i]) * 64];
        if (fresh) {
            lx = 0.0f;
            lz = 0.0f;
        }
        const int best = hard_decision(lx, lz, pr);
        const uint64_t bx = __ballot(best & 1);
        const uint64_t bz = __ballot(best >> 1);
        if (lane == 0) {
            ew0[v] = bx;
            ew1[v] = bz;
        }
        gb0[static_cast<size_t>(v) * 64] = lx + extrinsic_term(lz, pr.fx);
        gb1[static_cast<size_t>(v) * 64] = lz + extrinsic_term(lx, pr.fz);
    }
}

// Check update of deg <= MAXC inputs mu[] in place (the pair recursion of
// the LDS decoder's check_row), syndrome bit syn.
template <int MAXC>
__device__ __forceinline__ void check_update(float (&mu)[MAXC], int deg, uint32_t syn) {
    uint32_t neg = 0;
    float pe[MAXC + 1], po[MAXC + 1];
    pe[0] = 1.0f;
    po[0] = 0.0f;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        if (i < deg) {
            neg |= (__float_as_uint(mu[i]) >> 31) << i;
            float e = fast_exp2(-fabsf(mu[i]));
            mu[i] = e;
            pe[i + 1] = pe[i] + po[i] * e;
            po[i + 1] = pe[i] * e + po[i];
        }
    }
    const uint32_t negall = (__popc(neg) & 1u) ^ syn;
    float se = 1.0f, so = 0.0f;
#pragma unroll
    for (int i = MAXC - 1; i >= 0; --i) {
        if (i < deg) {
            float e = mu[i];
            float PE = pe[i] * se + po[i] * so;
            float PO = pe[i] * so + po[i] * se;
            float mag = fminf(fast_log2(PE) - fast_log2(PO), kLlrMax);
            uint32_t sg = negall ^ ((neg >> i) & 1u);
            mu[i] = sg ? -mag : mag;
            float ne = se + so * e;
            float no = se * e + so;
            se = ne;
            so = no;
        }
    }
}

// grid (ceil(chunks / 4), groups), block 256; chunk q < nch0 covers X checks
// [q*CPW, (q+1)*CPW), the rest Z checks.
template <int MAXC, int CPW>
__global__ void __launch_bounds__(256) gen_check_kernel(DevGraph g, GenState s, int nch0, int nch) {
    const int group = blockIdx.y;
    const uint64_t act = uni64(s.active[group]);
    if (!act) return;
    const int q = blockIdx.x * 4 + wave_id();
    if (q >= nch) return;
    const int lane = threadIdx.x & 63;
    const bool fresh = (uni64(s.fresh[group]) >> lane) & 1u;
    const int side = q < nch0 ? 0 : 1;
    const DevSide sd = side == 0 ? g.side[0] : g.side[1];
    const int c0 = (side == 0 ? q : q - nch0) * CPW;
    const int c1 = min(c0 + CPW, sd.m);
    float *lam = (side == 0 ? s.lam[0] : s.lam[1]) + static_cast<size_t>(group) * sd.E * 64 + lane;
    const float *gb = (side == 0 ? s.gb[0] : s.gb[1]) + static_cast<size_t>(group) * g.n * 64 + lane;
    const uint64_t *ew = (side == 0 ? s.estw[0] : s.estw[1]) + static_cast<size_t>(group) * g.n;
    const uint64_t *sw = (side == 0 ? s.synw[0] : s.synw[1]) + static_cast<size_t>(group) * sd.m;
    uint64_t bad = 0;
    for (int c = c0; c < c1; ++c) {
        const int e0 = sd.chk_off[c];
        const int deg = sd.chk_off[c + 1] - e0;
        float mu[MAXC];
        uint64_t par = 0;
#pragma unroll
        for (int i = 0; i < MAXC; ++i) {
            if (i < deg) {
                const int v = sd.edge_var[e0 + i];
                par ^= ew[v];
                const float old = fresh ? 0.0f : lam[static_cast<size_t>(e0 + i) * 64];
                mu[i] = gb[static_cast<size_t>(v) * 64] - old;
            }
        }
        const uint64_t syn = sw[c];
        bad |= par ^ syn;
        check_update<MAXC>(mu, deg, static_cast<uint32_t>((syn >> lane) & 1u));
#pragma unroll
        for (int i = 0; i < MAXC; ++i)
            if (i < deg) lam[static_cast<size_t>(e0 + i) * 64] = mu[i];
    }
    bad = uni64(bad);
    if (bad && lane == 0) atomicOr(reinterpret_cast<unsigned long long *>(s.bad + group),
                                   static_cast<unsigned long long>(bad));
}

// grid = groups, block 64 (lane = slot).  Decides which slots stop, records
// their outcome, and hands the next trials to stopped and empty slots.
__global__ void __launch_bounds__(64) gen_book_kernel(GenState s, TrialPlanes tp, int count,
                                                      NoiseTab nt, int max_iter, int warmup,
                                                      int interval) {
    const int group = blockIdx.x;
    const int lane = threadIdx.x;
    const int slot = group * 64 + lane;
    const uint64_t act = s.active[group];
    const uint64_t badm = s.bad[group];
    const bool mine = (act >> lane) & 1u;
    bool fin = false;
    int matched = 0;
    int k = 0;
    if (mine) {
        k = s.slot_iter[slot];
        const bool bad = (badm >> lane) & 1u;
        const bool test = k == max_iter || (k > warmup && ((k - warmup) % interval) == 0);
        if (k >= 1 && test && !bad) {
            fin = true;
            matched = 1;
        } else if (k >= max_iter) {
            fin = true;
        } else {
            s.slot_iter[slot] = k + 1;
        }
    }
    const uint64_t finm = __ballot(fin);
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (finm) {
        int base = 0;
        if (lane == 0) {
            base = atomicAdd(&s.ctrl[2], __popcll(finm));
            atomicAdd(&s.ctrl[1], __popcll(finm));
        }
        base = __shfl(base, 0);
        if (fin) {
            const int trial = s.slot_trial[slot];
            tp.iters[trial] = k;
            tp.syn_ok[trial] = static_cast<uint8_t>(matched);
            s.fin[base + __popcll(finm & below)] = make_int2(trial, slot);
        }
    }
    const bool free_slot = !mine || fin;
    const uint64_t freem = __ballot(free_slot);
    uint64_t newact = act & ~finm;
    uint64_t refill = 0;
    if (freem) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&s.ctrl[0], __popcll(freem));
        base = __shfl(base, 0);
        const int t = base + __popcll(freem & below);
        const bool take = free_slot && t < count;
        if (take) {
            s.slot_trial[slot] = t;
            s.slot_iter[slot] = 0;
            const float *q = nt.prior32 + kPrior32 * nt.entry(t);
            float *dst = s.prior32 + static_cast<size_t>(group) * kPrior32 * 64 + lane;
#pragma unroll
            for (int i = 0; i < kPrior32; ++i) dst[i * 64] = q[i];
        }
        refill = __ballot(take);
        newact |= refill;
    }
    if (lane == 0) {
        s.active[group] = newact;
        s.fresh[group] = refill;
        s.refill[group] = refill;
        s.bad[group] = 0;
    }
}

// grid = any, block 64: one stopped trial per block iteration; its hard
// decisions (bit `slot % 64` of every estimate word) become plane words.
__global__ void __launch_bounds__(64) gen_emit_kernel(DevGraph g, GenState s, TrialPlanes tp) {
    const int nfin = s.ctrl[2];
    const int lane = threadIdx.x;
    for (int i = blockIdx.x; i < nfin; i += gridDim.x) {
        const int2 f = s.fin[i];
        const int group = f.y >> 6, sl = f.y & 63;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const uint64_t *ew = s.estw[side] + static_cast<size_t>(group) * g.n;
            uint32_t *out = (side == 0 ? tp.est_x : tp.est_z) + static_cast<size_t>(f.x) * g.wn;
            for (int v0 = 0; v0 < g.n; v0 += 64) {
                const int v = v0 + lane;
                const uint64_t w = v < g.n ? ew[v] : 0ull;
                const uint64_t bits = __ballot((w >> sl) & 1ull);
                const int wd = (v0 >> 5) + lane;
                if (lane < 2 && wd < g.wn) out[wd] = static_cast<uint32_t>(bits >> (32 * lane));
            }
        }
    }
}

// grid (ceil(maxw / WPB), groups), block 64: syndrome words of the slots
// refilled by the book pass.  Lane j loads word w of its trial's plane; 32
// ballots turn the 64 x 32 bit block into the 32 check words.
constexpr int kLoadWords = 8;
__global__ void __launch_bounds__(64) gen_load_kernel(DevGraph g, GenState s, TrialPlanes tp) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) s.ctrl[2] = 0;  // emit list consumed
    const int group = blockIdx.y;
    const uint64_t rf = uni64(s.refill[group]);
    if (!rf) return;
    const int lane = threadIdx.x;
    const bool mine = (rf >> lane) & 1u;
    const int trial = mine ? s.slot_trial[group * 64 + lane] : 0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int m = g.side[side].m, wm = g.wm[side];
        const uint32_t *plane = (side == 0 ? tp.syn_x : tp.syn_z) + static_cast<size_t>(trial) * wm;
        uint64_t *sw = s.synw[side] + static_cast<size_t>(group) * m;
        const int w0 = blockIdx.x * kLoadWords;
        const int w1 = min(w0 + kLoadWords, wm);
        for (int w = w0; w < w1; ++w) {
            const uint32_t word = mine ? plane[w] : 0u;
            uint64_t col = 0;
#pragma unroll 8
            for (int b = 0; b < 32; ++b) {
                const uint64_t bal = __ballot((word >> b) & 1u);
                if (lane == b) col = bal;
            }
            const int c = 32 * w + lane;
            if (lane < 32 && c < m) sw[c] = (sw[c] & ~rf) | (col & rf);
        }
    }
}

template <int MAXC>
void launch_check(const DevGraph &g, const GenState &s, hipStream_t st) {
    constexpr int CPW = 4;
    const int nch0 = (g.side[0].m + CPW - 1) / CPW;
    const int nch = nch0 + (g.side[1].m + CPW - 1) / CPW;
    hipLaunchKernelGGL((gen_check_kernel<MAXC, CPW>), dim3((nch + 3) / 4, s.groups), dim3(256), 0,
                       st, g, s, nch0, nch);
}

}  // namespace

size_t gen_state_bytes(const DevGraph &g, int groups) {
    const size_t G = static_cast<size_t>(groups), S = 64 * G;
    size_t b = 0;
    b += S * (static_cast<size_t>(g.side[0].E) + g.side[1].E) * sizeof(float);  // lam
    b += 2 * S * static_cast<size_t>(g.n) * sizeof(float);                      // gb
    b += 2 * G * static_cast<size_t>(g.n) * 8;                                   // estw
    b += G * (static_cast<size_t>(g.side[0].m) + g.side[1].m) * 8;               // synw
    b += 4 * G * 8 + 2 * S * 4 + S * kPrior32 * 4 + S * 8 + 64;
    return b;
}

bool gen_supported(const DevGraph &g) { return g.max_cdeg <= 32 && g.n > 0; }

void gen_carve(const DevGraph &g, int groups, void *base, GenState &s) {
    const size_t G = static_cast<size_t>(groups), S = 64 * G;
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) & ~static_cast<size_t>(255);
        return q;
    };
    s.groups = groups;
    for (int k = 0; k < 2; ++k) s.lam[k] = reinterpret_cast<float *>(take(S * g.side[k].E * 4));
    for (int k = 0; k < 2; ++k) s.gb[k] = reinterpret_cast<float *>(take(S * g.n * 4));
    for (int k = 0; k < 2; ++k) s.estw[k] = reinterpret_cast<uint64_t *>(take(G * g.n * 8));
    for (int k = 0; k < 2; ++k) s.synw[k] = reinterpret_cast<uint64_t *>(take(G * g.side[k].m * 8));
    // bad | active | fresh | refill, contiguous (cleared by one fill)
    s.bad = reinterpret_cast<uint64_t *>(take(4 * G * 8));
    s.active = s.bad + G;
    s.fresh = s.bad + 2 * G;
    s.refill = s.bad + 3 * G;
    s.slot_trial = reinterpret_cast<int *>(take(S * 4));
    s.slot_iter = reinterpret_cast<int *>(take(S * 4));
    s.prior32 = reinterpret_cast<float *>(take(S * kPrior32 * 4));
    s.fin = reinterpret_cast<int2 *>(take(S * 8));
    s.ctrl = reinterpret_cast<int *>(take(64));
}

size_t gen_carve_bytes(const DevGraph &g, int groups) {
    // gen_carve rounds every array up to 256 B
    return gen_state_bytes(g, groups) + 16 * 256;
}

int gen_decode(const DevGraph &g, const GenState &s, TrialPlanes tp, int count, const NoiseTab &nt,
               int max_iter, int warmup, int interval, int *h_ctrl, hipEvent_t *evs,
               hipStream_t st, int64_t *steps_out) {
    *steps_out = 0;
    if (count <= 0) return 0;
    const int G = s.groups;
    // reset: every slot empty, counters zero (a fill kernel, not a memset,
    // so the sequence can be captured in a graph)
    launch_fill_i32(reinterpret_cast<int *>(s.bad), 0, 4 * 2 * G, st);  // bad, active, fresh, refill
    launch_fill_i32(s.ctrl, 0, 16, st);
    const int wmax = std::max(g.wm[0], g.wm[1]);
    const dim3 load_grid((wmax + kLoadWords - 1) / kLoadWords, G);
    const dim3 var_grid((g.n + 4 * 16 - 1) / (4 * 16), G);
    const int emit_grid = std::min(64 * G, 2048);
    interval = std::max(1, interval);
    warmup = std::max(0, warmup);
    hipLaunchKernelGGL(gen_book_kernel, dim3(G), dim3(64), 0, st, s, tp, count, nt, max_iter,
                       warmup, interval);
    hipLaunchKernelGGL(gen_load_kernel, load_grid, dim3(64), 0, st, g, s, tp);
    // Each trial occupies its slot for at most max_iter + 1 steps.
    const int64_t waves = (count + 64LL * G - 1) / (64LL * G);
    constexpr int kBatch = 4;
    const int64_t limit = (waves + 1) * (static_cast<int64_t>(max_iter) + 1) + 4 * kBatch;
    int64_t steps = 0;
    int pend = -1;  // event slot whose copy is pending
    for (;;) {
        for (int b = 0; b < kBatch; ++b) {
            hipLaunchKernelGGL((gen_var_kernel<16>), var_grid, dim3(256), 0, st, g, s);
            if (g.max_cdeg <= 8) launch_check<8>(g, s, st);
            else if (g.max_cdeg <= 16) launch_check<16>(g, s, st);
            else launch_check<32>(g, s, st);
            hipLaunchKernelGGL(gen_book_kernel, dim3(G), dim3(64), 0, st, s, tp, count, nt,
                               max_iter, warmup, interval);
            hipLaunchKernelGGL(gen_emit_kernel, dim3(emit_grid), dim3(64), 0, st, g, s, tp);
            hipLaunchKernelGGL(gen_load_kernel, load_grid, dim3(64), 0, st, g, s, tp);
        }
        steps += kBatch;
        // finished count after this batch, read back one batch later
        const int slot = static_cast<int>((steps / kBatch) & 1);
        hipError_t e = hipMemcpyAsync(h_ctrl + slot, s.ctrl + 1, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(evs[slot], st);
        if (e != hipSuccess) return static_cast<int>(e);
        if (pend >= 0) {
            e = hipEventSynchronize(evs[pend]);
            if (e != hipSuccess) return static_cast<int>(e);
            if (h_ctrl[pend] >= count) break;
        }
        pend = slot;
        if (steps > limit) return -1;
    }
    hipError_t e = hipStreamSynchronize(st);
    *steps_out = steps;
    return static_cast<int>(e);
}

}  // namespace jbp
