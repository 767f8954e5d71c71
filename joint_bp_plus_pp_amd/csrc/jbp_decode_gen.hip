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
//   load    syndrome words of the refilled slots (emit and load share one
//           launch)
// Steps go out in batches of four (launched, or replayed from a captured
// graph on request).
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

// Variable pass of block (bx, group), block 256: wave w handles VPW
// variables; their edge lists come from the padded table (MAXV entries per
// variable and side, -1 = none), so every gather of a variable is issued at
// once.
template <int VPW, int MAXV>
__device__ __forceinline__ void var_block(const DevGraph &g, const GenState &s, int bx, int group) {
    const uint64_t act = uni64(s.active[group]);
    if (!act) return;
    const int lane = threadIdx.x & 63;
    const bool fresh = (uni64(s.fresh[group]) >> lane) & 1u;
    const uint32_t keep = fresh ? 0u : ~0u;  // a fresh slot reads its old messages as 0
    const PriorF pr = slot_prior(s, group, lane);
    const int v0 = (bx * 4 + wave_id()) * VPW;
    const int v1 = min(v0 + VPW, g.n);
    const size_t E0 = g.side[0].E, E1 = g.side[1].E;
    const float *lam0 = s.lam[0] + static_cast<size_t>(group) * (E0 + 1) * 64 + lane;
    const float *lam1 = s.lam[1] + static_cast<size_t>(group) * (E1 + 1) * 64 + lane;
    float *gb0 = s.gb[0] + static_cast<size_t>(group) * g.n * 64 + lane;
    float *gb1 = s.gb[1] + static_cast<size_t>(group) * g.n * 64 + lane;
    uint64_t *ew0 = s.estw[0] + static_cast<size_t>(group) * g.n;
    uint64_t *ew1 = s.estw[1] + static_cast<size_t>(group) * g.n;
    // Batches of VB variables: all gathers of a batch before its stores (the
    // stores to gb cannot be proven disjoint from the next loads of lam).
    constexpr int VB = 4;
    for (int vb = v0; vb < v1; vb += VB) {
        float lx[VB], lz[VB];
#pragma unroll
        for (int j = 0; j < VB; ++j) {
            const int v = min(vb + j, v1 - 1);  // a short last batch repeats its final variable
            const int *p0 = g.vpad[0] + static_cast<size_t>(v) * MAXV;
            const int *p1 = g.vpad[1] + static_cast<size_t>(v) * MAXV;
            // padding entries point at edge E, whose message is always 0
            int i0[MAXV], i1[MAXV];
#pragma unroll
            for (int k = 0; k < MAXV; ++k) {
                i0[k] = p0[k];
                i1[k] = p1[k];
            }
            float x[MAXV], z[MAXV];
#pragma unroll
            for (int k = 0; k < MAXV; ++k) {
                x[k] = lam0[static_cast<size_t>(i0[k]) * 64];
                z[k] = lam1[static_cast<size_t>(i1[k]) * 64];
            }
            float sx = 0.0f, sz = 0.0f;
#pragma unroll
            for (int k = 0; k < MAXV; ++k) {
                sx += x[k];
                sz += z[k];
            }
            lx[j] = __uint_as_float(__float_as_uint(sx) & keep);
            lz[j] = __uint_as_float(__float_as_uint(sz) & keep);
        }
#pragma unroll
        for (int j = 0; j < VB; ++j) {
            const int v = vb + j;
            if (v < v1) {
                const int best = hard_decision(lx[j], lz[j], pr);
                const uint64_t bx = __ballot(best & 1);
                const uint64_t bz = __ballot(best >> 1);
                if (lane == 0) {
                    ew0[v] = bx;
                    ew1[v] = bz;
                }
                gb0[static_cast<size_t>(v) * 64] = lx[j] + extrinsic_term(lz[j], pr.fx);
                gb1[static_cast<size_t>(v) * 64] = lz[j] + extrinsic_term(lx[j], pr.fz);
            }
        }
    }
}

// grid (ceil(n / (4*VPW)), groups), block 256.
template <int VPW, int MAXV>
__global__ void __launch_bounds__(256) gen_var_kernel(DevGraph g, GenState s) {
    // the previous step's stopped-trial list has been emitted: clear it for
    // this step's book pass
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) s.ctrl[2] = 0;
    var_block<VPW, MAXV>(g, s, blockIdx.x, blockIdx.y);
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

// Damping of a variable-to-check message, the reference's scaled_mix
// (jointbp_ets.cpp:1903-1911): normalize((1 - d) new + d old) of the 4-ary
// messages.  A check reads only the pair sums of its side, and mixing
// commutes with forming them, so on the binary probabilities
// (q0, q1) = (1 - d) (n0, n1) + d (o0, o1), returned as log2(q0 / q1).
__device__ __forceinline__ float mix_llr(float mn, float mo, float d) {
    const float en = fast_exp2(-fabsf(mn)), eo = fast_exp2(-fabsf(mo));
    const float rn = __builtin_amdgcn_rcpf(1.0f + en), ro = __builtin_amdgcn_rcpf(1.0f + eo);
    const float n0 = mn >= 0.0f ? rn : en * rn, n1 = mn >= 0.0f ? en * rn : rn;
    const float o0 = mo >= 0.0f ? ro : eo * ro, o1 = mo >= 0.0f ? eo * ro : ro;
    const float q0 = (1.0f - d) * n0 + d * o0, q1 = (1.0f - d) * n1 + d * o1;
    return fast_log2(q0) - fast_log2(q1);
}

// One wavefront per group (lane = slot): decides which slots stop, records
// their outcome, and hands the next trials to stopped and empty slots.
__device__ void book_group(const GenState &s, const TrialPlanes &tp, int count, const NoiseTab &nt,
                           int max_iter, int warmup, int interval, int group, int lane) {
    const int slot = group * 64 + lane;
    const uint64_t act = s.active[group];
    // OR of the group's kBadWays words (lane i reads word i; set by atomics,
    // read at device scope), then clear them
    uint64_t badm = __hip_atomic_load(reinterpret_cast<unsigned long long *>(s.bad + group * kBadWays + lane),
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s.bad[group * kBadWays + lane] = 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = __shfl_xor(static_cast<uint32_t>(badm), m);
        const uint32_t hi = __shfl_xor(static_cast<uint32_t>(badm >> 32), m);
        badm |= (static_cast<uint64_t>(hi) << 32) | lo;
    }
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
        const int k = base + __popcll(freem & below);
        const bool take = free_slot && k < count;
        if (take) {
            const int t = tp.order ? tp.order[k] : k;
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
    }
}

// grid = groups, block 64.
__global__ void __launch_bounds__(64) gen_book_kernel(GenState s, TrialPlanes tp, int count,
                                                      NoiseTab nt, int max_iter, int warmup,
                                                      int interval) {
    book_group(s, tp, count, nt, max_iter, warmup, interval, blockIdx.x, threadIdx.x);
}

// Check messages of chunk q (CPW checks of one side) of one group.
template <int MAXC, int CPW, int DEG>
__device__ __forceinline__ void check_chunk(const DevGraph &g, const GenState &s, int group, int q,
                                            int nch0) {
    const int lane = threadIdx.x & 63;
    const bool fresh = (uni64(s.fresh[group]) >> lane) & 1u;
    const uint32_t keep = fresh ? 0u : ~0u;  // a fresh slot reads its old messages as 0
    const int side = q < nch0 ? 0 : 1;
    const DevSide sd = side == 0 ? g.side[0] : g.side[1];
    const int c0 = (side == 0 ? q : q - nch0) * CPW;
    const int c1 = min(c0 + CPW, sd.m);
    float *lam = (side == 0 ? s.lam[0] : s.lam[1]) + static_cast<size_t>(group) * (sd.E + 1) * 64 + lane;
    const float *gb = (side == 0 ? s.gb[0] : s.gb[1]) + static_cast<size_t>(group) * g.n * 64 + lane;
    const uint64_t *ew = (side == 0 ? s.estw[0] : s.estw[1]) + static_cast<size_t>(group) * g.n;
    const uint64_t *sw = (side == 0 ? s.synw[0] : s.synw[1]) + static_cast<size_t>(group) * sd.m;
    // damping: a slot's first step sends the prior message unmixed
    float *mp = s.mprev[side] ? s.mprev[side] + static_cast<size_t>(group) * (sd.E + 1) * 64 + lane : nullptr;
    auto damp = [&](float m, size_t e) {
        if (!mp) return m;
        const float v = fresh ? m : mix_llr(m, mp[e * 64], s.damping);
        mp[e * 64] = v;
        return v;
    };
    uint64_t bad = 0;
    if constexpr (DEG > 0) {
        // Load phase for all CPW checks, then update and store: stores of one
        // check cannot be proven disjoint from the next check's loads, so
        // interleaving them would serialize the gathers.
        float mu[CPW][DEG];
        uint64_t par[CPW];
        int e0s[CPW];
        int ev[CPW][DEG];
#pragma unroll
        for (int j = 0; j < CPW; ++j)
            e0s[j] = sd.chk_off[min(c0 + j, sd.m - 1)];  // a short last chunk repeats its final check (not stored)
#pragma unroll
        for (int j = 0; j < CPW; ++j)
#pragma unroll
            for (int i = 0; i < DEG; ++i) ev[j][i] = sd.edge_var[e0s[j] + i];
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            uint64_t pj = 0;
#pragma unroll
            for (int i = 0; i < DEG; ++i) {
                const int v = ev[j][i];
                pj ^= ew[v];
                const float old = __uint_as_float(
                    __float_as_uint(lam[static_cast<size_t>(e0s[j] + i) * 64]) & keep);
                mu[j][i] = damp(gb[static_cast<size_t>(v) * 64] - old, static_cast<size_t>(e0s[j] + i));
            }
            par[j] = pj;
        }
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            if (c0 + j < c1) {
                const uint64_t syn = sw[c0 + j];
                bad |= par[j] ^ syn;
                check_update<DEG>(mu[j], DEG, static_cast<uint32_t>((syn >> lane) & 1u));
#pragma unroll
                for (int i = 0; i < DEG; ++i) lam[static_cast<size_t>(e0s[j] + i) * 64] = mu[j][i];
            }
        }
    } else {
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
                    const float old = __uint_as_float(
                        __float_as_uint(lam[static_cast<size_t>(e0 + i) * 64]) & keep);
                    mu[i] = damp(gb[static_cast<size_t>(v) * 64] - old, static_cast<size_t>(e0 + i));
                }
            }
            const uint64_t syn = sw[c];
            bad |= par ^ syn;
            check_update<MAXC>(mu, deg, static_cast<uint32_t>((syn >> lane) & 1u));
#pragma unroll
            for (int i = 0; i < MAXC; ++i)
                if (i < deg) lam[static_cast<size_t>(e0 + i) * 64] = mu[i];
        }
    }
    // one of kBadWays words per group, chosen by block, so the wavefronts of
    // a group do not all update one address; the book pass ORs them
    bad = uni64(bad);
    if (bad && lane == 0)
        atomicOr(reinterpret_cast<unsigned long long *>(s.bad + group * kBadWays + blockIdx.x % kBadWays),
                 static_cast<unsigned long long>(bad));
}

// grid (ceil(chunks / 4), groups), block 256; chunk q < nch0 covers X checks
// [q*CPW, (q+1)*CPW), the rest Z checks.
// DEG > 0: every check has exactly DEG edges (no per-edge guards, so all
// loads of a check are issued together); DEG = 0: up to MAXC edges.
template <int MAXC, int CPW, int DEG>
__global__ void __launch_bounds__(256) gen_check_kernel(DevGraph g, GenState s, int nch0, int nch) {
    const int group = blockIdx.y;
    const uint64_t act = uni64(s.active[group]);
    const int q = blockIdx.x * 4 + wave_id();
    if (act && q < nch) check_chunk<MAXC, CPW, DEG>(g, s, group, q, nch0);
}

// Emit pass, block (x, y) of a (ceil(ceil(n/64) / 4), kEmitRows) grid of
// 256 threads: wave w turns variables [64 (4x + w), +64) of stopped trials
// y, y + kEmitRows, ... into plane words (bit `slot % 64` of each estimate
// word, one ballot).
constexpr int kEmitRows = 32;
__device__ __forceinline__ void emit_block(const DevGraph &g, const GenState &s, const TrialPlanes &tp,
                                           int bx, int by) {
    const int nfin = s.ctrl[2];
    if (by >= nfin) return;
    const int lane = threadIdx.x & 63;
    const int v0 = (bx * 4 + wave_id()) * 64;
    if (v0 >= g.n) return;
    const int v = v0 + lane;
    const int wd = (v0 >> 5) + lane;
    for (int i = by; i < nfin; i += kEmitRows) {
        const int2 f = s.fin[i];
        const int group = f.y >> 6, sl = f.y & 63;
        const uint64_t wx = v < g.n ? s.estw[0][static_cast<size_t>(group) * g.n + v] : 0ull;
        const uint64_t wz = v < g.n ? s.estw[1][static_cast<size_t>(group) * g.n + v] : 0ull;
        const uint64_t bx = __ballot((wx >> sl) & 1ull);
        const uint64_t bz = __ballot((wz >> sl) & 1ull);
        if (lane < 2 && wd < g.wn) {
            tp.est_x[static_cast<size_t>(f.x) * g.wn + wd] = static_cast<uint32_t>(bx >> (32 * lane));
            tp.est_z[static_cast<size_t>(f.x) * g.wn + wd] = static_cast<uint32_t>(bz >> (32 * lane));
        }
    }
}

// Load pass, block (x, group) of a (ceil(max(mX, mZ) / 256), groups) grid of
// 256 threads: syndrome words of the slots refilled by the book pass.  Thread c rebuilds word c of each side
// from the refilled slots' planes (a loop over the set bits of the refill
// mask; each plane word is read by 32 neighbouring threads at once).
__device__ __forceinline__ void load_block(const DevGraph &g, const GenState &s, const TrialPlanes &tp,
                                           int bx, int group) {
    const uint64_t rf = uni64(s.refill[group]);
    if (!rf) return;
    const int c = bx * 256 + threadIdx.x;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int m = g.side[side].m, wm = g.wm[side];
        if (c >= m) continue;
        const uint32_t *planes = side == 0 ? tp.syn_x : tp.syn_z;
        uint64_t acc = 0;
        for (uint64_t r = rf; r; r &= r - 1) {
            const int j = __builtin_ctzll(r);
            const int trial = uni(s.slot_trial[group * 64 + j]);
            acc |= static_cast<uint64_t>((planes[static_cast<size_t>(trial) * wm + (c >> 5)] >> (c & 31)) & 1u) << j;
        }
        uint64_t *sw = s.synw[side] + static_cast<size_t>(group) * m;
        sw[c] = (sw[c] & ~rf) | acc;
    }
}

// Emit and load in one launch: blocks [0, ex*kEmitRows) run the emit pass
// (block b as (b % ex, b / ex)), the rest the load pass (as (b' % lx, b' /
// lx)).  They touch disjoint data (estimate words and output planes; syndrome
// words).
__global__ void __launch_bounds__(256) gen_post_kernel(DevGraph g, GenState s, TrialPlanes tp, int ex,
                                                       int lx) {
    const int b = blockIdx.x;
    if (b < ex * kEmitRows) emit_block(g, s, tp, b % ex, b / ex);
    else load_block(g, s, tp, (b - ex * kEmitRows) % lx, (b - ex * kEmitRows) / lx);
}

// grid = groups, block 64: the extra edge E of each group and side, which
// padded variable entries read, holds message 0.
__global__ void __launch_bounds__(64) gen_zero_kernel(DevGraph g, GenState s) {
    const int group = blockIdx.x;
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const size_t E = g.side[k].E;
        s.lam[k][(static_cast<size_t>(group) * (E + 1) + E) * 64 + lane] = 0.0f;
    }
}

struct BookArgs {
    TrialPlanes tp;
    int count;
    NoiseTab nt;
    int max_iter, warmup, interval;
};

// ---- several codes in one engine (gen_decode_mix) ----
// Engine groups [c.group0, c.group0 + c.s.groups) belong to code c: every
// kernel of a step covers all codes, each block serving the code its group
// (or, in the post pass, its block range) falls in.  Each code keeps its
// own state, trials, channel table and counters, so a trial's decode is the
// same as in that code's own engine.  The code index of a kernel body is a
// template argument, so each code's fields are read from the kernel
// arguments at a fixed offset.

// Check-pass variants of a mixed engine.  Every variant compiled into a
// kernel costs that kernel its registers (a kernel holds the most any of its
// bodies needs), so the mixed kernels are instantiated per pair of codes'
// variants, from few: exact degree 7 (hypergraph products of (3,4)-regular
// codes), exact degree 12 (the included code) and a generic body for any
// degree up to 32; likewise the variable pass per pair of padded widths.
enum { kMixChk7, kMixChk12, kMixChkAny };
inline int mix_check_variant(const DevGraph &g, int &cpw) {
    if (g.min_cdeg == g.max_cdeg && g.max_cdeg == 7) {
        cpw = 4;
        return kMixChk7;
    }
    if (g.min_cdeg == g.max_cdeg && g.max_cdeg == 12) {
        cpw = 2;
        return kMixChk12;
    }
    cpw = 4;
    return kMixChkAny;
}

template <int C, int W>
__device__ __forceinline__ void var_code(const GenMix &m, int bx, int group) {
    const GenCode &k = m.c[C];
    if (bx == 0 && group == 0 && threadIdx.x == 0) k.s.ctrl[2] = 0;
    if (bx >= k.vx) return;
    var_block<16, W>(k.g, k.s, bx, group);
}

template <int C, int K>
__device__ __forceinline__ void check_code(const GenMix &m, int group, int q) {
    const GenCode &k = m.c[C];
    const uint64_t act = uni64(k.s.active[group]);
    if (!act || q >= k.nch) return;
    if constexpr (K == kMixChk7) check_chunk<7, 4, 7>(k.g, k.s, group, q, k.nch0);
    else if constexpr (K == kMixChk12) check_chunk<12, 2, 12>(k.g, k.s, group, q, k.nch0);
    else check_chunk<32, 4, 0>(k.g, k.s, group, q, k.nch0);
}

template <int C>
__device__ __forceinline__ void post_code(const GenMix &m, int b) {
    const GenCode &k = m.c[C];
    if (b < k.ex * kEmitRows) emit_block(k.g, k.s, k.tp, b % k.ex, b / k.ex);
    else {
        b -= k.ex * kEmitRows;
        load_block(k.g, k.s, k.tp, b % k.lx, b / k.lx);
    }
}

__device__ __forceinline__ int code_of_group(const GenMix &m, int gy) {
    return (m.n > 1 && gy >= m.c[1].group0) ? 1 : 0;
}

// Variable and check passes run on one-dimensional grids: code c owns the
// blocks [c.var_b0, + vx * groups) (and [c.chk_b0, + ceil(nch / 4) * groups)),
// group-major, so a code whose graph needs fewer blocks per group launches
// no idle ones.

// grid = both codes' var blocks, block 256; W0 / W1: the codes' padded widths
template <int W0, int W1>
__global__ void __launch_bounds__(256) gen_var_mix_kernel(GenMix m) {
    const int b = blockIdx.x;
    if (m.n > 1 && b >= m.c[1].var_b0) {
        const int l = b - m.c[1].var_b0;
        var_code<1, W1>(m, l % m.c[1].vx, l / m.c[1].vx);
    } else {
        const int l = b - m.c[0].var_b0;
        var_code<0, W0>(m, l % m.c[0].vx, l / m.c[0].vx);
    }
}

// grid = both codes' check blocks, block 256; K0 / K1: the codes' check
// variants
template <int K0, int K1>
__global__ void __launch_bounds__(256) gen_check_mix_kernel(GenMix m) {
    const int b = blockIdx.x;
    if (m.n > 1 && b >= m.c[1].chk_b0) {
        const int l = b - m.c[1].chk_b0, cx = (m.c[1].nch + 3) / 4;
        check_code<1, K1>(m, l / cx, (l % cx) * 4 + wave_id());
    } else {
        const int l = b - m.c[0].chk_b0, cx = (m.c[0].nch + 3) / 4;
        check_code<0, K0>(m, l / cx, (l % cx) * 4 + wave_id());
    }
}

// grid = all groups, block 64
__global__ void __launch_bounds__(64) gen_book_mix_kernel(GenMix m) {
    const int gy = blockIdx.x;
    if (code_of_group(m, gy) == 0) {
        const GenCode &k = m.c[0];
        book_group(k.s, k.tp, k.count, k.nt, k.max_iter, k.warmup, k.interval, gy - k.group0, threadIdx.x);
    } else {
        const GenCode &k = m.c[1];
        book_group(k.s, k.tp, k.count, k.nt, k.max_iter, k.warmup, k.interval, gy - k.group0, threadIdx.x);
    }
}

// grid = the codes' post blocks one after another, block 256
__global__ void __launch_bounds__(256) gen_post_mix_kernel(GenMix m) {
    const int b = blockIdx.x;
    if (m.n > 1 && b >= m.c[1].post0) post_code<1>(m, b - m.c[1].post0);
    else post_code<0>(m, b - m.c[0].post0);
}

template <int MAXC, int DEG>
void launch_check(const DevGraph &g, const GenState &s, const BookArgs &b, hipStream_t st) {
    constexpr int CPW = (DEG == 0 || DEG <= 8) ? 4 : 2;  // registers: CPW*DEG messages in flight
    const int nch0 = (g.side[0].m + CPW - 1) / CPW;
    const int nch = nch0 + (g.side[1].m + CPW - 1) / CPW;
    hipLaunchKernelGGL((gen_check_kernel<MAXC, CPW, DEG>), dim3((nch + 3) / 4, s.groups), dim3(256),
                       0, st, g, s, nch0, nch);
    hipLaunchKernelGGL(gen_book_kernel, dim3(s.groups), dim3(64), 0, st, s, b.tp, b.count, b.nt,
                       b.max_iter, b.warmup, b.interval);
}

void launch_check_any(const DevGraph &g, const GenState &s, const BookArgs &b, hipStream_t st) {
    if (g.min_cdeg == g.max_cdeg) {
        switch (g.max_cdeg) {
            case 6: return launch_check<6, 6>(g, s, b, st);
            case 7: return launch_check<7, 7>(g, s, b, st);
            case 8: return launch_check<8, 8>(g, s, b, st);
            case 12: return launch_check<12, 12>(g, s, b, st);
            default: break;
        }
    }
    if (g.max_cdeg <= 8) launch_check<8, 0>(g, s, b, st);
    else if (g.max_cdeg <= 16) launch_check<16, 0>(g, s, b, st);
    else launch_check<32, 0>(g, s, b, st);
}

void launch_var(const DevGraph &g, const GenState &s, hipStream_t st) {
    const dim3 grid((g.n + 4 * 16 - 1) / (4 * 16), s.groups);
    if (g.vpad_w == 4) hipLaunchKernelGGL((gen_var_kernel<16, 4>), grid, dim3(256), 0, st, g, s);
    else if (g.vpad_w == 8) hipLaunchKernelGGL((gen_var_kernel<16, 8>), grid, dim3(256), 0, st, g, s);
    else hipLaunchKernelGGL((gen_var_kernel<16, 16>), grid, dim3(256), 0, st, g, s);
}

}  // namespace

size_t gen_state_bytes(const DevGraph &g, int groups, bool damped) {
    const size_t G = static_cast<size_t>(groups), S = 64 * G;
    size_t b = 0;
    b += (damped ? 2 : 1) * S * (static_cast<size_t>(g.side[0].E) + g.side[1].E + 2) * sizeof(float);  // lam (+ zero edge), mprev
    b += 2 * S * static_cast<size_t>(g.n) * sizeof(float);                      // gb
    b += 2 * G * static_cast<size_t>(g.n) * 8;                                   // estw
    b += G * (static_cast<size_t>(g.side[0].m) + g.side[1].m) * 8;               // synw
    b += (kBadWays + 3) * G * 8 + 2 * S * 4 + S * kPrior32 * 4 + S * 8 + 64;
    return b;
}

bool gen_supported(const DevGraph &g) { return g.max_cdeg <= 32 && g.n > 0; }

void gen_carve(const DevGraph &g, int groups, void *base, GenState &s, bool damped) {
    const size_t G = static_cast<size_t>(groups), S = 64 * G;
    char *p = static_cast<char *>(base);
    auto take = [&](size_t bytes) {
        char *q = p;
        p += (bytes + 255) & ~static_cast<size_t>(255);
        return q;
    };
    s.groups = groups;
    for (int k = 0; k < 2; ++k) s.lam[k] = reinterpret_cast<float *>(take(S * (g.side[k].E + 1) * 4));
    for (int k = 0; k < 2; ++k) s.gb[k] = reinterpret_cast<float *>(take(S * g.n * 4));
    for (int k = 0; k < 2; ++k) s.estw[k] = reinterpret_cast<uint64_t *>(take(G * g.n * 8));
    for (int k = 0; k < 2; ++k) s.synw[k] = reinterpret_cast<uint64_t *>(take(G * g.side[k].m * 8));
    // bad | active | fresh | refill, contiguous (cleared by one fill)
    s.bad = reinterpret_cast<uint64_t *>(take((kBadWays + 3) * G * 8));
    s.active = s.bad + kBadWays * G;
    s.fresh = s.active + G;
    s.refill = s.fresh + G;
    s.slot_trial = reinterpret_cast<int *>(take(S * 4));
    s.slot_iter = reinterpret_cast<int *>(take(S * 4));
    s.prior32 = reinterpret_cast<float *>(take(S * kPrior32 * 4));
    s.fin = reinterpret_cast<int2 *>(take(S * 8));
    s.ctrl = reinterpret_cast<int *>(take(64));
    for (int k = 0; k < 2; ++k)
        s.mprev[k] = damped ? reinterpret_cast<float *>(take(S * (g.side[k].E + 1) * 4)) : nullptr;
}

size_t gen_carve_bytes(const DevGraph &g, int groups, bool damped) {
    // gen_carve rounds every array up to 256 B
    return gen_state_bytes(g, groups, damped) + 18 * 256;
}

int gen_decode(const DevGraph &g, const GenState &s, TrialPlanes tp, int count, const NoiseTab &nt,
               int max_iter, int warmup, int interval, int *h_ctrl, hipEvent_t *evs,
               hipStream_t st, int64_t *steps_out, bool use_graph) {
    *steps_out = 0;
    if (count <= 0) return 0;
    const int G = s.groups;
    // reset: every slot empty, counters zero (a fill kernel, not a memset,
    // so the sequence can be captured in a graph)
    launch_fill_i32(reinterpret_cast<int *>(s.bad), 0, 2 * (kBadWays + 3) * G, st);  // bad, active, fresh, refill
    launch_fill_i32(s.ctrl, 0, 16, st);
    hipLaunchKernelGGL(gen_zero_kernel, dim3(G), dim3(64), 0, st, g, s);
    const int mmax = std::max(g.side[0].m, g.side[1].m);
    const int lx = (mmax + 255) / 256;
    const int ex = ((g.n + 63) / 64 + 3) / 4;
    const dim3 post_grid(ex * kEmitRows + lx * G);
    const BookArgs bargs{tp, count, nt, max_iter, std::max(0, warmup), std::max(1, interval)};
    interval = std::max(1, interval);
    warmup = std::max(0, warmup);
    hipLaunchKernelGGL(gen_book_kernel, dim3(G), dim3(64), 0, st, s, tp, count, nt, max_iter,
                       warmup, interval);
    hipLaunchKernelGGL(gen_post_kernel, post_grid, dim3(256), 0, st, g, s, tp, ex, lx);
    // Each trial occupies its slot for at most max_iter + 1 steps.
    const int64_t waves = (count + 64LL * G - 1) / (64LL * G);
    constexpr int kBatch = 4;
    const int64_t limit = (waves + 1) * (static_cast<int64_t>(max_iter) + 1) + 4 * kBatch;
    // One batch of kBatch steps (kernels only).
    auto enqueue_batch = [&]() -> hipError_t {
        for (int b = 0; b < kBatch; ++b) {
            launch_var(g, s, st);
            launch_check_any(g, s, bargs, st);  // check pass, then the book pass
            hipLaunchKernelGGL(gen_post_kernel, post_grid, dim3(256), 0, st, g, s, tp, ex, lx);
        }
        return hipGetLastError();
    };
    // With use_graph the batch is replayed as a captured graph (kernel nodes
    // only); a capture failure falls back to launches.
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess) {
            const hipError_t ce = enqueue_batch();
            const hipError_t ee = hipStreamEndCapture(st, &graph);
            if (ce != hipSuccess || ee != hipSuccess || !graph ||
                hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
                exec = nullptr;
            if (graph) (void)hipGraphDestroy(graph);
        }
        (void)hipGetLastError();
    }
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
    };
    int64_t steps = 0;
    int pend = -1;  // event slot whose copy is pending
    for (;;) {
        steps += kBatch;
        // the finished count after this batch, read back one batch later;
        // slot alternates between two copies and events
        const int slot = static_cast<int>((steps / kBatch) & 1);
        hipError_t e = exec ? hipGraphLaunch(exec, st) : enqueue_batch();
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_ctrl + slot, s.ctrl + 1, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(evs[slot], st);
        if (e != hipSuccess) {
            cleanup();
            return static_cast<int>(e);
        }
        if (pend >= 0) {
            e = hipEventSynchronize(evs[pend]);
            if (e != hipSuccess) {
                cleanup();
                return static_cast<int>(e);
            }
            if (h_ctrl[pend] >= count) break;
        }
        pend = slot;
        if (steps > limit) {
            (void)hipStreamSynchronize(st);
            cleanup();
            return -1;
        }
    }
    hipError_t e = hipStreamSynchronize(st);
    cleanup();
    *steps_out = steps;
    return static_cast<int>(e);
}

int gen_decode_mix(GenMix &m, int *h_ctrl, hipEvent_t *evs, hipStream_t st, int64_t *steps_out) {
    *steps_out = 0;
    if (m.n < 1 || m.n > kMaxMixCodes) return -2;
    int groups = 0, vblocks = 0, cblocks = 0, post = 0;
    int64_t limit = 0;
    constexpr int kBatch = 4;
    for (int c = 0; c < m.n; ++c) {
        GenCode &k = m.c[c];
        if (k.count <= 0 || k.s.groups <= 0) return -2;
        k.group0 = groups;
        groups += k.s.groups;
        k.vx = (k.g.n + 4 * 16 - 1) / (4 * 16);
        k.var_b0 = vblocks;
        vblocks += k.vx * k.s.groups;
        k.cvar = mix_check_variant(k.g, k.cpw);
        k.nch0 = (k.g.side[0].m + k.cpw - 1) / k.cpw;
        k.nch = k.nch0 + (k.g.side[1].m + k.cpw - 1) / k.cpw;
        k.chk_b0 = cblocks;
        cblocks += (k.nch + 3) / 4 * k.s.groups;
        k.lx = (std::max(k.g.side[0].m, k.g.side[1].m) + 255) / 256;
        k.ex = ((k.g.n + 63) / 64 + 3) / 4;
        k.post0 = post;
        post += k.ex * kEmitRows + k.lx * k.s.groups;
        k.warmup = std::max(0, k.warmup);
        k.interval = std::max(1, k.interval);
        // reset: every slot empty, counters zero (fill kernels)
        launch_fill_i32(reinterpret_cast<int *>(k.s.bad), 0, 2 * (kBadWays + 3) * k.s.groups, st);
        launch_fill_i32(k.s.ctrl, 0, 16, st);
        hipLaunchKernelGGL(gen_zero_kernel, dim3(k.s.groups), dim3(64), 0, st, k.g, k.s);
        // each trial occupies its slot for at most max_iter + 1 steps
        const int64_t waves = (k.count + 64LL * k.s.groups - 1) / (64LL * k.s.groups);
        limit = std::max(limit, (waves + 1) * (static_cast<int64_t>(k.max_iter) + 1) + 4 * kBatch);
    }
    hipLaunchKernelGGL(gen_book_mix_kernel, dim3(groups), dim3(64), 0, st, m);
    hipLaunchKernelGGL(gen_post_mix_kernel, dim3(post), dim3(256), 0, st, m);
    // the instantiations for this pair of codes (a lone code fills both)
    const GenCode &k0 = m.c[0], &k1 = m.c[m.n > 1 ? 1 : 0];
    auto wsel = [](int w) { return w == 4 ? 0 : (w == 8 ? 1 : 2); };
    using Kern = void (*)(GenMix);
    static const Kern var_tab[3][3] = {
        {gen_var_mix_kernel<4, 4>, gen_var_mix_kernel<4, 8>, gen_var_mix_kernel<4, 16>},
        {gen_var_mix_kernel<8, 4>, gen_var_mix_kernel<8, 8>, gen_var_mix_kernel<8, 16>},
        {gen_var_mix_kernel<16, 4>, gen_var_mix_kernel<16, 8>, gen_var_mix_kernel<16, 16>}};
    static const Kern chk_tab[3][3] = {
        {gen_check_mix_kernel<kMixChk7, kMixChk7>, gen_check_mix_kernel<kMixChk7, kMixChk12>,
         gen_check_mix_kernel<kMixChk7, kMixChkAny>},
        {gen_check_mix_kernel<kMixChk12, kMixChk7>, gen_check_mix_kernel<kMixChk12, kMixChk12>,
         gen_check_mix_kernel<kMixChk12, kMixChkAny>},
        {gen_check_mix_kernel<kMixChkAny, kMixChk7>, gen_check_mix_kernel<kMixChkAny, kMixChk12>,
         gen_check_mix_kernel<kMixChkAny, kMixChkAny>}};
    const Kern var_k = var_tab[wsel(k0.g.vpad_w)][wsel(k1.g.vpad_w)];
    const Kern chk_k = chk_tab[k0.cvar][k1.cvar];
    int64_t steps = 0;
    int pend = -1;
    for (;;) {
        steps += kBatch;
        for (int b = 0; b < kBatch; ++b) {
            hipLaunchKernelGGL(var_k, dim3(vblocks), dim3(256), 0, st, m);
            hipLaunchKernelGGL(chk_k, dim3(cblocks), dim3(256), 0, st, m);
            hipLaunchKernelGGL(gen_book_mix_kernel, dim3(groups), dim3(64), 0, st, m);
            hipLaunchKernelGGL(gen_post_mix_kernel, dim3(post), dim3(256), 0, st, m);
        }
        hipError_t e = hipGetLastError();
        // each code's finished count after this batch, read back one batch
        // later (two copies and events alternate)
        const int slot = static_cast<int>((steps / kBatch) & 1);
        for (int c = 0; c < m.n && e == hipSuccess; ++c)
            e = hipMemcpyAsync(h_ctrl + slot * kMaxMixCodes + c, m.c[c].s.ctrl + 1, sizeof(int),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(evs[slot], st);
        if (e != hipSuccess) return static_cast<int>(e);
        if (pend >= 0) {
            e = hipEventSynchronize(evs[pend]);
            if (e != hipSuccess) return static_cast<int>(e);
            bool all = true;
            for (int c = 0; c < m.n; ++c) all = all && h_ctrl[pend * kMaxMixCodes + c] >= m.c[c].count;
            if (all) break;
        }
        pend = slot;
        if (steps > limit) {
            (void)hipStreamSynchronize(st);
            return -1;
        }
    }
    *steps_out = steps;
    return static_cast<int>(hipStreamSynchronize(st));
}

}  // namespace jbp
