// FP32 production decoder for codes built from affine permutation blocks
// (the params-file codes of jointbp_ets.cpp:98-160, 628-648): check r*P + x
// meets variable cc*P + (a[r][cc]*x + b[r][cc] mod P) for every column block.
//
// One workgroup of P threads owns one trial at a time and keeps its whole
// message state on chip; it pulls the next trial from a device counter when
// the current one stops (persistent kernel, refill on completion, so trials
// of uneven length never idle a compute unit).
//
// Message representation.  The reference's X-side check message is
// {a, b, a, b} and an X check reads only q0 = m0+m2, q1 = m1+m3 of a
// variable message (jointbp_ets.cpp:2378-2408); likewise {a, a, b, b} and
// m0+m1 / m2+m3 on the Z side.  Each directed edge message is therefore one
// binary log-likelihood ratio, kept here in base 2 (bits):
//   c2v   lambda = log2(P(bit=0) / P(bit=1)) from the check
//   v2c   mu     = log2(q0 / q1)
// The joint variable update (jointbp_ets.cpp:2451-2519) becomes, with
// Lx = sum_r lambda_x[r], Lz = sum_r lambda_z[r] and prior pi over (I,X,Z,Y):
//   est      = argmax(log2 pi_I + Lx + Lz, log2 pi_X + Lz, log2 pi_Z + Lx, log2 pi_Y)
//   mu_x[r]  = Lx - lambda_x[r] + log2((pi_I 2^Lz + pi_Z) / (pi_X 2^Lz + pi_Y))
//   mu_z[r]  = Lz - lambda_z[r] + log2((pi_I 2^Lx + pi_X) / (pi_Z 2^Lx + pi_Y))
// and the check update (jointbp_ets.cpp:2373-2449) is the reference's
// even/odd parity recursion run on pairs (1, 2^-|mu|), with the signs and
// the syndrome bit applied to the result.
//
// On-chip layout.  Thread t is the owner of the variables (cc, y = t) for
// all column blocks and keeps their incoming check messages lambda[side][r][cc]
// in registers.  Thread t is also the owner of the checks (r, x = t).  The
// two views meet in one LDS exchange buffer B[r][cc][x] (J*L*P floats): the
// variable owner writes mu to B at x = ainv*(y - b) mod P, the check owner
// reads its row B[r][*][x], writes lambda back in place, and the variable
// owner reads it from the same slot.  Because P is a multiple of 64 and the
// block multiplier is odd, a wavefront's 64 permuted addresses fall in 64
// distinct LDS banks.
//
// Syndrome test for free: the variable owner stores the hard-decision bit of
// the edge's variable in the least-significant mantissa bit of mu (a 2^-23
// relative perturbation, below FP32 rounding of the update itself), so each
// check owner recovers the parity of est over its row while reading its
// inputs.  The test of iteration k's estimate therefore completes during the
// check pass of iteration k+1; a trial that converged at k keeps est_k.
#include "jbp_bp32.h"
#include "jbp_common.h"
#include "jbp_learned.h"

namespace jbp {
namespace {

typedef __attribute__((address_space(3))) float lds_float;

struct QcDev {
    const int *ainv;  // [2][J][L]
    const int *b;     // [2][J][L]
    int P;
};

__device__ __forceinline__ float with_lsb(float v, uint32_t bit) {
    return __uint_as_float((__float_as_uint(v) & ~1u) | bit);
}

// Check update of one row of L inputs (stride P floats apart), in place.
// One XOR over the inputs' bit patterns gives both parities at once: its
// sign bit is the parity of negative inputs and its least-significant bit
// that of the estimate bits the variable owners stored there.  The forward
// pass parks e_i = 2^-|mu_i| in the row slot it read, carrying mu_i's sign,
// so the backward pass needs no register copy of either.
template <int L, int P>
__device__ __forceinline__ uint32_t check_row(float *row, uint32_t syn) {
    uint32_t acc = 0;
    float pe[L + 1], po[L + 1];
    pe[0] = 1.0f;
    po[0] = 0.0f;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        float mu = row[i * P];
        uint32_t u = __float_as_uint(mu);
        acc ^= u;
        float e = fast_exp2(-fabsf(mu));
        row[i * P] = __uint_as_float((__float_as_uint(e) & 0x7fffffffu) | (u & 0x80000000u));
        if (i == 0) {  // (1, 0) then (1, e): no arithmetic
            pe[1] = 1.0f;
            po[1] = e;
        } else {
            pe[i + 1] = pe[i] + po[i] * e;
            po[i + 1] = pe[i] * e + po[i];
        }
    }
    uint32_t sall = acc ^ (syn << 31);  // sign of the product of the others, times syn
    asm volatile("" : "+v"(sall));      // one register, so each output is one xor and one insert
    float se = 1.0f, so = 0.0f;
#pragma unroll
    for (int i = L - 1; i >= 0; --i) {
        const float es = row[i * P];
        const float e = fabsf(es);
        // the suffix starts at (1, 0): the last input's products need no arithmetic
        float PE = i == L - 1 ? pe[i] : pe[i] * se + po[i] * so;
        float PO = i == L - 1 ? po[i] : pe[i] * so + po[i] * se;
        // PE >= PO (both recursions keep even >= odd), so mag >= 0 and its
        // sign bit is free for the output sign
        float mag = fminf(fast_log2(PE) - fast_log2(PO), kLlrMax);
        row[i * P] = __uint_as_float(__float_as_uint(mag) |
                                     ((sall ^ __float_as_uint(es)) & 0x80000000u));
        if (i == L - 1) {
            se = 1.0f;
            so = e;
        } else {
            float ne = se + so * e;
            float no = se * e + so;
            se = ne;
            so = no;
        }
    }
    return acc & 1u;
}

// Damping of the learned decoder, d = sigmoid(W2 . tanh(W1 (x0, x1) + b1) + b2),
// evaluated by the first wavefront: lane i < 16 forms hidden unit i, and an
// xor-shuffle tree sums the 16 products.  Every lane of wave 0 returns d.
__device__ __forceinline__ float mlp_damping_wave(const float *mlp, float x0, float x1, int lane) {
    constexpr float kLog2e = 1.4426950408889634f;
    float v = 0.0f;
    if (lane < kMlpHidden) {
        // tanh(a) = 1 - 2 / (e^{2a} + 1), with a clamped where tanh is +-1 in FP32
        float a = fminf(fmaxf(mlp[2 * lane] * x0 + mlp[2 * lane + 1] * x1 + mlp[2 * kMlpHidden + lane],
                              -15.0f), 15.0f);
        float h = 1.0f - 2.0f * __builtin_amdgcn_rcpf(fast_exp2(2.0f * kLog2e * a) + 1.0f);
        v = mlp[3 * kMlpHidden + lane] * h;
    }
#pragma unroll
    for (int m = kMlpHidden / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    const float z = fminf(fmaxf(v + mlp[4 * kMlpHidden], -30.0f), 30.0f);
    return __builtin_amdgcn_rcpf(1.0f + fast_exp2(-kLog2e * z));
}

// One decoding phase of a trial on the workgroup: flooding iterations from
// check messages 0 until both syndromes match or kmax iterations have run.
// WEIGHTED applies the learned weights and damping of iteration k+1 (the
// training kernels' decoder, in base 2).  Returns the iteration index at
// which it stopped; `matched` says whether that estimate satisfied both
// syndromes; estx / estz hold its hard decisions.  DAMPED (with WEIGHTED)
// adds the MLP damping, which keeps the previous check messages live across
// each check pass; without it the weighted phase has the register footprint
// of the plain one.
template <int P, int J, int L, bool WEIGHTED, bool DAMPED>
__device__ __forceinline__ int qc_phase(float *B, float *GZ, int *s_cnt, float *s_d, float *s_w,
                                        uint32_t (&aw)[2][J][(L + 1) / 2], float (&lam)[2][J][L],
                                        uint32_t synx, uint32_t synz, const PriorF &pr, int kmax,
                                        int warmup, int interval, const LearnedInfer &li, float u0,
                                        uint32_t &estx, uint32_t &estz, int &matched) {
    constexpr int W = (L + 1) / 2;
    constexpr int HB = J * L / 2;  // blocks in the lower half of B
    const int t = threadIdx.x;
    const int lane = t & 63;
    // Make the packed addresses opaque to the optimiser at chosen points, so
    // that unpacked values are recomputed instead of being kept live (and
    // spilled) across the check passes or the iteration loop.
    auto launder = [&]() {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < J; ++r)
#pragma unroll
                for (int w = 0; w < W; ++w) asm volatile("" : "+v"(aw[s][r][w]));
    };
    // Exchange slot of variable (cc, t) in block (r, cc) of side s: a 16-bit
    // LDS address (B's base + x*4, plus the lower half's size for blocks >=
    // HB) and the block's position within its half, a constant that the LDS
    // access takes as its immediate offset.
    auto slot = [&](int s, int r, int cc) -> lds_float & {
        const uint32_t a16 = (cc & 1) ? (aw[s][r][cc >> 1] >> 16) : (aw[s][r][cc >> 1] & 0xffffu);
        const int blk = r * L + cc;
        return *reinterpret_cast<lds_float *>(
            static_cast<uintptr_t>(a16 + static_cast<uint32_t>((blk % HB) * P * 4)));
    };
    // Iteration-major weights (for iteration k: wc X | wc Z | wv X | wv Z, J*L
    // each) are staged in an LDS double buffer s_w[2][4*J*L]: iteration k reads
    // half k&1 while the next iteration's weights are loaded into the other.
    constexpr int NW = 4 * J * L;
    auto stage = [&](int kk) {
        if (t < NW && kk < kmax) s_w[(kk & 1) * NW + t] = li.params[kk * NW + t];
    };
    if (WEIGHTED) stage(0);
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc) lam[s][r][cc] = 0.0f;
    float ukm1 = u0;  // unsatisfied fraction of est_{k-1} (est_0 for k <= 1)
    int k = 0;
    __syncthreads();
    for (;;) {
        launder();
        const bool live = !WEIGHTED || k < kmax;  // weights exist for iteration k+1
        const float *wk = s_w + (k & 1) * NW;
        if (WEIGHTED) stage(k + 1);
        // Variable pass k: estimate, extrinsic X messages (+ est bit) to B.
        estx = 0;
        estz = 0;
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
            float lx = 0.0f, lz = 0.0f;
#pragma unroll
            for (int r = 0; r < J; ++r) {
                lx += lam[0][r][cc];
                lz += lam[1][r][cc];
            }
            const int best = hard_decision(lx, lz, pr);
            estx |= static_cast<uint32_t>(best & 1) << cc;
            estz |= static_cast<uint32_t>(best >> 1) << cc;
            float gx = lx + extrinsic_term(lz, pr.fx);
            GZ[cc * P + t] = lz + extrinsic_term(lx, pr.fz);
            uint32_t bx = static_cast<uint32_t>(best & 1);
#pragma unroll
            for (int r = 0; r < J; ++r) {
                float mu = gx - lam[0][r][cc];
                if (WEIGHTED && live) mu *= wk[2 * J * L + r * L + cc];
                slot(0, r, cc) = with_lsb(mu, bx);
            }
        }
        if (WEIGHTED && t < 64) {
            float d = 0.0f;
            if (DAMPED && live)
                d = mlp_damping_wave(li.params + 4 * J * L * li.T,
                                     static_cast<float>(k + 1) / static_cast<float>(li.T),
                                     ukm1, lane);
            if (t == 0) {
                *s_cnt = 0;
                *s_d = d;
            }
        }
        // One register per mask: without this the per-column bits are
        // forwarded to their later uses individually and spill.
        asm volatile("" : "+v"(estx), "+v"(estz));
        __syncthreads();
        // X check pass: rows (r, t).
        uint32_t bad = 0;
#pragma unroll 1
        for (int r = 0; r < J; ++r) {
            uint32_t sb = (synx >> r) & 1u;
            bad |= (check_row<L, P>(B + r * L * P + t, sb) ^ sb) << r;
        }
        if (WEIGHTED) {
            uint32_t c = 0;
#pragma unroll
            for (int r = 0; r < J; ++r) c += __popcll(__ballot((bad >> r) & 1u));
            if (lane == 0 && c) atomicAdd(s_cnt, static_cast<int>(c));
        }
        const int badx = __syncthreads_or(static_cast<int>(bad));
        launder();
        // learned: lambda = (1 - d) * wc * CN + d * lambda(t-1); CN was clamped
        // in check_row, before the weight, as in training
        const float dmp = DAMPED ? *s_d : 0.0f;
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                float v = slot(0, r, cc);
                if (WEIGHTED && live) v *= wk[r * L + cc];
                lam[0][r][cc] = DAMPED ? (1.0f - dmp) * v + dmp * lam[0][r][cc] : v;
            }
            // damping keeps the old messages live: bound the loads in flight
            // to one row so they do not spill
            if (DAMPED) __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
            uint32_t bz = (estz >> cc) & 1u;
            const float gz = GZ[cc * P + t];
#pragma unroll
            for (int r = 0; r < J; ++r) {
                float mu = gz - lam[1][r][cc];
                if (WEIGHTED && live) mu *= wk[3 * J * L + r * L + cc];
                slot(1, r, cc) = with_lsb(mu, bz);
            }
        }
        __syncthreads();
        bad = 0;
#pragma unroll 1
        for (int r = 0; r < J; ++r) {
            uint32_t sb = (synz >> r) & 1u;
            bad |= (check_row<L, P>(B + r * L * P + t, sb) ^ sb) << r;
        }
        if (WEIGHTED) {
            uint32_t c = 0;
#pragma unroll
            for (int r = 0; r < J; ++r) c += __popcll(__ballot((bad >> r) & 1u));
            if (lane == 0 && c) atomicAdd(s_cnt, static_cast<int>(c));
        }
        const int badz = __syncthreads_or(static_cast<int>(bad));
        if (k >= 1) {
            bool test = WEIGHTED || (k == kmax) || (k > warmup && ((k - warmup) % interval) == 0);
            if (test && !badx && !badz) {
                matched = 1;
                return k;
            }
        }
        if (k >= kmax) {
            matched = 0;
            return k;
        }
        if (WEIGHTED) ukm1 = static_cast<float>(*s_cnt) / static_cast<float>(2 * J * P);
        launder();
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                float v = slot(1, r, cc);
                if (WEIGHTED && live) v *= wk[J * L + r * L + cc];
                lam[1][r][cc] = DAMPED ? (1.0f - dmp) * v + dmp * lam[1][r][cc] : v;
            }
            // damping keeps the old messages live: bound the loads in flight
            // to one row so they do not spill
            if (DAMPED) __builtin_amdgcn_sched_barrier(0);
        ++k;
        __syncthreads();
    }
}


// ---------------------------------------------------------------------------
// Learned min-sum phase (the fast learned decoder).
//
// Normalised min-sum with the max-log joint variable node (the max-product
// form of the updates above), learned weights per (side, block, iteration)
// on both message directions and optional MLP damping; the training kernels
// (jbp_learned.hip, kind 1) define it.  In base 2 (all updates are
// homogeneous, so the weights are the same in any base):
//   S      = sum_r wc[r] * lambda[r]                (per side)
//   F      = max(lp_I + S', lp_Z) - max(lp_X + S', lp_Y)   (X side; S' = the
//            other side's S; Z side with X and Z swapped), evaluated as the
//            clamp it equals (ms_marginal)
//   est    = sign bits of G = S + F
//   mu[r]  = wv[r] * (G - wc[r] * lambda[r])
//   lambda = (-1)^syn prod_{k != e} sign(mu_k) * min(64, min_{k != e} |mu_k|)
// X and Z travel together as FP16 pairs: word B[r][cc][x] holds, in its low
// half, the message of X check (r, x) at position cc and, in its high half,
// that of Z check (r, x), so a check owner updates its X and Z checks with
// the same packed instructions, and a variable owner keeps its 36 incoming
// (X, Z) pairs in 36 registers.  The estimate bits ride in the least
// significant bit of each half of mu (bit 0 and bit 16 of the word).
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) _Float16 lds_h16;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ h2 as_h2(uint32_t u) { return __builtin_bit_cast(h2, u); }
__device__ __forceinline__ uint32_t h2_bits(h2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as_u16x2(uint32_t u) { return __builtin_bit_cast(u16x2, u); }
__device__ __forceinline__ uint32_t u16x2_bits(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
// magnitudes (sign bits clear) compare as unsigned integers
__device__ __forceinline__ u16x2 umin2(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }

constexpr uint32_t kMsClampPair = 0x54005400u;  // (64.0, 64.0) in FP16

// The channel term F(L) = max(a + L, b) - max(c + L, d) of one side (a..d
// its log-prior terms) is a clamp of a line of slope +-1: when
// b - a <= d - c it is min(max(L + (a - d), b - d), a - c), otherwise
// min(max((b - c) - L, a - c), b - d).  Per half of an (X, Z) pair: the
// slope s, offset k and bounds lo <= hi, so the marginal G = S + F(S') is
// four packed operations (S' = S with its halves swapped).
struct MsChan {
    h2 s, k, lo, hi;
};
// one half: (s, k, lo, hi)
__device__ __forceinline__ void ms_chan_half(float a, float b, float c, float d, float (&v)[4]) {
    const bool rising = b - a <= d - c;
    v[0] = rising ? 1.0f : -1.0f;
    v[1] = rising ? a - d : b - c;
    v[2] = rising ? b - d : a - c;
    v[3] = rising ? a - c : b - d;
}
__device__ __forceinline__ MsChan ms_chan(const PriorF &pr) {
    // X half: the Z-side sum enters with (a, b, c, d) = (I, Z, X, Y) terms;
    // Z half: (I, X, Z, Y)
    float x[4], z[4];
    ms_chan_half(pr.lp[0], pr.lp[2], pr.lp[1], pr.lp[3], x);
    ms_chan_half(pr.lp[0], pr.lp[1], pr.lp[2], pr.lp[3], z);
    MsChan m;
    m.s = h2{(_Float16)x[0], (_Float16)z[0]};
    m.k = h2{(_Float16)x[1], (_Float16)z[1]};
    m.lo = h2{(_Float16)x[2], (_Float16)z[2]};
    m.hi = h2{(_Float16)x[3], (_Float16)z[3]};
    return m;
}
__device__ __forceinline__ h2 ms_marginal(h2 S, const MsChan &m) {
    const h2 Sw = __builtin_shufflevector(S, S, 1, 0);
    const h2 F = __builtin_elementwise_min(
        __builtin_elementwise_max(__builtin_elementwise_fma(Sw, m.s, m.k), m.lo), m.hi);
    return S + F;
}

template <int P, int J, int L, bool DAMPED>
__device__ __forceinline__ int qc_phase_ms(uint32_t *Bw, int *s_cnt, float *s_d, uint32_t *s_w,
                                           uint32_t (&aw)[2][J][(L + 1) / 2], uint32_t synx,
                                           uint32_t synz, const PriorF &pr, int kmax,
                                           const LearnedInfer &li, float u0, uint32_t &estx,
                                           uint32_t &estz, int &matched) {
    constexpr int HB = J * L / 2;
    const int t = threadIdx.x;
    const int lane = t & 63;
    // Byte address of the word holding variable (cc, t)'s message in block
    // (r, cc) of side s (its Z half is 2 bytes further on), without the
    // block's offset inside its half of B, which the LDS instructions take as
    // their immediate.  Unpacked once per trial into 72 registers; the packed
    // copy is rebuilt from them on exit, so it holds no registers meanwhile.
    uint32_t ad[2][J][L];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc)
                ad[s][r][cc] = (cc & 1) ? (aw[s][r][cc >> 1] >> 16) : (aw[s][r][cc >> 1] & 0xffffu);
    auto repack = [&]() {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < J; ++r)
#pragma unroll
                for (int w = 0; w < (L + 1) / 2; ++w)
                    aw[s][r][w] = ad[s][r][2 * w] | (2 * w + 1 < L ? ad[s][r][2 * w + 1] << 16 : 0u);
    };
    auto at = [&](int s, int r, int cc) -> uint32_t {
        return ad[s][r][cc] + static_cast<uint32_t>(((r * L + cc) % HB) * P * 4) + (s ? 2u : 0u);
    };
    auto ldh = [](uint32_t a) -> _Float16 { return *reinterpret_cast<lds_h16 *>(static_cast<uintptr_t>(a)); };
    auto sth = [](uint32_t a, _Float16 v) { *reinterpret_cast<lds_h16 *>(static_cast<uintptr_t>(a)) = v; };
    auto stage = [&](int kk) {
        if (t < kMsRec && kk <= kmax) s_w[(kk & 1) * kMsRec + t] = li.ms_rec[kk * kMsRec + t];
    };
    const MsChan ch = ms_chan(pr);  // channel term, (X side, Z side) pairs
    // damped: the previous check messages, mixed into the new ones
    h2 lam[DAMPED ? J : 1][DAMPED ? L : 1];
    if (DAMPED) {
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc) lam[r][cc] = h2{(_Float16)0.0f, (_Float16)0.0f};
    }
    stage(0);
    if (DAMPED && t == 0) {
        s_cnt[0] = 0;
        s_cnt[1] = 0;
    }
    // lambda_0 = 0: clear the exchange words the first variable pass reads
    {
        lds_u32 *bw = reinterpret_cast<lds_u32 *>(reinterpret_cast<uintptr_t>(Bw));
#pragma unroll
        for (int i = 0; i < J * L; ++i) bw[i * P + t] = 0u;
    }
    float ukm1 = u0;
    float dprev = 0.0f;  // damping of the check messages being read (iteration k)
    int k = 0;
    __syncthreads();
    for (;;) {
        stage(k + 1);
        const uint32_t *rec = s_w + (k & 1) * kMsRec;
        // Variable pass k: read lambda_k (left in place by the check pass;
        // zeros before the first), form the estimate bits and mu, write mu
        // back to the same halves.
        uint32_t acc_est = 0;
        // software-pipelined over column blocks: the next block's check
        // messages are in flight while this block's arithmetic runs
        h2 nxt[J];
#pragma unroll
        for (int r = 0; r < J; ++r) nxt[r] = h2{ldh(at(0, r, 0)), ldh(at(1, r, 0))};
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
            h2 lv[J];
#pragma unroll
            for (int r = 0; r < J; ++r) lv[r] = nxt[r];
            if (cc + 1 < L) {
#pragma unroll
                for (int r = 0; r < J; ++r) nxt[r] = h2{ldh(at(0, r, cc + 1)), ldh(at(1, r, cc + 1))};
            }
            const uint4 w0 = *reinterpret_cast<const uint4 *>(rec + cc * kMsRecCC);
            const uint4 w1 = *reinterpret_cast<const uint4 *>(rec + cc * kMsRecCC + 4);
            const h2 wc[3] = {as_h2(w0.x), as_h2(w0.y), as_h2(w0.z)};
            const h2 wv[3] = {as_h2(w0.w), as_h2(w1.x), as_h2(w1.y)};
            h2 S;
            if (DAMPED) {
                // lambda_k = (1 - d_k) wc_{k-1} CN + d_k lambda_{k-1}
                const h2 dd = {(_Float16)dprev, (_Float16)dprev};
                const h2 c1 = {(_Float16)(1.0f - dprev), (_Float16)(1.0f - dprev)};
#pragma unroll
                for (int r = 0; r < J; ++r) {
                    lam[r][cc] = __builtin_elementwise_fma(c1 * wc[r], lv[r], dd * lam[r][cc]);
                    lv[r] = lam[r][cc];
                }
                S = lv[0] + lv[1] + lv[2];
            } else {
                S = wc[0] * lv[0];
#pragma unroll
                for (int r = 1; r < J; ++r) S = __builtin_elementwise_fma(wc[r], lv[r], S);
            }
            const h2 G = ms_marginal(S, ch);
            const uint32_t eb = (h2_bits(G) >> 15) & 0x00010001u;
            acc_est |= eb << cc;
#pragma unroll
            for (int r = 0; r < J; ++r) {
                const h2 m = DAMPED ? (G - lv[r]) : __builtin_elementwise_fma(-wc[r], lv[r], G);
                const h2 mb = as_h2((h2_bits(wv[r] * m) & 0xfffefffeu) | eb);
                sth(at(0, r, cc), mb.x);
                sth(at(1, r, cc), mb.y);
            }
        }
        if (DAMPED && t < 64) {
            const float d = mlp_damping_wave(li.params + 4 * J * L * li.T,
                                             static_cast<float>(k + 1) / static_cast<float>(li.T),
                                             ukm1, lane);
            if (t == 0) *s_d = d;
        }
        __syncthreads();
        if (DAMPED && t == 0) s_cnt[(k + 1) & 1] = 0;
        // check pass: X check (r, t) in the low halves, Z check (r, t) in the high
        uint32_t bad = 0;
        int nbad = 0;
#pragma unroll 1
        for (int r = 0; r < J; ++r) {
            lds_u32 *row = reinterpret_cast<lds_u32 *>(reinterpret_cast<uintptr_t>(Bw)) + r * L * P + t;
            const uint32_t synw = ((synx >> r) & 1u) | (((synz >> r) & 1u) << 16);
            uint32_t u[L];
            u16x2 pre[L];
            u16x2 run = as_u16x2(kMsClampPair);
            uint32_t acc = 0;
#pragma unroll
            for (int i = 0; i < L; ++i) {
                u[i] = row[i * P];
                acc ^= u[i];
                pre[i] = run;
                run = umin2(run, as_u16x2(u[i] & 0x7fff7fffu));
            }
            const uint32_t par = (acc ^ synw) & 0x00010001u;  // estimate parity vs syndrome
            bad |= par;
            if (DAMPED) nbad += __popc(par);
            const uint32_t sall = acc ^ (synw << 15);
            u16x2 suf = as_u16x2(kMsClampPair);
#pragma unroll
            for (int i = L - 1; i >= 0; --i) {
                const uint32_t mag = u16x2_bits(umin2(pre[i], suf));
                suf = umin2(suf, as_u16x2(u[i] & 0x7fff7fffu));
                row[i * P] = mag | ((sall ^ u[i]) & 0x80008000u);
            }
        }
        if (DAMPED) {
            // unsatisfied checks of est_k: a wave sum, one LDS atomic per wave
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) nbad += __shfl_xor(nbad, m);
            if (lane == 0 && nbad) atomicAdd(&s_cnt[k & 1], nbad);
        }
        const int any_bad = __syncthreads_or(static_cast<int>(bad));
        estx = acc_est & 0xffffu;
        estz = acc_est >> 16;
        if (k >= 1 && !any_bad) {
            matched = 1;
            repack();
            return k;
        }
        if (k >= kmax) {
            matched = 0;
            repack();
            return k;
        }
        if (DAMPED) {
            ukm1 = static_cast<float>(s_cnt[k & 1]) / static_cast<float>(2 * J * P);
            dprev = *s_d;
        }
        ++k;
    }
}

// Layered learned min-sum phase (kind kLearnedLMS, the training kernels'
// learned_*_lms_kernel in base 2): block row r of both sides is layer r, and
// iteration k updates the layers in turn, each from the messages the layers
// before it left.  Check-centric form: thread t owns the X check (r, t) and
// the Z check (r, t) of every row and keeps their weighted outgoing messages
// lambda (36 (X, Z) FP16 pairs) in registers; the per-variable sums
//   S(v) = (S_x(v), S_z(v)) = sums of v's current weighted check messages
// live in LDS as FP16 pairs, in three buffers: layer r reads buffer r and
// writes buffer r+1 mod 3, so every LDS offset is a constant.  Layer
// r, for the X check's variable v and the Z check's variable v' at each
// position cc (one packed lane each):
//   U = (S_x(v), S_z(v')),  W = (S_z(v), S_x(v'))      (from the read buffer)
//   F = the joint channel term of W (a clamp, LmsChan),  G = U + F
//   mu = wv[k][r][cc] * (U - lambda + F),  lambda' = wc[k][r][cc] * CN(mu)
//   U - lambda + lambda' -> the write buffer's S_x(v) and S_z(v')
// Every variable meets exactly one X and one Z check of a block row, so each
// half of each sum is written once per layer, and one barrier separates
// layers.  (The sums are updated in place rather than re-added from the
// three rows; the difference is FP16 rounding.)
//
// Convergence.  The sign bits of G are the estimate this layer starts from;
// their parity over a check is compared with its syndrome bit.  After three
// consecutive layers whose rows all accept, the estimate of the current sums
// is verified against every row (verify) and, if it holds, the trial stops
// with that estimate; iterations = the number begun.  A trial whose layers
// settle on a valid estimate thus stops within the same iteration as a test
// at iteration boundaries would, without running the remaining layers.
// This lane's index in its wavefront, computed where it is used (volatile,
// so that neither it nor anything derived from it is hoisted into a
// long-lived register: the layered kernel's registers are all taken by its
// messages and slot addresses, and a hoisted per-thread constant ends up
// spilled to scratch and reloaded with a memory round trip).
__device__ __forceinline__ uint32_t lane_now() {
    uint32_t lane;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    return lane;
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS (and
// scalar) operations, not for its global stores, which no thread of the
// workgroup reads back (__syncthreads would wait for them: trial outcomes
// would then stall every trial's end).  The empty asm statements keep the
// compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt not waited for
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// The joint channel term of one half is F(W) = lo + d * clamp01(s*W + k0).
// The sums are stored offset by lo of their own side (S' = S + lo), so
// that G = U' + d*c and mu' = (U' - lambda) + d*c are one FMA each; the
// clamp argument, which reads the other side's offset sums W', uses
// k = k0 - s * (lo of the other side).
struct LmsChan {
    h2 s, k, d, lo;  // per half (X low, Z high)
};
// one half: F(L) = max(a + L, b) - max(c + L, d) = min(max(sl*L + k, lo), hi)
__device__ __forceinline__ void lms_chan_half(float a, float b, float c, float d, float (&v)[4]) {
    a = fmaxf(a, -100.0f);
    b = fmaxf(b, -100.0f);
    c = fmaxf(c, -100.0f);
    d = fmaxf(d, -100.0f);
    const bool rising = b - a <= d - c;
    const float sl = rising ? 1.0f : -1.0f;
    const float k = rising ? a - d : b - c;
    const float lo = rising ? b - d : a - c;
    const float hi = rising ? a - c : b - d;
    const float w = hi - lo;
    const bool flat = !(w > 1e-3f);
    v[0] = flat ? 0.0f : sl / w;
    v[1] = flat ? 0.0f : (k - lo) / w;
    v[2] = flat ? 0.0f : w;
    v[3] = lo;
}
__device__ __forceinline__ LmsChan lms_chan(const PriorF &pr) {
    float x[4], z[4];
    lms_chan_half(pr.lp[0], pr.lp[2], pr.lp[1], pr.lp[3], x);  // X half: (I, Z, X, Y)
    lms_chan_half(pr.lp[0], pr.lp[1], pr.lp[2], pr.lp[3], z);  // Z half: (I, X, Z, Y)
    LmsChan m;
    m.s = h2{(_Float16)x[0], (_Float16)z[0]};
    // the X half's clamp reads Z sums (offset by the Z lo), and vice versa
    m.k = h2{(_Float16)(x[1] - x[0] * z[3]), (_Float16)(z[1] - z[0] * x[3])};
    m.d = h2{(_Float16)x[2], (_Float16)z[2]};
    m.lo = h2{(_Float16)x[3], (_Float16)z[3]};
    return m;
}
// clamp01(s * W' + k) of the packed lanes whose other-side offset sums are W'
__device__ __forceinline__ h2 lms_c(h2 W, const LmsChan &ch) {
    return __builtin_elementwise_min(
        __builtin_elementwise_max(__builtin_elementwise_fma(W, ch.s, ch.k), h2{(_Float16)0.0f, (_Float16)0.0f}),
        h2{(_Float16)1.0f, (_Float16)1.0f});
}
// the marginal G of lanes with own-side offset sums U' and other-side W'
__device__ __forceinline__ h2 lms_G(h2 U, h2 W, const LmsChan &ch) {
    return __builtin_elementwise_fma(lms_c(W, ch), ch.d, U);
}

typedef __attribute__((address_space(4))) const uint32_t cu32;

// LDS layout of the sums: variable (cc, y) of buffer bf is word
// cc*3P + bf*P + y.  A thread's packed slot address (16 bits) holds the LDS
// base of the sums plus 4*y plus lms_split(cc), so that the rest of the
// offset, lms_off(bf, cc), fits the LDS instruction's 16-bit immediate.
template <int P>
__device__ __forceinline__ constexpr uint32_t lms_split(int cc) {
    return cc * 3 * P * 4 > 45056 ? static_cast<uint32_t>(cc * 3 * P * 4 - 45056) : 0u;
}
template <int P>
__device__ __forceinline__ constexpr uint32_t lms_off(int bf, int cc) {
    return static_cast<uint32_t>((cc * 3 * P + bf * P) * 4) - lms_split<P>(cc);
}

template <int P, int J, int L>
__device__ __forceinline__ int qc_phase_lms(uint32_t sb, int t0, uint32_t (&aw)[J][L], uint32_t synx,
                                            uint32_t synz, const PriorF &pr, int kmax, const cu32 *rec,
                                            int *s_flag, uint32_t &estx, uint32_t &estz, int &matched) {
    static_assert(lms_off<P>(2, L - 1) + 2 < 65536 && (P - 1) * 4 + lms_split<P>(L - 1) + 1024 < 65536,
                  "sum-buffer offsets must fit the LDS immediate and 16-bit slot addresses");
    // t0: the wavefront's first thread (uniform); thread t0 + lane_now()
    constexpr uint32_t kSign = 0x80008000u, kMag = 0x7fff7fffu;
    auto lds32 = [](uint32_t a) -> lds_u32 * { return reinterpret_cast<lds_u32 *>(static_cast<uintptr_t>(a)); };
    auto sth = [](uint32_t a, _Float16 v) { *reinterpret_cast<lds_h16 *>(static_cast<uintptr_t>(a)) = v; };
    auto launder = [&](int r) {
#pragma unroll
        for (int cc = 0; cc < L; ++cc) asm volatile("" : "+v"(aw[r][cc]));
    };
    const LmsChan ch = lms_chan(pr);
    // syndrome bits of the checks (r, t): X at bit r, Z at bit 16 + r; row
    // r's pair moved to the sign positions (bits 15 and 31) where used
    const uint32_t syn2 = synx | (synz << 16);
    auto synw = [&](int r) { return (syn2 << (15 - r)) & kSign; };
    h2 lam[J][L];
#pragma unroll
    for (int r = 0; r < J; ++r)
#pragma unroll
        for (int cc = 0; cc < L; ++cc) lam[r][cc] = h2{(_Float16)0.0f, (_Float16)0.0f};
#pragma unroll
    for (int cc = 0; cc < L; ++cc) *lds32(sb + (cc * 3 * P + t0 + lane_now()) * 4) = h2_bits(ch.lo);  // S' = lo
    lds_barrier();

    // parity of the estimate of buffer bf's sums over every row, vs the syndrome
    auto verify = [&](int bf) -> uint32_t {
        uint32_t vb = 0;
#pragma unroll
        for (int r = 0; r < J; ++r) {
            launder(r);
            uint32_t wx[L], wz[L];
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                wx[cc] = *lds32((aw[r][cc] & 0xffffu) + lms_off<P>(bf, cc));
                wz[cc] = *lds32((aw[r][cc] >> 16) + lms_off<P>(bf, cc));
            }
            uint32_t acc = 0;
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                const h2 U = as_h2((wx[cc] & 0xffffu) | (wz[cc] & 0xffff0000u));
                const h2 W = as_h2(__builtin_amdgcn_alignbit(wz[cc], wx[cc], 16));
                acc ^= h2_bits(lms_G(U, W, ch));
            }
            vb |= (acc & kSign) ^ synw(r);
        }
        return vb;
    };

    int nv = 0;
    uint32_t hist = 7;  // bit i: a row rejected the estimate of the layer i steps back
    // weights of layer (k, r): (wc X, wc Z) per column block at w[cc], (wv X,
    // wv Z) at w[12 + cc]; loaded one layer ahead into scalar registers
    uint32_t wnext[2 * L];
#pragma unroll
    for (int i = 0; i < 2 * L; ++i) wnext[i] = rec[i];
    for (int k = 0; k < kmax; ++k) {
#pragma unroll
        for (int r = 0; r < J; ++r) {
            const int j = k * J + r;
            uint32_t wcur[2 * L];
#pragma unroll
            for (int i = 0; i < 2 * L; ++i) wcur[i] = wnext[i];
            launder(r);
            uint32_t wx[L], wz[L];
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                wx[cc] = *lds32((aw[r][cc] & 0xffffu) + lms_off<P>(r, cc));
                wz[cc] = *lds32((aw[r][cc] >> 16) + lms_off<P>(r, cc));
            }
            h2 E[L], mu[L];
            u16x2 pre[L];
            u16x2 run = as_u16x2(kMsClampPair);
            // acc: sign parity of G (the estimate), sg: of mu; folded two
            // positions at a time (three-input XORs)
            uint32_t acc = 0, sg = 0, gprev = 0, mprev = 0;
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                const h2 U = as_h2((wx[cc] & 0xffffu) | (wz[cc] & 0xffff0000u));
                const h2 W = as_h2(__builtin_amdgcn_alignbit(wz[cc], wx[cc], 16));
                const h2 c = lms_c(W, ch);
                E[cc] = U - lam[r][cc];
                const uint32_t gb = h2_bits(__builtin_elementwise_fma(c, ch.d, U));
                mu[cc] = as_h2(wcur[L + cc]) * __builtin_elementwise_fma(c, ch.d, E[cc]);
                const uint32_t mb = h2_bits(mu[cc]);
                if (cc & 1) {
                    acc = __builtin_amdgcn_bitop3_b32(acc, gprev, gb, 0x96);  // a ^ b ^ c
                    sg = __builtin_amdgcn_bitop3_b32(sg, mprev, mb, 0x96);
                } else {
                    gprev = gb;
                    mprev = mb;
                }
                pre[cc] = run;
                run = umin2(run, as_u16x2(h2_bits(mu[cc]) & kMag));
            }
            {
                // the next layer's record (issued after this layer's reads;
                // record T, past the last layer, exists)
                const cu32 *wn = rec + (j + 1) * 32;
#pragma unroll
                for (int i = 0; i < 2 * L; ++i) wnext[i] = wn[i];
            }
            const uint32_t bad = (acc & kSign) ^ synw(r);
            const uint32_t sall = (sg & kSign) ^ synw(r);
            const int wb = (r + 1) % J;  // the buffer this layer writes
            u16x2 suf = as_u16x2(kMsClampPair);
#pragma unroll
            for (int cc = L - 1; cc >= 0; --cc) {
                const uint32_t m = h2_bits(mu[cc]) & kMag;
                const uint32_t mag = u16x2_bits(umin2(pre[cc], suf));
                suf = umin2(suf, as_u16x2(m));
                const h2 p = as_h2(wcur[cc]) * as_h2(mag);
                const h2 lnew = as_h2(h2_bits(p) ^ ((h2_bits(mu[cc]) ^ m) ^ sall));
                const h2 S = E[cc] + lnew;
                lam[r][cc] = lnew;
                sth((aw[r][cc] & 0xffffu) + lms_off<P>(wb, cc), S.x);
                sth((aw[r][cc] >> 16) + lms_off<P>(wb, cc) + 2, S.y);
            }
            if (__ballot(bad != 0) != 0 && lane_now() == 0) s_flag[j & 3] = 1;
            if (t0 == 0 && lane_now() == 0) s_flag[(j + 2) & 3] = 0;
            lds_barrier();
            const int f = __builtin_amdgcn_readfirstlane(s_flag[j & 3]);
            hist = (hist << 1) | (f ? 1u : 0u);
            if ((hist & 7u) == 0) {
                const uint32_t vb = verify(wb);
                const int vi = 4 + (nv & 1);
                if (t0 == 0 && lane_now() == 0) s_flag[4 + ((nv + 1) & 1)] = 0;
                if (__ballot(vb != 0) != 0 && lane_now() == 0) s_flag[vi] = 1;
                lds_barrier();
                const int vf = __builtin_amdgcn_readfirstlane(s_flag[vi]);
                ++nv;
                if (!vf) {
                    // the estimate of the variables (cc, t) from the same sums
                    estx = 0;
                    estz = 0;
                    const uint32_t own = sb + (wb * P + t0 + lane_now()) * 4;
#pragma unroll
                    for (int cc = 0; cc < L; ++cc) {
                        const h2 S = as_h2(*lds32(own + cc * 3 * P * 4));
                        const uint32_t g = h2_bits(lms_G(S, __builtin_shufflevector(S, S, 1, 0), ch));
                        estx |= ((g >> 15) & 1u) << cc;
                        estz |= (g >> 31) << cc;
                    }
                    matched = 1;
                    return k + 1;
                }
                hist = 1;
            }
        }
    }
    matched = 0;
    return kmax;
}

// Exchange-slot addresses of thread t's variables (cc, t): row index
// x = ainv*(t - b) mod P in block (side, r, cc), whose slot is
// B[(r*L + cc)*P + x], kept as a 16-bit LDS address (two per register: see
// qc_phase); blocks of the upper half of B carry the half's offset.
template <int P, int J, int L>
__device__ __forceinline__ void qc_slot_addresses(const QcDev &qc, const float *B,
                                                  uint32_t (&aw)[2][J][(L + 1) / 2]) {
    constexpr int HB = J * L / 2;
    static_assert((P - 1) * 4 + HB * P * 4 + 4096 < 65536 && (HB - 1) * P * 4 < 65536,
                  "exchange addresses must fit 16 bits and the LDS immediate offset");
    const int t = threadIdx.x;
    const uint32_t bbase = static_cast<uint32_t>((uintptr_t)((const lds_float *)B));  // LDS address of B
    constexpr int W = (L + 1) / 2;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int w = 0; w < W; ++w) aw[s][r][w] = 0;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int r = 0; r < J; ++r)
#pragma unroll
            for (int cc = 0; cc < L; ++cc) {
                int idx = (s * J + r) * L + cc;
                int d = (t - qc.b[idx]) % P;
                if (d < 0) d += P;
                uint32_t x = static_cast<uint32_t>((qc.ainv[idx] * d) % P);
                const int blk = r * L + cc;
                const uint32_t a16 =
                    bbase + x * 4u + (blk >= HB ? static_cast<uint32_t>(HB * P * 4) : 0u);
                aw[s][r][cc >> 1] |= a16 << (16 * (cc & 1));
            }
}

// Channel constants of one trial, uniform across the workgroup.
__device__ __forceinline__ PriorF trial_prior(const NoiseTab &nt, int trial) {
    PriorF pr;
    const int ent = __builtin_amdgcn_readfirstlane(nt.entry(trial));
    const float *q = nt.prior32 + kPrior32 * ent;
    auto uni = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        pr.lp[i] = uni(q[i]);
        pr.fx[i] = uni(q[4 + i]);
        pr.fz[i] = uni(q[8 + i]);
    }
    return pr;
}

// Estimate bit planes and outcome of one trial: variable (cc, t) is bit
// cc*P + t; P % 64 == 0 makes each wavefront's 64 variables one aligned
// 64-bit span.
template <int P, int L>
__device__ __forceinline__ void write_outcome(const DevGraph &g, const TrialPlanes &tp, int trial, int t,
                                              uint32_t estx, uint32_t estz, int iterations,
                                              int matched) {
    const int lane = t & 63;
    // 64-bit word (cc*P + t)/64 = cc*P/64 + t/64: one base per wavefront,
    // the column block's part in the store's immediate offset
    const size_t base = static_cast<size_t>(trial) * g.wn / 2 + (t >> 6);
    uint64_t *ox = reinterpret_cast<uint64_t *>(tp.est_x) + base;
    uint64_t *oz = reinterpret_cast<uint64_t *>(tp.est_z) + base;
#pragma unroll
    for (int cc = 0; cc < L; ++cc) {
        const uint64_t bx = __ballot((estx >> cc) & 1u);
        const uint64_t bz = __ballot((estz >> cc) & 1u);
        if (lane == 0) {
            // one 64-bit store per plane (a store costs about as much with
            // one active lane as with all of them)
            ox[cc * (P / 64)] = bx;
            oz[cc * (P / 64)] = bz;
        }
    }
    if (t == 0) {
        tp.iters[trial] = iterations;
        tp.syn_ok[trial] = static_cast<uint8_t>(matched);
    }
}

// Syndrome bits of checks (r, t): bit r.
template <int P, int J>
__device__ __forceinline__ void trial_syndromes(const DevGraph &g, const TrialPlanes &tp, int trial, int t,
                                                uint32_t &synx, uint32_t &synz) {
    synx = 0;
    synz = 0;
#pragma unroll
    for (int r = 0; r < J; ++r) {
        int c = r * P + t;
        synx |= ((tp.syn_x[static_cast<size_t>(trial) * g.wm[0] + (c >> 5)] >> (c & 31)) & 1u) << r;
        synz |= ((tp.syn_z[static_cast<size_t>(trial) * g.wm[1] + (c >> 5)] >> (c & 31)) & 1u) << r;
    }
}

// Plain BP (LEARNED = false) or the sum-product learned decoder: the first
// li.T iterations use the learned weights and damping; a trial that has not
// converged by then restarts plain BP from the channel prior for up to
// max_iter iterations, and its iteration count is li.T plus the plain
// iterations.  count_ptr (plain only, may be null): the trial count lives on
// the device (a list of trials left over by an earlier launch), and
// iter_offset is added to every reported iteration count.  fb_list (plain
// only, may be null): a trial still unconverged after max_iter iterations is
// appended there (fb_ctr[0] of them) instead of finishing, for a later
// launch with a larger max_iter to decode again from the prior.
template <int P, int J, int L, bool LEARNED, bool DAMPED>
__device__ __forceinline__ void qc_f32_body(DevGraph g, QcDev qc, TrialPlanes tp, int count, NoiseTab nt,
                                            int max_iter, int warmup, int interval, int *counter,
                                            LearnedInfer li, const int *count_ptr, int iter_offset,
                                            int *fb_list, int *fb_ctr) {
    static_assert(P % 64 == 0 && P <= 1024, "one thread per block row, whole wavefronts");
    extern __shared__ float B[];  // [J][L][P] exchange buffer, then GZ[L][P]
    float *GZ = B + J * L * P;   // Z-side extrinsic base per variable, parked between passes
    __shared__ int s_trial;
    __shared__ int s_cnt[2];     // learned: unsatisfied checks of the estimate under test
    __shared__ float s_d;        // learned: damping of the current iteration
    __shared__ float s_w[LEARNED ? 2 * 4 * J * L : 1];  // learned: weights, double buffer
    const int t = threadIdx.x;
    if (count_ptr) count = *count_ptr;
    uint32_t aw[2][J][(L + 1) / 2];
    qc_slot_addresses<P, J, L>(qc, B, aw);

    for (;;) {
        if (t == 0) {
            // the k-th trial handed out: order[k] (a permutation of the
            // chunk, or the list of trials a previous launch left over)
            const int k = atomicAdd(counter, 1);
            s_trial = k < count ? (tp.order ? tp.order[k] : k) : -1;
        }
        __syncthreads();
        const int trial = s_trial;
        __syncthreads();
        if (trial < 0) return;
        const PriorF pr = trial_prior(nt, trial);
        uint32_t synx, synz;  // bit r: syndrome of check (r, t)
        trial_syndromes<P, J>(g, tp, trial, static_cast<int>(threadIdx.x), synx, synz);
        float lam[2][J][L];
        uint32_t estx = 0, estz = 0;  // bit cc: hard decision of variable (cc, t)
        int matched = 0, iterations = 0;
        if (LEARNED) {
            // unsatisfied fraction of est_0 (all identity): the syndrome weight
            if (t == 0) s_cnt[0] = 0;
            __syncthreads();
            atomicAdd(&s_cnt[0], __popc(synx) + __popc(synz));
            __syncthreads();
            const float u0 = static_cast<float>(s_cnt[0]) / static_cast<float>(2 * J * P);
            __syncthreads();
            int k = qc_phase<P, J, L, true, DAMPED>(B, GZ, s_cnt, &s_d, s_w, aw, lam, synx, synz, pr,
                                                    li.T, 0, 1, li, u0, estx, estz, matched);
            if (matched) {
                iterations = k;
            } else {
                k = qc_phase<P, J, L, false, false>(B, GZ, s_cnt, &s_d, s_w, aw, lam, synx, synz, pr,
                                                    max_iter, warmup, interval, li, 0.0f, estx, estz,
                                                    matched);
                iterations = li.T + k;
            }
        } else {
            iterations = qc_phase<P, J, L, false, false>(B, GZ, s_cnt, &s_d, s_w, aw, lam, synx, synz,
                                                         pr, max_iter, warmup, interval, li, 0.0f,
                                                         estx, estz, matched) + iter_offset;
            if (fb_list && !matched) {
                if (t == 0) fb_list[atomicAdd(fb_ctr, 1)] = trial;
                __syncthreads();
                continue;
            }
        }
        write_outcome<P, L>(g, tp, trial, static_cast<int>(threadIdx.x), estx, estz, iterations, matched);
        __syncthreads();
    }
}

// Plain BP: at most 160 VGPRs (three 768-thread wavefronts' worth per SIMD,
// the few spills outside the iteration loop), which leaves room for a
// sampling wavefront of the next chunk beside a resident workgroup.
template <int P, int J, int L>
__global__ void __launch_bounds__(P) __attribute__((amdgpu_num_vgpr(80)))
decode_qc_plain_kernel(DevGraph g, QcDev qc, TrialPlanes tp, int count, NoiseTab nt, int max_iter, int warmup,
                       int interval, int *counter, LearnedInfer li, const int *count_ptr, int iter_offset,
                       int *fb_list, int *fb_ctr) {
    qc_f32_body<P, J, L, false, false>(g, qc, tp, count, nt, max_iter, warmup, interval, counter, li, count_ptr,
                                       iter_offset, fb_list, fb_ctr);
}
// The learned sum-product kinds (LEARNED = true).
template <int P, int J, int L, bool LEARNED, bool DAMPED>
__global__ void __launch_bounds__(P) decode_qc_f32_kernel(DevGraph g, QcDev qc, TrialPlanes tp, int count,
                                                         NoiseTab nt, int max_iter, int warmup, int interval,
                                                         int *counter, LearnedInfer li, const int *count_ptr,
                                                         int iter_offset, int *fb_list, int *fb_ctr) {
    qc_f32_body<P, J, L, LEARNED, DAMPED>(g, qc, tp, count, nt, max_iter, warmup, interval, counter, li,
                                          count_ptr, iter_offset, fb_list, fb_ctr);
}

// The learned flooding min-sum decoder (li.T iterations of qc_phase_ms).  A trial
// that converges is finished here; the others are appended to fb_list
// (fb_ctr[0] of them) for plain BP, launched next on the same stream, which
// decodes them from the channel prior (reported iterations li.T + its own).
// Its exchange buffer is B only (J*L*P words).
template <int P, int J, int L, bool DAMPED>
__global__ void __launch_bounds__(P) decode_qc_ms_kernel(DevGraph g, QcDev qc, TrialPlanes tp,
                                                        int count, NoiseTab nt, int *counter,
                                                        LearnedInfer li, int *fb_list, int *fb_ctr) {
    static_assert(P % 64 == 0 && P <= 1024, "one thread per block row, whole wavefronts");
    extern __shared__ float B[];  // [J][L][P] words: (X, Z) FP16 pairs
    __shared__ int s_trial;
    __shared__ int s_cnt[2];
    __shared__ float s_d;
    __shared__ __align__(16) uint32_t s_w[2 * kMsRec];
    const int t = threadIdx.x;
    uint32_t aw[2][J][(L + 1) / 2];
    qc_slot_addresses<P, J, L>(qc, B, aw);
    for (;;) {
        if (t == 0) {
            // the k-th trial handed out: order[k] (a permutation of the
            // chunk, or the list of trials a previous launch left over)
            const int k = atomicAdd(counter, 1);
            s_trial = k < count ? (tp.order ? tp.order[k] : k) : -1;
        }
        __syncthreads();
        const int trial = s_trial;
        __syncthreads();
        if (trial < 0) return;
        const PriorF pr = trial_prior(nt, trial);
        uint32_t synx, synz;
        trial_syndromes<P, J>(g, tp, trial, static_cast<int>(threadIdx.x), synx, synz);
        float u0 = 0.0f;
        if (DAMPED) {
            // unsatisfied fraction of est_0 (all identity): the syndrome weight
            if (t == 0) s_cnt[0] = 0;
            __syncthreads();
            atomicAdd(&s_cnt[0], __popc(synx) + __popc(synz));
            __syncthreads();
            u0 = static_cast<float>(s_cnt[0]) / static_cast<float>(2 * J * P);
            __syncthreads();
        }
        uint32_t estx = 0, estz = 0;
        int matched = 0;
        const int k = qc_phase_ms<P, J, L, DAMPED>(reinterpret_cast<uint32_t *>(B), s_cnt, &s_d, s_w, aw,
                                                   synx, synz, pr, li.T, li, u0, estx, estz, matched);
        if (matched) {
            write_outcome<P, L>(g, tp, trial, static_cast<int>(threadIdx.x), estx, estz, k, 1);
        } else if (t == 0) {
            fb_list[atomicAdd(fb_ctr, 1)] = trial;
        }
        __syncthreads();
    }
}

// Sum-buffer slot addresses of thread t's checks: position cc of the X
// check (r, t) is variable (cc, a*t + b mod P) of block (X, r, cc); likewise
// the Z check.  One register per (r, cc): the X check's in the low 16 bits,
// the Z check's in the high 16 (see lms_off).
template <int P, int J, int L>
__device__ __forceinline__ void lms_slot_addresses(const int *qc_tab, uint32_t sb,
                                                   uint32_t (&aw)[J][L]) {
    // the (a, b) tables of both sides pass through LDS (the sum buffers are
    // not in use yet), so they are read per thread rather than held in
    // scalar registers
    const int t = threadIdx.x;
    constexpr int JL = J * L;
    static_assert(4 * JL <= P, "one thread per table entry");
    if (t < 4 * JL) {
        const int s = t / (2 * JL), rest = t % (2 * JL);  // rest: a (0..JL-1) then b
        *reinterpret_cast<lds_u32 *>(static_cast<uintptr_t>(sb + t * 4)) =
            static_cast<uint32_t>(qc_tab[(3 * s + rest / JL) * JL + rest % JL]);
    }
    lds_barrier();
#pragma unroll
    for (int r = 0; r < J; ++r)
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
            uint32_t a16[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const lds_u32 *row = reinterpret_cast<lds_u32 *>(static_cast<uintptr_t>(sb + s * 2 * JL * 4));
                const int a = static_cast<int>(row[r * L + cc]);
                const int b = static_cast<int>(row[JL + r * L + cc]);
                const uint32_t y = static_cast<uint32_t>((a * t + b) % P);
                a16[s] = sb + y * 4u + lms_split<P>(cc);
            }
            aw[r][cc] = a16[0] | (a16[1] << 16);
        }
    lds_barrier();
}

// The layered learned min-sum decoder (li.T iterations of qc_phase_lms).
// A trial that converges is finished here; the others are appended to
// fb_list (fb_ctr[0] of them) for plain BP, launched next, which decodes
// them from the channel prior (reported iterations li.T + its own).  LDS:
// the three sum buffers (3*L*P words).
template <int P, int J, int L>
__global__ void __launch_bounds__(P) __attribute__((amdgpu_num_vgpr(80))) decode_qc_lms_kernel(DevGraph g, TrialPlanes tp, int count,
                                                         NoiseTab nt, int *counter, LearnedInfer li,
                                                         int *fb_list, int *fb_ctr) {
    static_assert(P % 64 == 0 && P <= 1024, "one thread per check of a block row, whole wavefronts");
    extern __shared__ uint32_t Sbuf[];
    __shared__ int s_trial;
    __shared__ int s_flag[6];
    const int t0 = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) & ~63);  // wavefront base
    const uint32_t sb = static_cast<uint32_t>((uintptr_t)((lds_u32 *)Sbuf));  // LDS address of the sums
    uint32_t aw[J][L];
    lms_slot_addresses<P, J, L>(g.qc_tab, sb, aw);
    const bool first = t0 == 0;
    const cu32 *rec = (const cu32 *)li.ms_rec;  // scalar loads
    for (;;) {
        if (first && lane_now() == 0) {
            const int k = atomicAdd(counter, 1);
            s_trial = k < count ? (tp.order ? tp.order[k] : k) : -1;
        }
        lds_barrier();
        const int trial = __builtin_amdgcn_readfirstlane(s_trial);
        if (first && lane_now() < 6) s_flag[lane_now()] = 0;
        lds_barrier();
        if (trial < 0) return;
        const PriorF pr = trial_prior(nt, trial);
        uint32_t synx, synz;
        trial_syndromes<P, J>(g, tp, trial, t0 + static_cast<int>(lane_now()), synx, synz);
        uint32_t estx = 0, estz = 0;
        int matched = 0;
        const int k = qc_phase_lms<P, J, L>(sb, t0, aw, synx, synz, pr, li.T, rec, s_flag, estx, estz, matched);
        if (matched) {
            write_outcome<P, L>(g, tp, trial, t0 + static_cast<int>(lane_now()), estx, estz, k, 1);
        } else if (first && lane_now() == 0) {
            fb_list[atomicAdd(fb_ctr, 1)] = trial;
        }
        lds_barrier();
    }
}

}  // namespace

void prior32_from(const Prior &pr, float out[kPrior32]) {
    const double *p = pr.p;
    const int fx[4] = {0, 2, 1, 3}, fz[4] = {0, 1, 2, 3};
    for (int i = 0; i < 4; ++i) {
        out[i] = static_cast<float>(std::log2(p[i]));
        out[4 + i] = static_cast<float>(p[fx[i]]);
        out[8 + i] = static_cast<float>(p[fz[i]]);
    }
}

bool qc_f32_supported(const QcParams &qc, int device, char *why, size_t len) {
    auto fail = [&](const char *msg) {
        if (why && len) snprintf(why, len, "%s", msg);
        return false;
    };
    if (qc.P != 768 || qc.J != 3 || qc.L != 12)
        return fail("LDS decoder is instantiated for P=768, J=3, L=12");
    size_t lds = static_cast<size_t>(qc.J + 1) * qc.L * qc.P * sizeof(float);
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess)
        return fail("cannot query LDS size");
    if (lds + 64 > static_cast<size_t>(max_lds)) return fail("exchange buffer exceeds LDS");
    return true;
}

void launch_decode_qc_f32(const DevGraph &g, const QcParams &qc, const int *d_tables,
                          TrialPlanes tp, int count, const NoiseTab &nt, int max_iter, int check_warmup,
                          int check_interval, int slots, int *counter, hipStream_t st,
                          const LearnedInfer *learned, int *fb, hipEvent_t phase_end,
                          hipStream_t fb_st, int cap_iter) {
    if (count <= 0) return;
    QcDev dev;
    dev.P = qc.P;
    dev.ainv = d_tables;                  // [2][J][L]
    dev.b = d_tables + 2 * qc.J * qc.L;   // [2][J][L]
    size_t lds = static_cast<size_t>(qc.J + 1) * qc.L * qc.P * sizeof(float);
    int grid = slots < count ? slots : count;
    int interval = check_interval > 0 ? check_interval : 1;
    int warmup = check_warmup > 0 ? check_warmup : 0;
    LearnedInfer li;
    if (learned) li = *learned;
    auto launch = [&](auto kern, const int *count_ptr, int offset, int *ctr, TrialPlanes planes,
                      hipStream_t on, int iters, int *list, int *list_ctr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(768), lds, on, g, dev, planes, count, nt, iters,
                           warmup, interval, ctr, li, count_ptr, offset, list, list_ctr);
    };
    // the second launch of a two-phase decode: plain BP from the prior over
    // the list the first launch left in fb (fb: [0] its count, [1] the work
    // counter, fb + 4 the list), on fb_st after phase_end when given
    auto second = [&](int offset) {
        if (phase_end) (void)hipEventRecord(phase_end, st);
        hipStream_t on = st;
        if (fb_st && phase_end) {
            (void)hipStreamWaitEvent(fb_st, phase_end, 0);
            on = fb_st;
        }
        TrialPlanes fbp = tp;
        fbp.order = fb + 4;
        li = LearnedInfer();
        launch(decode_qc_plain_kernel<768, 3, 12>, fb, offset, fb + 1, fbp, on, max_iter,
               nullptr, nullptr);
    };
    if (learned && learned->T > 0 && (learned->kind == kLearnedMS || learned->kind == kLearnedLMS)) {
        // min-sum phase, then plain BP over the trials it left unconverged
        launch_fill_i32(fb, 0, 2, st);
        const size_t ms_lds = static_cast<size_t>(qc.J) * qc.L * qc.P * sizeof(uint32_t);
        auto ms = [&](auto kern) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(ms_lds));
            hipLaunchKernelGGL(kern, dim3(grid), dim3(768), ms_lds, st, g, dev, tp, count, nt, counter, li,
                               fb + 4, fb);
        };
        if (learned->kind == kLearnedLMS) {
            const size_t lms_lds = static_cast<size_t>(3) * qc.L * qc.P * sizeof(uint32_t);
            auto kern = decode_qc_lms_kernel<768, 3, 12>;
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lms_lds));
            hipLaunchKernelGGL(kern, dim3(grid), dim3(768), lms_lds, st, g, tp, count, nt, counter, li, fb + 4,
                               fb);
        } else if (learned->use_mlp) {
            ms(decode_qc_ms_kernel<768, 3, 12, true>);
        } else {
            ms(decode_qc_ms_kernel<768, 3, 12, false>);
        }
        second(learned->T);
    } else if (learned && learned->T > 0) {
        if (learned->use_mlp)
            launch(decode_qc_f32_kernel<768, 3, 12, true, true>, nullptr, 0, counter, tp, st, max_iter,
                   nullptr, nullptr);
        else
            launch(decode_qc_f32_kernel<768, 3, 12, true, false>, nullptr, 0, counter, tp, st, max_iter,
                   nullptr, nullptr);
    } else if (cap_iter > 0 && cap_iter < max_iter && fb) {
        // plain BP in two launches: up to cap_iter iterations for every
        // trial, then, for those still unconverged, the whole decode again
        // from the prior up to max_iter.  BP is deterministic, so each trial
        // ends exactly as in one launch; the first launch ends at the pace of
        // the bulk instead of its slowest trial.
        launch_fill_i32(fb, 0, 2, st);
        launch(decode_qc_plain_kernel<768, 3, 12>, nullptr, 0, counter, tp, st, cap_iter,
               fb + 4, fb);
        second(0);
    } else {
        launch(decode_qc_plain_kernel<768, 3, 12>, nullptr, 0, counter, tp, st, max_iter,
               nullptr, nullptr);
    }
}

}  // namespace jbp
