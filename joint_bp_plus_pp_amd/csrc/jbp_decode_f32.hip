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

// LEARNED: the first li.T iterations use the learned weights and damping; a
// trial that has not converged by then restarts plain BP from the channel
// prior for up to max_iter iterations, and its iteration count is li.T plus
// the plain iterations.
template <int P, int J, int L, bool LEARNED, bool DAMPED>
__global__ void __launch_bounds__(P) decode_qc_f32_kernel(DevGraph g, QcDev qc, TrialPlanes tp,
                                                         int count, NoiseTab nt, int max_iter,
                                                         int warmup, int interval, int *counter,
                                                         LearnedInfer li) {
    static_assert(P % 64 == 0 && P <= 1024, "one thread per block row, whole wavefronts");
    extern __shared__ float B[];  // [J][L][P] exchange buffer, then GZ[L][P]
    float *GZ = B + J * L * P;   // Z-side extrinsic base per variable, parked between passes
    __shared__ int s_trial;
    __shared__ int s_cnt;        // learned: unsatisfied checks of the estimate under test
    __shared__ float s_d;        // learned: damping of the current iteration
    __shared__ float s_w[LEARNED ? 2 * 4 * J * L : 1];  // learned: weights, double buffer
    const int t = threadIdx.x;
    const int lane = t & 63;

    // Row index x = ainv*(t - b) mod P of this thread's variable (cc, t) in
    // block (side, r, cc); its exchange slot is B[(r*L + cc)*P + x], kept as
    // a 16-bit LDS address (two per register, see qc_phase).
    constexpr int HB = J * L / 2;
    static_assert((P - 1) * 4 + HB * P * 4 + 4096 < 65536 && (HB - 1) * P * 4 < 65536,
                  "exchange addresses must fit 16 bits and the LDS immediate offset");
    const uint32_t bbase = static_cast<uint32_t>((uintptr_t)((lds_float *)B));  // LDS address of B
    constexpr int W = (L + 1) / 2;
    uint32_t aw[2][J][W];
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

    for (;;) {
        if (t == 0) {
            const int k = atomicAdd(counter, 1);
            s_trial = (tp.order && k < count) ? tp.order[k] : k;
        }
        __syncthreads();
        const int trial = s_trial;
        __syncthreads();
        if (trial >= count) return;
        // this trial's channel constants, uniform across the workgroup
        PriorF pr;
        {
            const int ent = __builtin_amdgcn_readfirstlane(nt.entry(trial));
            const float *q = nt.prior32 + kPrior32 * ent;
            auto uni = [](float x) {
                return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
            };
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                pr.lp[i] = uni(q[i]);
                pr.fx[i] = uni(q[4 + i]);
                pr.fz[i] = uni(q[8 + i]);
            }
        }

        uint32_t synx = 0, synz = 0;  // bit r: syndrome of check (r, t)
#pragma unroll
        for (int r = 0; r < J; ++r) {
            int c = r * P + t;
            synx |= ((tp.syn_x[static_cast<size_t>(trial) * g.wm[0] + (c >> 5)] >> (c & 31)) & 1u) << r;
            synz |= ((tp.syn_z[static_cast<size_t>(trial) * g.wm[1] + (c >> 5)] >> (c & 31)) & 1u) << r;
        }
        float lam[2][J][L];
        uint32_t estx = 0, estz = 0;  // bit cc: hard decision of variable (cc, t)
        int matched = 0, iterations = 0;
        if (LEARNED) {
            // unsatisfied fraction of est_0 (all identity): the syndrome weight
            if (t == 0) s_cnt = 0;
            __syncthreads();
            atomicAdd(&s_cnt, __popc(synx) + __popc(synz));
            __syncthreads();
            const float u0 = static_cast<float>(s_cnt) / static_cast<float>(2 * J * P);
            int k = qc_phase<P, J, L, true, DAMPED>(B, GZ, &s_cnt, &s_d, s_w, aw, lam, synx, synz, pr, li.T, 0,
                                            1, li, u0, estx, estz, matched);
            if (matched) {
                iterations = k;
            } else {
                k = qc_phase<P, J, L, false, false>(B, GZ, &s_cnt, &s_d, s_w, aw, lam, synx, synz, pr,
                                             max_iter, warmup, interval, li, 0.0f, estx, estz,
                                             matched);
                iterations = li.T + k;
            }
        } else {
            iterations = qc_phase<P, J, L, false, false>(B, GZ, &s_cnt, &s_d, s_w, aw, lam, synx, synz, pr,
                                                  max_iter, warmup, interval, li, 0.0f, estx,
                                                  estz, matched);
        }

        // est bit planes: variable (cc, t) is bit cc*P + t; P % 64 == 0
        // makes each wavefront's 64 variables one aligned 64-bit span.
        uint32_t *ox = tp.est_x + static_cast<size_t>(trial) * g.wn;
        uint32_t *oz = tp.est_z + static_cast<size_t>(trial) * g.wn;
#pragma unroll
        for (int cc = 0; cc < L; ++cc) {
            uint64_t bx = __ballot((estx >> cc) & 1u);
            uint64_t bz = __ballot((estz >> cc) & 1u);
            if (lane == 0) {
                int w = (cc * P + t) >> 5;
                ox[w] = static_cast<uint32_t>(bx);
                ox[w + 1] = static_cast<uint32_t>(bx >> 32);
                oz[w] = static_cast<uint32_t>(bz);
                oz[w + 1] = static_cast<uint32_t>(bz >> 32);
            }
        }
        if (t == 0) {
            tp.iters[trial] = iterations;
            tp.syn_ok[trial] = static_cast<uint8_t>(matched);
        }
        __syncthreads();
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
                          const LearnedInfer *learned) {
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
    auto launch = [&](auto kern) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(768), lds, st, g, dev, tp, count, nt, max_iter,
                           warmup, interval, counter, li);
    };
    if (learned && learned->T > 0) {
        if (learned->use_mlp) launch(decode_qc_f32_kernel<768, 3, 12, true, true>);
        else launch(decode_qc_f32_kernel<768, 3, 12, true, false>);
    } else {
        launch(decode_qc_f32_kernel<768, 3, 12, false, false>);
    }
}

}  // namespace jbp
