// Learned min-sum training in FP32 — the flooding kind (kind 1, with its MLP
// damping) and the layered kind (kind 2, the headline decoder's model) — for
// the permutation-block code of the params files (P = 768, J = 3, L = 12):
// forward passes with their tape and adjoint passes, on chip.  No
// counterpart in the reference program.
//
// The models are the generic training kernels' (jbp_learned.hip).  Flooding:
//   mu_e     = wv[g][b][t] * (S_g(v) - lambda_e + F_g(S_other(v)))
//   lambda_e = (1 - d_t) * wc[g][b][t] * CN_e(mu) + d_t * lambda_e(t-1)
// with the min-sum check node clamped at 64 ln 2, F(L) = max(lI + L, lZ) -
// max(lX + L, lY) in natural logs, d_t from the 2-16-1 MLP of (t / T,
// unsatisfied fraction of the estimate two iterations back), and the BCE of
// the X / Z marginals averaged over iterations 1..T.  The layered kind runs
// the same updates layer by layer (below).
//
// Layout: one 768-thread workgroup per trial; thread t owns the variables
// (cc, t) and the checks (r, t).  The variable owner keeps the 72 messages
// lambda of its variables' edges in registers (forward) or their gradients
// (adjoint).  The two views meet in an LDS exchange buffer, variable-major:
// the edge of row r and column block cc at variable (cc, y) sits at [r][cc]
// [y], so the variable owner's slots are fixed offsets from one base, and
// the check owner (r, x) forms the address of its input i from the block's
// permutation, y = (a x + b) mod P, when it needs it (four instructions per
// input, in place of 36 registers of addresses).  The tape is [trial][k =
// 0..T][72 edges][768 threads] floats, written as whole wavefront rows.  The
// forward exchanges mu (check owners read it and write the clamped outputs
// back), with the estimate bit of a variable in the least significant bit
// of every mu it sends, so the check owners count unsatisfied checks while
// reading their rows; it also records every check row (where its two
// smallest input magnitudes are, the signs, the two magnitudes).  The
// adjoint writes each check's raw outputs from its record, then exchanges
// the gradient of the raw outputs and the gradient of mu.  The flooding
// adjoint holds one side's gradients in registers and parks the other's in
// the tape row its loss term has consumed.  Weight gradients of a block row
// are summed over the wavefront in one transposing reduction and
// accumulated per iteration and block in LDS, then added to the global
// gradient once per workgroup.
#include "jbp_common.h"
#include "jbp_learned.h"

namespace jbp {
namespace {

constexpr int QP = 768, QJ = 3, QL = 12, QNB = QJ * QL, QHB = QNB / 2, QK = 2 * QNB;
constexpr float kLmaxQ = 44.3614195558365f;  // 64 ln 2

typedef __attribute__((address_space(3))) float lds_f32;

__device__ __forceinline__ float ldsf(uint32_t a) {
    return *reinterpret_cast<lds_f32 *>(static_cast<uintptr_t>(a));
}
__device__ __forceinline__ void stf(uint32_t a, float v) {
    *reinterpret_cast<lds_f32 *>(static_cast<uintptr_t>(a)) = v;
}
__device__ __forceinline__ float with_lsb(float v, uint32_t bit) {
    return __uint_as_float((__float_as_uint(v) & ~1u) | bit);
}
// F(L) = max(lk0 + L, lk1) - max(lk2 + L, lk3) and its derivative
__device__ __forceinline__ float fterm(float L, const float *lk) {
    const float u = lk[0] + L, v = lk[2] + L;
    return (u > lk[1] ? u : lk[1]) - (v > lk[3] ? v : lk[3]);
}
__device__ __forceinline__ float fterm_d(float L, const float *lk) {
    return (lk[0] + L > lk[1] ? 1.0f : 0.0f) - (lk[2] + L > lk[3] ? 1.0f : 0.0f);
}
// The logistic function and softplus from one base-2 exponential of -|x|
// (hardware exp2 / log2 / reciprocal, within about an ulp), without
// branches: sigma(x) = 1 / (1 + e) for x >= 0 and e / (1 + e) below, with
// e = exp(-|x|); softplus(x) = max(x, 0) + ln(1 + e).
constexpr float kLog2e = 1.44269504088896341f, kLn2 = 0.693147180559945309f;
__device__ __forceinline__ float sigmf(float x) {
    const float e = __builtin_amdgcn_exp2f(-fabsf(x) * kLog2e);
    const float s = __builtin_amdgcn_rcpf(1.0f + e);
    return x >= 0.0f ? s : e * s;
}
__device__ __forceinline__ float softplusf(float x) {
    const float e = __builtin_amdgcn_exp2f(-fabsf(x) * kLog2e);
    return fmaxf(x, 0.0f) + __builtin_amdgcn_logf(1.0f + e) * kLn2;
}

// Sum over the wavefront, in every lane (row sums by DPP, then the four
// rows' sums by readlane).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<0xb1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4e>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror: every lane holds its row's sum
    const int b = __float_as_int(v);
    return (__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
           (__int_as_float(__builtin_amdgcn_readlane(b, 32)) + __int_as_float(__builtin_amdgcn_readlane(b, 48)));
}

// Sums over the wavefront of 12 per-lane values at once (a transposing
// reduction): each step halves the values a lane holds, the lanes on one
// side of a lane-index bit keeping the first half summed with their
// partners', the others the second half.  Bits 5 and 4 use the lane-swap
// instructions (permlane32 / permlane16 swap: the two results' sum holds the
// first operand's pair sums in the lower lanes and the second's in the
// upper), bits 3 and 2 two masked row shifts, bits 1 and 0 a butterfly.  On
// return a lane holds the sum of value slot12(lane) (when that is >= 0; the
// four lanes of a quad hold the same sum).
__device__ __forceinline__ float swap_sum32(float a, float b) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ __forceinline__ float swap_sum16(float a, float b) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
// lanes with the bit clear keep a plus the partner's a (row_shl), the others
// b plus the partner's b (row_shr); LO / HI are the banks of each side
template <int SHL, int SHR, int LO, int HI>
__device__ __forceinline__ float shift_sum(float a, float b, bool hi) {
    float t = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a), SHL, 0xf, LO, false));
    t = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(t), __float_as_int(b), SHR, 0xf, HI, false));
    return (hi ? b : a) + t;
}
__device__ __forceinline__ float reduce12(const float (&v)[12], uint32_t lane) {
    float w[6], x[3];
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = swap_sum32(v[k], v[k + 6]);
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = swap_sum16(w[k], w[k + 3]);
    const bool b3 = (lane >> 3) & 1u, b2 = (lane >> 2) & 1u;
    // bit 3: pairs (0, 2) and (1, padding)
    const float y0 = shift_sum<0x108, 0x118, 0x3, 0xc>(x[0], x[2], b3);
    const float y1 = shift_sum<0x108, 0x118, 0x3, 0xc>(x[1], 0.0f, b3);
    // bit 2: the pair (y0, y1)
    float z = shift_sum<0x104, 0x114, 0x5, 0xa>(y0, y1, b2);
    z += dpp_f<0xb1>(z);  // quad_perm [1,0,3,2]
    z += dpp_f<0x4e>(z);  // quad_perm [2,3,0,1]
    return z;
}
// The value slot whose sum reduce12 leaves in this lane, or -1 (padding).
__device__ __forceinline__ int slot12(uint32_t lane) {
    const int b2 = (lane >> 2) & 1, b3 = (lane >> 3) & 1, b4 = (lane >> 4) & 1, b5 = (lane >> 5) & 1;
    const int s3 = b2 + 2 * b3;  // slot among x[0..2] (3 = padding)
    return s3 < 3 ? s3 + 3 * b4 + 6 * b5 : -1;
}

// The variable owner's slots: B[r][cc][t] from a base opaque to the
// optimiser (re-formed at each pass, so that per-slot addresses are not
// kept live); blocks of the upper half from a second base, so every offset
// fits the 16-bit immediate.
struct VarSlots {
    uint32_t lo, hi;
};
__device__ __forceinline__ VarSlots var_slots(uint32_t bbase) {
    uint32_t a = bbase + static_cast<uint32_t>(threadIdx.x) * 4u;
    asm volatile("" : "+v"(a));
    return {a, a + static_cast<uint32_t>(QHB * QP * 4)};
}
__device__ __forceinline__ uint32_t vslot(const VarSlots &v, int r, int cc) {
    const int blk = r * QL + cc;
    return (blk < QHB ? v.lo : v.hi) + static_cast<uint32_t>((blk % QHB) * QP * 4);
}
// The check owner's input i of row r (side s): variable i*P + y with
// y = (a x + b) mod P, a and b of block r*L + i from the code's table
// ([side][a | b | ainv][r*L + cc]).  a x + b < 2^20, where
// floor(y / 768) = mulhi(y, 5592406) exactly.
// The table is read through the constant address space: it is never
// written while the kernels run, so its uniform reads are scalar loads
// (vector loads would share the vector memory counter with the tape's
// loads and stores, and waiting for them would wait for those too).
typedef __attribute__((address_space(4))) const int ctab_i32;
__device__ __forceinline__ uint32_t check_slot(uint32_t bbase, const int *tab, int s, int r, int i, int x) {
    constexpr int JL = QJ * QL;
    const int blk = r * QL + i;
    const ctab_i32 *ct = (const ctab_i32 *)tab;
    const uint32_t y = static_cast<uint32_t>(ct[(3 * s) * JL + blk] * x + ct[(3 * s + 1) * JL + blk]);
    const uint32_t q = __umulhi(y, 5592406u);
    return bbase + static_cast<uint32_t>(blk * QP * 4) + (y - q * static_cast<uint32_t>(QP)) * 4u;
}
static_assert(QP == 768, "check_slot's division constant is for P = 768");

struct QcTrial {
    float lx[4], lz[4];     // log prior terms: X side (I, Z, X, Y), Z side (I, X, Z, Y)
    uint32_t ebx, ebz;      // error bits of the variables (cc, t): bit cc
    uint32_t synx, synz;    // syndrome bits of the checks (r, t): bit r
};
__device__ __forceinline__ QcTrial qc_trial(const DevGraph &g, const TrialPlanes &tp, const double *prior, int b) {
    QcTrial q;
    const double *p4 = prior + 4 * static_cast<size_t>(b);
    const int ox[4] = {0, 2, 1, 3};
    // per-trial constants, uniform over the workgroup (scalar registers)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        q.lx[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(static_cast<float>(log(p4[ox[i]])))));
        q.lz[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(static_cast<float>(log(p4[i])))));
    }
    // an opaque thread index, so that the addresses below are formed here,
    // once per trial, rather than hoisted out of the trial loop and spilled
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    q.ebx = q.ebz = q.synx = q.synz = 0;
#pragma unroll
    for (int cc = 0; cc < QL; ++cc) {
        const int v = cc * QP + t;
        const size_t w = static_cast<size_t>(b) * g.wn + (v >> 5);
        q.ebx |= ((tp.err_x[w] >> (v & 31)) & 1u) << cc;
        q.ebz |= ((tp.err_z[w] >> (v & 31)) & 1u) << cc;
    }
#pragma unroll
    for (int r = 0; r < QJ; ++r) {
        const int c = r * QP + t;
        q.synx |= ((tp.syn_x[static_cast<size_t>(b) * g.wm[0] + (c >> 5)] >> (c & 31)) & 1u) << r;
        q.synz |= ((tp.syn_z[static_cast<size_t>(b) * g.wm[1] + (c >> 5)] >> (c & 31)) & 1u) << r;
    }
    return q;
}

__device__ float mlp_fwd(const float *mlp, float x0, float x1, float *h) {
    float z = mlp[4 * kMlpHidden];
    for (int i = 0; i < kMlpHidden; ++i) {
        h[i] = tanhf(mlp[2 * i] * x0 + mlp[2 * i + 1] * x1 + mlp[2 * kMlpHidden + i]);
        z += mlp[3 * kMlpHidden + i] * h[i];
    }
    return sigmf(z);
}

// The tape of one trial through a buffer resource: element (edge e, thread
// t) of row k sits at byte k * 72 * 768 * 4 + e * 768 * 4 + t * 4; the
// thread's part is the one vector offset, the rest is uniform (scalar
// offset), so no per-edge addresses are held in vector registers.
typedef __amdgpu_buffer_rsrc_t tape_rsrc;
__device__ __forceinline__ tape_rsrc make_tape(const float *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), 0, 0x7fffffff, 0x00020000);
}
constexpr uint32_t kRowBytes = static_cast<uint32_t>(QK) * QP * 4;
__device__ __forceinline__ uint32_t tape_off(int k, int s, int r, int cc) {
    return static_cast<uint32_t>(k) * kRowBytes + static_cast<uint32_t>(((s * QJ + r) * QL + cc) * QP * 4);
}
// Row k's scalar offset, opaque to the optimiser: each pass forms its
// edges' offsets from it (one scalar add each) instead of sharing 72
// offsets across passes, which would be kept in scalar registers and
// spilled.
__device__ __forceinline__ uint32_t tape_row(int k) {
    uint32_t o = static_cast<uint32_t>(k) * kRowBytes;
    asm volatile("" : "+s"(o));
    return o;
}
__device__ __forceinline__ uint32_t edge_off(int s, int r, int cc) {
    return static_cast<uint32_t>(((s * QJ + r) * QL + cc) * QP * 4);
}
__device__ __forceinline__ float tape_ld(tape_rsrc rs, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}
__device__ __forceinline__ void tape_st(tape_rsrc rs, uint32_t voff, uint32_t soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, voff, soff, 0);
}

// A scheduling fence between the tape reads of successive column blocks: it
// bounds the loads in flight (and their registers) to a few blocks' worth.
__device__ __forceinline__ void tape_fence() { __builtin_amdgcn_sched_barrier(0); }

// Check records: for every iteration, side and block row, the forward
// pass's check (r, x) leaves three words for the adjoint — where the
// smallest input magnitude is (i1), where the smallest of the others is
// (i2), which inputs and raw outputs are negative (bits 8..19, 20..31), and
// the two smallest magnitudes — so that the adjoint writes the raw outputs
// without forming mu again.  Per trial [iteration - 1][side * 3 + row][word]
// [768], coalesced like the tape.
constexpr uint32_t kChkIterBytes = 6u * 3u * QP * 4u;
__device__ __forceinline__ uint32_t chk_word(int s, int r, int word) {
    return static_cast<uint32_t>(((s * QJ + r) * 3 + word) * QP * 4);
}
__device__ __forceinline__ uint32_t chk_iter(int it) {
    uint32_t o = static_cast<uint32_t>(it - 1) * kChkIterBytes;
    asm volatile("" : "+s"(o));
    return o;
}
__device__ __forceinline__ tape_rsrc make_chk(uint32_t *p) {
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ uint32_t chk_ld(tape_rsrc rs, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0);
}
__device__ __forceinline__ void chk_st(tape_rsrc rs, uint32_t voff, uint32_t soff, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, rs, voff, soff, 0);
}
// the record of one check row: i1, i2, signs of the inputs and of the raw
// outputs (raw_i < 0), from the inputs' signs and the two minima
__device__ __forceinline__ uint32_t chk_info(int i1, int i2, uint32_t neg, uint32_t negall, float m1, float m2) {
    // raw_i = (-1)^(negall ^ neg_i) * (i == i1 ? m2 : m1) is negative when
    // that sign is and the magnitude is not zero
    constexpr uint32_t kAll = (1u << QL) - 1u;
    const uint32_t sgn = (negall ? ~neg : neg) & kAll;
    const uint32_t b1 = 1u << i1;
    const uint32_t nz = ((m1 > 0.0f ? kAll : 0u) & ~b1) | (m2 > 0.0f ? b1 : 0u);
    return static_cast<uint32_t>(i1) | (static_cast<uint32_t>(i2) << 4) | (neg << 8) | ((sgn & nz) << 20);
}
// raw output i of a recorded check row
__device__ __forceinline__ float chk_raw(uint32_t info, float m1, float m2, int i) {
    const float mag = i == static_cast<int>(info & 15u) ? m2 : m1;
    return ((info >> (20 + i)) & 1u) ? -mag : mag;
}

// Pin one side's messages: the update that forms them is completed here,
// not sunk past the next barrier with its LDS operands held in registers.
__device__ __forceinline__ void pin_side(float (&m)[QJ][QL]) {
#pragma unroll
    for (int r = 0; r < QJ; ++r)
#pragma unroll
        for (int cc = 0; cc < QL; ++cc) asm volatile("" : "+v"(m[r][cc]));
}

__device__ __forceinline__ void pin_row(float (&m)[QL]) {
#pragma unroll
    for (int cc = 0; cc < QL; ++cc) asm volatile("" : "+v"(m[cc]));
}

// The weights of one iteration staged in LDS: wc X | wc Z | wv X | wv Z, 36
// blocks each (the parameter vector holds [4][36][T]).
constexpr int QW = 4 * QNB;
__device__ __forceinline__ void stage_weights(float *dst, const float *params, int T, int it) {
    const int i = threadIdx.x;
    if (i < QW && it >= 1 && it <= T) dst[i] = params[static_cast<size_t>(i) * T + it - 1];
}

// Check rows of side s: min-sum on the 12 inputs of each check (r, t), in
// place; returns the number of those checks whose inputs' estimate bits
// (their least significant bits) disagree with the syndrome.
__device__ __forceinline__ int qc_check_rows_fwd(uint32_t bbase, const int *tab, int s, uint32_t syn,
                                                 tape_rsrc chk, uint32_t voff, uint32_t crow) {
    int x = threadIdx.x;
    asm volatile("" : "+v"(x));
    int unsat = 0;
#pragma unroll 1
    for (int r = 0; r < QJ; ++r) {
        float mu[QL];
        uint32_t par = 0, neg = 0, negall = (syn >> r) & 1u;
        float m1 = HUGE_VALF, m2 = HUGE_VALF;
        int i1 = -1, i2 = -1;
#pragma unroll
        for (int i = 0; i < QL; ++i) mu[i] = ldsf(check_slot(bbase, tab, s, r, i, x));
#pragma unroll
        for (int i = 0; i < QL; ++i) {
            par ^= __float_as_uint(mu[i]) & 1u;
            const float a = fabsf(mu[i]);
            const uint32_t ng = mu[i] < 0.0f ? 1u : 0u;
            neg |= ng << i;
            negall ^= ng;
            // the two smallest magnitudes and the first index of the
            // smallest, without branches
            const bool lt1 = a < m1, lt2 = a < m2;
            i2 = lt1 ? i1 : (lt2 ? i : i2);  // the first index of the minimum over the others
            m2 = lt1 ? m1 : (lt2 ? a : m2);
            i1 = lt1 ? i : i1;
            m1 = lt1 ? a : m1;
        }
        unsat += (par != ((syn >> r) & 1u)) ? 1 : 0;
        chk_st(chk, voff, crow + chk_word(s, r, 0), chk_info(i1, i2, neg, negall, m1, m2));
        chk_st(chk, voff, crow + chk_word(s, r, 1), __float_as_uint(m1));
        chk_st(chk, voff, crow + chk_word(s, r, 2), __float_as_uint(m2));
#pragma unroll
        for (int i = 0; i < QL; ++i) {
            const float mag = i == i1 ? m2 : m1;
            const float out = mag < kLmaxQ ? mag : kLmaxQ;
            stf(check_slot(bbase, tab, s, r, i, x), ((negall ^ (neg >> i)) & 1u) ? -out : out);
        }
    }
    return unsat;
}

// mu of the variable (cc, t)'s edges on side s from its messages (the
// forward's registers or the tape) and the staged weights w; the adjoint
// forms it the same way, bit for bit.  The estimate bit of the side rides in
// the least significant bit.
__device__ __forceinline__ void qc_mu_side(int s, const float (&lx3)[QJ], const float (&lz3)[QJ], const QcTrial &q,
                                           const float *w, int cc, float (&mu)[QJ]) {
    const float sx = (lx3[0] + lx3[1]) + lx3[2];
    const float sz = (lz3[0] + lz3[1]) + lz3[2];
    const float fx = fterm(sz, q.lx), fz = fterm(sx, q.lz);
    const uint32_t ex = (sx + fx < 0.0f) ? 1u : 0u, ez = (sz + fz < 0.0f) ? 1u : 0u;
#pragma unroll
    for (int r = 0; r < QJ; ++r) {
        const float a = s == 0 ? (sx - lx3[r]) + fx : (sz - lz3[r]) + fz;
        mu[r] = with_lsb(w[2 * QNB + s * QNB + r * QL + cc] * a, s == 0 ? ex : ez);
    }
}

__global__ void __launch_bounds__(QP) learned_qc_fwd_kernel(DevGraph g, const float *params, int T, int use_mlp,
                                                           TrialPlanes tp, float *act, uint32_t *chk,
                                                           float *tape_d, float *tape_u,
                                                           const double *prior, float *loss, int *conv_out,
                                                           int batch, int *counter) {
    extern __shared__ float B[];  // [J][L][P] | GZ [L][P]
    float *GZ = B + QNB * QP;     // the Z side's channel term of each variable, kept over the X update
    __shared__ int s_trial;
    __shared__ int s_cnt[2];
    __shared__ float s_d, s_loss;
    __shared__ float s_ubar[kMaxLearnedT + 1];
    __shared__ float s_w[2][QW];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t bbase = static_cast<uint32_t>((uintptr_t)((lds_f32 *)B));
    const float *mlp = params + 4 * QNB * T;
    const size_t row = static_cast<size_t>(QK) * QP;
    const float scale = 1.0f / (static_cast<float>(batch) * static_cast<float>(T) * static_cast<float>(2 * g.n));
    const float mtot = static_cast<float>(2 * QJ * QP);
    for (;;) {
        if (t == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_trial);
        __syncthreads();
        if (b >= batch) return;
        QcTrial q = qc_trial(g, tp, prior, b);
        const tape_rsrc tape = make_tape(act + static_cast<size_t>(b) * (T + 1) * row);
        const tape_rsrc chkr = make_chk(chk + static_cast<size_t>(b) * T * (kChkIterBytes / 4));
        const uint32_t voff = static_cast<uint32_t>(t) * 4u;
        float lam[2][QJ][QL];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    lam[s][r][cc] = 0.0f;
                    tape_st(tape, voff, tape_off(0, s, r, cc), 0.0f);
                }
        stage_weights(s_w[1], params, T, 1);
        if (t == 0) {
            s_cnt[0] = s_cnt[1] = 0;
            s_loss = 0.0f;
        }
        __syncthreads();
        {
            const float w0 = wave_sum(static_cast<float>(__popc(q.synx) + __popc(q.synz)));
            if (lane == 0) atomicAdd(&s_cnt[0], static_cast<int>(w0));
        }
        __syncthreads();
        if (t == 0) s_ubar[0] = static_cast<float>(s_cnt[0]) / mtot;
        int conv = 0;
        float lsum = 0.0f;
        for (int it = 1; it <= T; ++it) {
            VarSlots vs = var_slots(bbase);
            const float *w = s_w[it & 1];
            stage_weights(s_w[(it + 1) & 1], params, T, it + 1);
            // counter it & 1 counts the estimate of iteration it - 1; the
            // other one was last read before this iteration's first barrier
            // the damping from the MLP on the first wavefront, one hidden
            // unit per lane (the same units as the adjoint's; their sum
            // over the lanes)
            if (t < 64) {
                const float x0 = static_cast<float>(it) / static_cast<float>(T);
                const float x1 = s_ubar[it >= 2 ? it - 2 : 0];
                float d = 0.0f;
                if (use_mlp) {
                    float v = 0.0f;
                    if (t < kMlpHidden)
                        v = mlp[3 * kMlpHidden + t] *
                            tanhf(mlp[2 * t] * x0 + mlp[2 * t + 1] * x1 + mlp[2 * kMlpHidden + t]);
#pragma unroll
                    for (int m = kMlpHidden / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m);
                    d = sigmf(mlp[4 * kMlpHidden] + v);
                }
                if (t == 0) {
                    s_d = d;
                    tape_d[static_cast<size_t>(b) * T + it - 1] = d;
                    tape_u[static_cast<size_t>(b) * T + it - 1] = x1;
                    s_cnt[it & 1] = 0;
                }
            }
            // variable pass, X side (the Z side's channel terms and estimate
            // bits kept for after the X update)
            uint32_t ez_bits = 0;
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float lx3[QJ] = {lam[0][0][cc], lam[0][1][cc], lam[0][2][cc]};
                const float lz3[QJ] = {lam[1][0][cc], lam[1][1][cc], lam[1][2][cc]};
                float mu[QJ];
                qc_mu_side(0, lx3, lz3, q, w, cc, mu);
#pragma unroll
                for (int r = 0; r < QJ; ++r) stf(vslot(vs, r, cc), mu[r]);
                const float sx = (lx3[0] + lx3[1]) + lx3[2];
                const float sz = (lz3[0] + lz3[1]) + lz3[2];
                const float fz = fterm(sx, q.lz);
                ez_bits |= ((sz + fz < 0.0f) ? 1u : 0u) << cc;
                GZ[cc * QP + t] = fz;
            }
            asm volatile("" : "+v"(ez_bits));
            __syncthreads();
            const float d = s_d;
            int unsat = qc_check_rows_fwd(bbase, g.qc_tab, 0, q.synx, chkr, voff, chk_iter(it));
            __syncthreads();
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    const float out = ldsf(vslot(vs, r, cc));
                    lam[0][r][cc] = (1.0f - d) * (w[r * QL + cc] * out) + d * lam[0][r][cc];
                }
            pin_side(lam[0]);
            __syncthreads();
            vs = var_slots(bbase);
            // Z side, from the Z messages (unchanged) and the kept terms
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float sz = (lam[1][0][cc] + lam[1][1][cc]) + lam[1][2][cc];
                const uint32_t ez = (ez_bits >> cc) & 1u;
                const float fz = GZ[cc * QP + t];
#pragma unroll
                for (int r = 0; r < QJ; ++r) {
                    const float a = (sz - lam[1][r][cc]) + fz;
                    stf(vslot(vs, r, cc), with_lsb(w[3 * QNB + r * QL + cc] * a, ez));
                }
            }
            __syncthreads();
            unsat += qc_check_rows_fwd(bbase, g.qc_tab, 1, q.synz, chkr, voff, chk_iter(it));
            __syncthreads();
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    const float out = ldsf(vslot(vs, r, cc));
                    lam[1][r][cc] = (1.0f - d) * (w[QNB + r * QL + cc] * out) + d * lam[1][r][cc];
                }
            pin_side(lam[1]);
            {
                const float wsum = wave_sum(static_cast<float>(unsat));
                if (lane == 0 && wsum != 0.0f) atomicAdd(&s_cnt[it & 1], static_cast<int>(wsum));
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) tape_st(tape, voff, tape_off(it, s, r, cc), lam[s][r][cc]);
            // loss on the marginals after iteration it
            asm volatile("" : "+v"(q.ebx), "+v"(q.ebz));
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float sx = (lam[0][0][cc] + lam[0][1][cc]) + lam[0][2][cc];
                const float sz = (lam[1][0][cc] + lam[1][1][cc]) + lam[1][2][cc];
                const float mx = sx + fterm(sz, q.lx), mz = sz + fterm(sx, q.lz);
                const float kx = static_cast<float>(2 * static_cast<int>((q.ebx >> cc) & 1u) - 1);
                const float kz = static_cast<float>(2 * static_cast<int>((q.ebz >> cc) & 1u) - 1);
                lsum += softplusf(kx * mx) + softplusf(kz * mz);
            }
            asm volatile("" : "+v"(lsum));  // summed here, not carried into the next iteration's passes
            __syncthreads();
            const int cnt = s_cnt[it & 1];  // unsatisfied checks of the estimate of iteration it - 1
            if (t == 0 && it >= 2) s_ubar[it - 1] = static_cast<float>(cnt) / mtot;
            if (it >= 2 && conv == 0 && cnt == 0) conv = it - 1;
        }
        // the estimate after the last iteration: its bits through the exchange
        // buffer, counted by the check owners
        if (conv == 0) {
            int unsat = 0;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const VarSlots vs = var_slots(bbase);
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    const float sx = (lam[0][0][cc] + lam[0][1][cc]) + lam[0][2][cc];
                    const float sz = (lam[1][0][cc] + lam[1][1][cc]) + lam[1][2][cc];
                    const uint32_t e = s == 0 ? ((sx + fterm(sz, q.lx) < 0.0f) ? 1u : 0u)
                                              : ((sz + fterm(sx, q.lz) < 0.0f) ? 1u : 0u);
#pragma unroll
                    for (int r = 0; r < QJ; ++r) stf(vslot(vs, r, cc), __uint_as_float(e));
                }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < QJ; ++r) {
                    uint32_t par = 0;
#pragma unroll
                    for (int i = 0; i < QL; ++i)
                        par ^= __float_as_uint(ldsf(check_slot(bbase, g.qc_tab, s, r, i, t))) & 1u;
                    unsat += (par != (((s == 0 ? q.synx : q.synz) >> r) & 1u)) ? 1 : 0;
                }
                __syncthreads();
            }
            if (t == 0) s_cnt[(T + 1) & 1] = 0;
            __syncthreads();
            const float wsum = wave_sum(static_cast<float>(unsat));
            if (lane == 0 && wsum != 0.0f) atomicAdd(&s_cnt[(T + 1) & 1], static_cast<int>(wsum));
            __syncthreads();
            if (s_cnt[(T + 1) & 1] == 0) conv = T;
        }
        {
            const float ws = wave_sum(lsum);
            if (lane == 0) atomicAdd(&s_loss, ws);
        }
        __syncthreads();
        if (t == 0) {
            atomicAdd(loss, s_loss * scale);
            conv_out[b] = conv;
        }
        __syncthreads();
    }
}

// One side of the adjoint of iteration it (steps 2 to 6 below), with that
// side's gradients g = dL/dlambda_it in registers: on return g holds the
// side's part of dL/dlambda_(it-1) without the sums' cross terms, acc[cc]
// the side's share of its sum's gradient (sum over the column's 3 edges of
// wv * dL/dmu), and gd has the side's share of dL/dd added.
template <int S>
__device__ __forceinline__ void qc_bwd_side(float (&g)[QJ][QL], float (&acc)[QL], float &gd, float (&sx)[QL],
                                            float (&sz)[QL], uint32_t bbase, const int *tab, const QcTrial &q,
                                            tape_rsrc tape, tape_rsrc chkr, uint32_t voff, int it,
                                            const float *w, float *gw_it, float d, int lane) {
    // l: the side's messages before the iteration (tape row it-1), loaded
    // for (4) and again for (6), each issued before the barrier ahead of the
    // pass so that they arrive during it
    float l[QJ][QL];
    auto ld = [&](int r, int cc) { return l[r][cc]; };
    auto reload = [&]() {
        const uint32_t prow = tape_row(it - 1);
#pragma unroll
        for (int r = 0; r < QJ; ++r)
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) l[r][cc] = tape_ld(tape, voff, prow + edge_off(S, r, cc));
    };
    // (3) the check rows' raw outputs, from the forward pass's records of
    // them (the same minima, signs and positions it used)
    uint32_t info[QJ];
    {
        float m1r[QJ], m2r[QJ];
        const uint32_t crow = chk_iter(it);
#pragma unroll
        for (int r = 0; r < QJ; ++r) {
            info[r] = chk_ld(chkr, voff, crow + chk_word(S, r, 0));
            m1r[r] = __uint_as_float(chk_ld(chkr, voff, crow + chk_word(S, r, 1)));
            m2r[r] = __uint_as_float(chk_ld(chkr, voff, crow + chk_word(S, r, 2)));
        }
        int x = threadIdx.x;
        asm volatile("" : "+v"(x));
#pragma unroll
        for (int r = 0; r < QJ; ++r)
#pragma unroll
            for (int i = 0; i < QL; ++i) stf(check_slot(bbase, tab, S, r, i, x), chk_raw(info[r], m1r[r], m2r[r], i));
    }
    reload();
    __syncthreads();
    VarSlots vs = var_slots(bbase);
    // the sums re-enter opaque, so that (6) does not keep values from the
    // caller's passes live through (3) to (5)
    pin_row(sx);
    pin_row(sz);
    // (4) damping and check weights, per edge of the variable owner; the wc
    // gradients of a block row are summed over the wavefront together
    const int wslot = (lane & 3) == 0 ? slot12(static_cast<uint32_t>(lane)) : -1;
#pragma unroll
    for (int r = 0; r < QJ; ++r) {
        float pw[QL];
#pragma unroll
        for (int cc = 0; cc < QL; ++cc) {
            const float raw = ldsf(vslot(vs, r, cc));
            const float mag = fabsf(raw);
            const float cl = mag < kLmaxQ ? mag : kLmaxQ;
            const float out = raw < 0.0f ? -cl : cl;
            const float wc = w[S * QNB + r * QL + cc];
            const float ge = g[r][cc];
            gd += ge * (ld(r, cc) - wc * out);
            // formed here, not sunk to the pass's end with every edge's terms
            asm volatile("" : "+v"(gd));
            const float gtil = (1.0f - d) * ge;
            pw[cc] = gtil * out;
            g[r][cc] = d * ge;
            asm volatile("" : "+v"(g[r][cc]));
            stf(vslot(vs, r, cc), mag > kLmaxQ ? 0.0f : wc * gtil);
        }
        const float sum = reduce12(pw, static_cast<uint32_t>(lane));
        if (wslot >= 0) atomicAdd(&gw_it[S * QNB + r * QL + wslot], sum);
    }
    pin_side(g);
    asm volatile("" : "+v"(gd));  // summed here, not sunk past the next passes with its terms held
    __syncthreads();
    // (5) check adjoint: the gradient of each output goes to the input
    // attaining the minimum over the others
    int x = threadIdx.x;
    asm volatile("" : "+v"(x));
#pragma unroll 1
    for (int r = 0; r < QJ; ++r) {
        const uint32_t inf = r == 0 ? info[0] : (r == 1 ? info[1] : info[2]);
        uint32_t ad[QL];
#pragma unroll
        for (int i = 0; i < QL; ++i) ad[i] = check_slot(bbase, tab, S, r, i, x);
        const int i1 = static_cast<int>(inf & 15u), i2 = static_cast<int>((inf >> 4) & 15u);
        float to1 = 0.0f, to2 = 0.0f;
#pragma unroll
        for (int i = 0; i < QL; ++i) {
            const float ge = ldsf(ad[i]);
            const float se = ((inf >> (20 + i)) & 1u) ? -ge : ge;
            if (i == i1) to2 += se;
            else to1 += se;
        }
#pragma unroll
        for (int i = 0; i < QL; ++i) {
            float o = 0.0f;
            if (i == i1) o = ((inf >> (8 + i)) & 1u) ? -to1 : to1;
            else if (i == i2) o = ((inf >> (8 + i)) & 1u) ? -to2 : to2;
            stf(ad[i], o);
        }
    }
    reload();
    __syncthreads();
    vs = var_slots(bbase);
    // (6) the variable update's adjoint: wv gradients (a block row's summed
    // over the wavefront together) and the own message's share; the sums'
    // shares are the caller's
#pragma unroll
    for (int cc = 0; cc < QL; ++cc) acc[cc] = 0.0f;
#pragma unroll
    for (int r = 0; r < QJ; ++r) {
        float pv[QL];
#pragma unroll
        for (int cc = 0; cc < QL; ++cc) {
            const float f = S == 0 ? fterm(sz[cc], q.lx) : fterm(sx[cc], q.lz);
            const float gmu = ldsf(vslot(vs, r, cc));
            const float a = ((S == 0 ? sx[cc] : sz[cc]) - ld(r, cc)) + f;
            pv[cc] = gmu * a;
            const float ga = w[2 * QNB + S * QNB + r * QL + cc] * gmu;
            acc[cc] += ga;
            g[r][cc] -= ga;
            asm volatile("" : "+v"(g[r][cc]), "+v"(acc[cc]));
        }
        const float sum = reduce12(pv, static_cast<uint32_t>(lane));
        if (wslot >= 0) atomicAdd(&gw_it[2 * QNB + S * QNB + r * QL + wslot], sum);
    }
    pin_side(g);
    pin_row(acc);
    __syncthreads();
}

__global__ void __launch_bounds__(QP) learned_qc_bwd_kernel(DevGraph g, const float *params, int T, int use_mlp,
                                                           TrialPlanes tp, const float *act, uint32_t *chk,
                                                           const float *tape_d,
                                                           const float *tape_u, const double *prior, float *grad,
                                                           int batch, int *counter) {
    extern __shared__ float dyn[];  // B [J][L][P] | weight gradients [T][4][nb]
    float *s_gw = dyn + QNB * QP;
    __shared__ int s_trial;
    __shared__ float s_gd;
    __shared__ float s_gmlp[kMlpParams];
    __shared__ float s_gz[kMaxLearnedT];
    __shared__ float s_w[2][QW];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t bbase = static_cast<uint32_t>((uintptr_t)((lds_f32 *)dyn));
    const float *mlp = params + 4 * QNB * T;
    const size_t row = static_cast<size_t>(QK) * QP;
    const float scale = 1.0f / (static_cast<float>(batch) * static_cast<float>(T) * static_cast<float>(2 * g.n));
    for (int i = t; i < QW * T; i += QP) s_gw[i] = 0.0f;
    for (int i = t; i < kMlpParams; i += QP) s_gmlp[i] = 0.0f;
    if (t == 0) s_gd = 0.0f;
    __syncthreads();
    for (;;) {
        if (t == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_trial);
        __syncthreads();
        if (b >= batch) break;
        QcTrial q = qc_trial(g, tp, prior, b);
        // The tape is walked from row T down; once row k's loss term is
        // read (at iteration k) the row parks the side not in registers:
        // dL/dlambda_(k-1) of the X side (without its pending cross term,
        // held in cross) while the Z side is undone, then the Z side's for
        // iteration k-1.
        const tape_rsrc tape = make_tape(act + static_cast<size_t>(b) * (T + 1) * row);
        const tape_rsrc chkr = make_chk(chk + static_cast<size_t>(b) * T * (kChkIterBytes / 4));
        const uint32_t voff = static_cast<uint32_t>(t) * 4u;
        float cross[QL];  // the X side's pending share of the sums' gradients, per column block
#pragma unroll
        for (int cc = 0; cc < QL; ++cc) cross[cc] = 0.0f;
        // the sums S_x, S_z of each variable (cc, t) over the messages of tape
        // row k: row T's here, then each iteration forms row it-1's, which
        // the next iteration's loss term reads
        float sxc[QL], szc[QL];
        {
            const uint32_t krow = tape_row(T);
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                auto ldc = [&](int s, int r) { return tape_ld(tape, voff, krow + edge_off(s, r, cc)); };
                sxc[cc] = (ldc(0, 0) + ldc(0, 1)) + ldc(0, 2);
                szc[cc] = (ldc(1, 0) + ldc(1, 1)) + ldc(1, 2);
                tape_fence();
            }
        }
        stage_weights(s_w[T & 1], params, T, T);
        __syncthreads();
        for (int it = T; it >= 1; --it) {
            const float *w = s_w[it & 1];
            float *gw_it = s_gw + (it - 1) * QW;  // this iteration's weight gradients, ordered as w
            stage_weights(s_w[(it - 1) & 1], params, T, it - 1);
            const float d = tape_d[static_cast<size_t>(b) * T + it - 1];
            float gd = 0.0f;
            float gv[QJ][QL];  // the resident side's dL/dlambda
            // X side: parked gradient plus its cross term, plus the loss term
            // of iteration it (whose Z part is held for the Z side); row it-1
            // in the same round trip: the X messages (kept for the X side's
            // passes) and the sums, for both sides and the next loss term
            // (at it = T nothing is parked: row T is read and discarded)
            const bool parked = it < T;
            float lr[QJ][QL];
            float zr[QJ][QL];
            {
                const uint32_t krow = tape_row(parked ? it + 1 : it);
                const uint32_t prow = tape_row(it - 1);
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) {
                        gv[r][cc] = tape_ld(tape, voff, krow + edge_off(0, r, cc));
                        lr[r][cc] = tape_ld(tape, voff, prow + edge_off(0, r, cc));
                        zr[r][cc] = tape_ld(tape, voff, prow + edge_off(1, r, cc));
                    }
            }
            // (the X messages and the Z messages serve the sums only)
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) gv[r][cc] = (parked ? gv[r][cc] : 0.0f) + cross[cc];
            float gsz[QL];
            asm volatile("" : "+v"(q.ebx), "+v"(q.ebz));
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float sx = sxc[cc], sz = szc[cc];
                const float mx = sx + fterm(sz, q.lx), mz = sz + fterm(sx, q.lz);
                const float kx = static_cast<float>(2 * static_cast<int>((q.ebx >> cc) & 1u) - 1);
                const float kz = static_cast<float>(2 * static_cast<int>((q.ebz >> cc) & 1u) - 1);
                const float glx = scale * kx * sigmf(kx * mx), glz = scale * kz * sigmf(kz * mz);
                const float gsx = glx + glz * fterm_d(sx, q.lz);
                gsz[cc] = glz + glx * fterm_d(sz, q.lx);
#pragma unroll
                for (int r = 0; r < QJ; ++r) gv[r][cc] += gsx;
            }
            pin_side(gv);
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                sxc[cc] = (lr[0][cc] + lr[1][cc]) + lr[2][cc];
                szc[cc] = (zr[0][cc] + zr[1][cc]) + zr[2][cc];
            }
            // the loss term's Z part waits in row it (free from here on: the
            // loss term read its sums from registers)
            {
                const uint32_t krow = tape_row(it);
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) tape_st(tape, voff, krow + edge_off(1, 0, cc), gsz[cc]);
            }
            float accx[QL];
            qc_bwd_side<0>(gv, accx, gd, sxc, szc, bbase, g.qc_tab, q, tape, chkr, voff, it, w, gw_it, d, lane);
            {
                const uint32_t krow = tape_row(it);
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) tape_st(tape, voff, krow + edge_off(0, r, cc), gv[r][cc]);
            }
            // Z side: parked gradient plus the loss term, in one round trip
            {
                float lz[QL];
                const uint32_t jrow = tape_row(it);
                const uint32_t krow = tape_row(parked ? it + 1 : it);
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) lz[cc] = tape_ld(tape, voff, jrow + edge_off(1, 0, cc));
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) gv[r][cc] = tape_ld(tape, voff, krow + edge_off(1, r, cc));
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) gv[r][cc] = (parked ? gv[r][cc] : 0.0f) + lz[cc];
            }
            pin_side(gv);
            float accz[QL];
            qc_bwd_side<1>(gv, accz, gd, sxc, szc, bbase, g.qc_tab, q, tape, chkr, voff, it, w, gw_it, d, lane);
            // the sums' shares: S_x reaches the X messages directly and the
            // Z side's channel term, S_z likewise
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float fdx = fterm_d(szc[cc], q.lx);  // d a_x / d S_z
                const float fdz = fterm_d(sxc[cc], q.lz);  // d a_z / d S_x
                cross[cc] = accx[cc] + accz[cc] * fdz;
#pragma unroll
                for (int r = 0; r < QJ; ++r) gv[r][cc] += accz[cc] + accx[cc] * fdx;
            }
            {
                const uint32_t krow = tape_row(it);
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) tape_st(tape, voff, krow + edge_off(1, r, cc), gv[r][cc]);
            }
            // (7) the damping's gradient, dL/dz of the MLP output, for the
            // MLP's adjoint after the trial
            if (use_mlp) {
                const float ws = wave_sum(gd);
                if (lane == 0) atomicAdd(&s_gd, ws);
                __syncthreads();
                if (t == 0) {
                    s_gz[it - 1] = s_gd * d * (1.0f - d);
                    s_gd = 0.0f;
                }
            }
            __syncthreads();
        }
        // the MLP's adjoint: lane j of the first wavefront takes iteration
        // j + 1 and the 65 parameter gradients are summed over the lanes
        if (use_mlp && t < 64) {
            const bool on = t < T;
            const float x0 = static_cast<float>(t + 1) / static_cast<float>(T);
            const float x1 = on ? tape_u[static_cast<size_t>(b) * T + t] : 0.0f;
            const float gz = on ? s_gz[t] : 0.0f;
            float h[kMlpHidden];
            mlp_fwd(mlp, x0, x1, h);
            const float v = wave_sum(gz);
            if (t == 0) s_gmlp[4 * kMlpHidden] += v;
            for (int i = 0; i < kMlpHidden; ++i) {
                const float ga = gz * mlp[3 * kMlpHidden + i] * (1.0f - h[i] * h[i]);
                const float g3 = wave_sum(gz * h[i]), g0 = wave_sum(ga * x0), g1 = wave_sum(ga * x1),
                            g2 = wave_sum(ga);
                if (t == 0) {
                    s_gmlp[3 * kMlpHidden + i] += g3;
                    s_gmlp[2 * i] += g0;
                    s_gmlp[2 * i + 1] += g1;
                    s_gmlp[2 * kMlpHidden + i] += g2;
                }
            }
        }
    }
    __syncthreads();
    // the gradient vector is [wc X | wc Z | wv X | wv Z][36][T]: entry
    // kk*T + j is s_gw[j][kk]
    float *gmlp = grad + 4 * QNB * T;
    for (int i = t; i < QW * T; i += QP) {
        const int kk = i / T, j = i - kk * T;
        const float v = s_gw[j * QW + kk];
        if (v != 0.0f) atomicAdd(&grad[i], v);
    }
    for (int i = t; i < kMlpParams; i += QP)
        if (use_mlp && s_gmlp[i] != 0.0f) atomicAdd(&gmlp[i], s_gmlp[i]);
}

// ---- layered min-sum (kind 2): block row r of both sides is layer r ----
//
// Iteration t updates the layers in turn: layer r forms every variable's
// sums from the current messages (rows before r already of iteration t),
// sends mu = wv * (S - lambda + F) on row r of both sides, and replaces row
// r's messages by wc * CN(mu).  The iteration converges when every layer
// starts from the same estimate as layer 0 and every check accepts that
// estimate (jbp_learned.hip's layered kernels; no damping).  The exchange
// buffer holds one layer: B[side][cc][y] for row r (73.7 KB).  Layer 0's
// estimate bits ride in mu's least significant bit at every layer, so each
// layer's check owners count its checks against that estimate.

__device__ __forceinline__ VarSlots var_slots_l(uint32_t bbase) {
    uint32_t a = bbase + static_cast<uint32_t>(threadIdx.x) * 4u;
    asm volatile("" : "+v"(a));
    return {a, a + static_cast<uint32_t>(QL * QP * 4)};
}
// variable (cc, t)'s slot of side s in the layer buffer
__device__ __forceinline__ uint32_t lslot(const VarSlots &v, int s, int cc) {
    return (s == 0 ? v.lo : v.hi) + static_cast<uint32_t>(cc * QP * 4);
}
// check (r, x)'s input i of side s in the layer buffer
__device__ __forceinline__ uint32_t lcheck_slot(uint32_t bbase, const int *tab, int s, int r, int i, int x) {
    constexpr int JL = QJ * QL;
    const int blk = r * QL + i;
    const ctab_i32 *ct = (const ctab_i32 *)tab;
    const uint32_t y = static_cast<uint32_t>(ct[(3 * s) * JL + blk] * x + ct[(3 * s + 1) * JL + blk]);
    const uint32_t q = __umulhi(y, 5592406u);
    return bbase + static_cast<uint32_t>((s * QL + i) * QP * 4) + (y - q * static_cast<uint32_t>(QP)) * 4u;
}

// One check row of the layer (side s): min-sum in place, clamped outputs;
// returns 1 when its inputs' least significant bits disagree with the
// syndrome bit.
__device__ __forceinline__ uint32_t lms_check_row_fwd(uint32_t bbase, const int *tab, int s, int r, int x,
                                                      uint32_t synbit, tape_rsrc chk, uint32_t voff,
                                                      uint32_t crow) {
    float mu[QL];
    uint32_t par = 0, neg = 0, negall = synbit;
    float m1 = HUGE_VALF, m2 = HUGE_VALF;
    int i1 = -1, i2 = -1;
#pragma unroll
    for (int i = 0; i < QL; ++i) mu[i] = ldsf(lcheck_slot(bbase, tab, s, r, i, x));
#pragma unroll
    for (int i = 0; i < QL; ++i) {
        par ^= __float_as_uint(mu[i]) & 1u;
        const float a = fabsf(mu[i]);
        const uint32_t ng = mu[i] < 0.0f ? 1u : 0u;
        neg |= ng << i;
        negall ^= ng;
        const bool lt1 = a < m1, lt2 = a < m2;
        i2 = lt1 ? i1 : (lt2 ? i : i2);  // the first index of the minimum over the others
        m2 = lt1 ? m1 : (lt2 ? a : m2);
        i1 = lt1 ? i : i1;
        m1 = lt1 ? a : m1;
    }
    chk_st(chk, voff, crow + chk_word(s, r, 0), chk_info(i1, i2, neg, negall, m1, m2));
    chk_st(chk, voff, crow + chk_word(s, r, 1), __float_as_uint(m1));
    chk_st(chk, voff, crow + chk_word(s, r, 2), __float_as_uint(m2));
#pragma unroll
    for (int i = 0; i < QL; ++i) {
        const float mag = i == i1 ? m2 : m1;
        const float out = mag < kLmaxQ ? mag : kLmaxQ;
        stf(lcheck_slot(bbase, tab, s, r, i, x), ((negall ^ (neg >> i)) & 1u) ? -out : out);
    }
    return par != synbit ? 1u : 0u;
}

__global__ void __launch_bounds__(QP) learned_qc_lms_fwd_kernel(DevGraph g, const float *params, int T,
                                                               TrialPlanes tp, float *act, uint32_t *chk,
                                                               const double *prior,
                                                               float *loss, int *conv_out, int batch,
                                                               int *counter) {
    extern __shared__ float B[];  // two layer buffers [2][side][L][P]
    __shared__ int s_trial;
    __shared__ float s_loss;
    __shared__ float s_w[2][QW];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t bbase0 = static_cast<uint32_t>((uintptr_t)((lds_f32 *)B));
    const size_t row = static_cast<size_t>(QK) * QP;
    const float scale = 1.0f / (static_cast<float>(batch) * static_cast<float>(T) * static_cast<float>(2 * g.n));
    for (;;) {
        if (t == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_trial);
        __syncthreads();
        if (b >= batch) return;
        QcTrial q = qc_trial(g, tp, prior, b);
        const tape_rsrc tape = make_tape(act + static_cast<size_t>(b) * (T + 1) * row);
        const tape_rsrc chkr = make_chk(chk + static_cast<size_t>(b) * T * (kChkIterBytes / 4));
        const uint32_t voff = static_cast<uint32_t>(t) * 4u;
        float lam[2][QJ][QL];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    lam[s][r][cc] = 0.0f;
                    tape_st(tape, voff, tape_off(0, s, r, cc), 0.0f);
                }
        stage_weights(s_w[1], params, T, 1);
        if (t == 0) s_loss = 0.0f;
        __syncthreads();
        int conv = 0;
        float lsum = 0.0f;
        for (int it = 1; it <= T; ++it) {
            const float *w = s_w[it & 1];
            stage_weights(s_w[(it + 1) & 1], params, T, it + 1);
            uint32_t e0x = 0, e0z = 0;  // layer 0's estimate
            uint32_t flag = 0;          // an unstable estimate or an unsatisfied check
#pragma unroll
            for (int r = 0; r < QJ; ++r) {
                const uint32_t bbase = bbase0 + static_cast<uint32_t>((r & 1) * 2 * QL * QP * 4);
                const VarSlots vs = var_slots_l(bbase);
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    const float sx = (lam[0][0][cc] + lam[0][1][cc]) + lam[0][2][cc];
                    const float sz = (lam[1][0][cc] + lam[1][1][cc]) + lam[1][2][cc];
                    const float fx = fterm(sz, q.lx), fz = fterm(sx, q.lz);
                    const uint32_t ex = (sx + fx < 0.0f) ? 1u : 0u, ez = (sz + fz < 0.0f) ? 1u : 0u;
                    if (r == 0) {
                        e0x |= ex << cc;
                        e0z |= ez << cc;
                    } else {
                        flag |= (ex ^ ((e0x >> cc) & 1u)) | (ez ^ ((e0z >> cc) & 1u));
                    }
                    const float ax = (sx - lam[0][r][cc]) + fx, az = (sz - lam[1][r][cc]) + fz;
                    stf(lslot(vs, 0, cc), with_lsb(w[2 * QNB + r * QL + cc] * ax, (e0x >> cc) & 1u));
                    stf(lslot(vs, 1, cc), with_lsb(w[3 * QNB + r * QL + cc] * az, (e0z >> cc) & 1u));
                }
                asm volatile("" : "+v"(e0x), "+v"(e0z), "+v"(flag));
                __syncthreads();
                {
                    int x = t;
                    asm volatile("" : "+v"(x));
                    const uint32_t crow = chk_iter(it);
                    flag |= lms_check_row_fwd(bbase, g.qc_tab, 0, r, x, (q.synx >> r) & 1u, chkr, voff, crow);
                    flag |= lms_check_row_fwd(bbase, g.qc_tab, 1, r, x, (q.synz >> r) & 1u, chkr, voff, crow);
                }
                __syncthreads();
                const VarSlots vr = var_slots_l(bbase);
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) {
                    lam[0][r][cc] = w[r * QL + cc] * ldsf(lslot(vr, 0, cc));
                    lam[1][r][cc] = w[QNB + r * QL + cc] * ldsf(lslot(vr, 1, cc));
                }
                pin_side(lam[0]);
                pin_side(lam[1]);
            }
            const int any = __syncthreads_or(static_cast<int>(flag));
            if (any == 0 && conv == 0) conv = it;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int r = 0; r < QJ; ++r)
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) tape_st(tape, voff, tape_off(it, s, r, cc), lam[s][r][cc]);
            // loss on the marginals at the end of the iteration
            asm volatile("" : "+v"(q.ebx), "+v"(q.ebz));
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float sx = (lam[0][0][cc] + lam[0][1][cc]) + lam[0][2][cc];
                const float sz = (lam[1][0][cc] + lam[1][1][cc]) + lam[1][2][cc];
                const float mx = sx + fterm(sz, q.lx), mz = sz + fterm(sx, q.lz);
                const float kx = static_cast<float>(2 * static_cast<int>((q.ebx >> cc) & 1u) - 1);
                const float kz = static_cast<float>(2 * static_cast<int>((q.ebz >> cc) & 1u) - 1);
                lsum += softplusf(kx * mx) + softplusf(kz * mz);
            }
            asm volatile("" : "+v"(lsum));
        }
        {
            const float ws = wave_sum(lsum);
            if (lane == 0) atomicAdd(&s_loss, ws);
        }
        __syncthreads();
        if (t == 0) {
            atomicAdd(loss, s_loss * scale);
            conv_out[b] = conv;
        }
        __syncthreads();
    }
}

// The layered adjoint.  Iteration t is undone layer by layer (r = 2, 1, 0)
// with the gradients of all 72 messages in registers.  The state before step
// (t, r) — rows before r from tape row t, the rest from row t-1 — is read
// again for each layer (its sums are kept through the layer's passes, row r's
// own messages read again before the last pass); layer 0's estimate, which
// rode in every mu of the iteration, comes from row t-1's sums.
__global__ void __launch_bounds__(QP) learned_qc_lms_bwd_kernel(DevGraph g, const float *params, int T,
                                                               TrialPlanes tp, float *act, uint32_t *chk,
                                                               const double *prior, float *grad, int batch,
                                                               int *counter) {
    extern __shared__ float dyn[];  // layer buffer [side][L][P] | weight gradients [T][4][nb]
    float *s_gw = dyn + 2 * QL * QP;
    __shared__ int s_trial;
    __shared__ float s_w[2][QW];
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t bbase = static_cast<uint32_t>((uintptr_t)((lds_f32 *)dyn));
    const size_t row = static_cast<size_t>(QK) * QP;
    const float scale = 1.0f / (static_cast<float>(batch) * static_cast<float>(T) * static_cast<float>(2 * g.n));
    for (int i = t; i < QW * T; i += QP) s_gw[i] = 0.0f;
    __syncthreads();
    for (;;) {
        if (t == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(s_trial);
        __syncthreads();
        if (b >= batch) break;
        QcTrial q = qc_trial(g, tp, prior, b);
        const tape_rsrc tape = make_tape(act + static_cast<size_t>(b) * (T + 1) * row);
        const tape_rsrc chkr = make_chk(chk + static_cast<size_t>(b) * T * (kChkIterBytes / 4));
        const uint32_t voff = static_cast<uint32_t>(t) * 4u;
        float G[2][QJ][QL];  // dL / d lambda of the state after the step being undone
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < QJ; ++r)
#pragma unroll
                for (int cc = 0; cc < QL; ++cc) G[s][r][cc] = 0.0f;
        // sums of tape row k for the loss term of iteration k: row T's here,
        // then each iteration's layer 0 forms row t-1's
        float sxc[QL], szc[QL];
        {
            const uint32_t krow = tape_row(T);
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                auto ldc = [&](int s, int r) { return tape_ld(tape, voff, krow + edge_off(s, r, cc)); };
                sxc[cc] = (ldc(0, 0) + ldc(0, 1)) + ldc(0, 2);
                szc[cc] = (ldc(1, 0) + ldc(1, 1)) + ldc(1, 2);
                if ((cc & 3) == 3) tape_fence();
            }
        }
        stage_weights(s_w[T & 1], params, T, T);
        __syncthreads();
        const int wslot = (lane & 3) == 0 ? slot12(static_cast<uint32_t>(lane)) : -1;
        for (int it = T; it >= 1; --it) {
            const float *w = s_w[it & 1];
            float *gw_it = s_gw + (it - 1) * QW;
            stage_weights(s_w[(it - 1) & 1], params, T, it - 1);
            // the loss term at the end of iteration it
            asm volatile("" : "+v"(q.ebx), "+v"(q.ebz));
#pragma unroll
            for (int cc = 0; cc < QL; ++cc) {
                const float sx = sxc[cc], sz = szc[cc];
                const float mx = sx + fterm(sz, q.lx), mz = sz + fterm(sx, q.lz);
                const float kx = static_cast<float>(2 * static_cast<int>((q.ebx >> cc) & 1u) - 1);
                const float kz = static_cast<float>(2 * static_cast<int>((q.ebz >> cc) & 1u) - 1);
                const float glx = scale * kx * sigmf(kx * mx), glz = scale * kz * sigmf(kz * mz);
                const float gsx = glx + glz * fterm_d(sx, q.lz);
                const float gsz = glz + glx * fterm_d(sz, q.lx);
#pragma unroll
                for (int r = 0; r < QJ; ++r) {
                    G[0][r][cc] += gsx;
                    G[1][r][cc] += gsz;
                }
            }
            pin_side(G[0]);
            pin_side(G[1]);
#pragma unroll
            for (int r = QJ - 1; r >= 0; --r) {
                // (2) the sums of the state before step (it, r); (3) row r's
                // raw check outputs from the forward pass's records
                float Sx[QL], Sz[QL];
                uint32_t info0, info1;
                {
                    float m1r[2], m2r[2];
                    const uint32_t hrow = chk_iter(it);
                    info0 = chk_ld(chkr, voff, hrow + chk_word(0, r, 0));
                    m1r[0] = __uint_as_float(chk_ld(chkr, voff, hrow + chk_word(0, r, 1)));
                    m2r[0] = __uint_as_float(chk_ld(chkr, voff, hrow + chk_word(0, r, 2)));
                    info1 = chk_ld(chkr, voff, hrow + chk_word(1, r, 0));
                    m1r[1] = __uint_as_float(chk_ld(chkr, voff, hrow + chk_word(1, r, 1)));
                    m2r[1] = __uint_as_float(chk_ld(chkr, voff, hrow + chk_word(1, r, 2)));
                    const uint32_t crow = tape_row(it), prow = tape_row(it - 1);
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) {
                        float x[2][QJ];
#pragma unroll
                        for (int s = 0; s < 2; ++s)
#pragma unroll
                            for (int rr = 0; rr < QJ; ++rr)
                                x[s][rr] = tape_ld(tape, voff, (rr < r ? crow : prow) + edge_off(s, rr, cc));
                        Sx[cc] = (x[0][0] + x[0][1]) + x[0][2];
                        Sz[cc] = (x[1][0] + x[1][1]) + x[1][2];
                        if ((cc & 3) == 3) tape_fence();
                    }
                    int x = t;
                    asm volatile("" : "+v"(x));
#pragma unroll
                    for (int i = 0; i < QL; ++i) {
                        stf(lcheck_slot(bbase, g.qc_tab, 0, r, i, x), chk_raw(info0, m1r[0], m2r[0], i));
                        stf(lcheck_slot(bbase, g.qc_tab, 1, r, i, x), chk_raw(info1, m1r[1], m2r[1], i));
                    }
                }
                pin_row(Sx);
                pin_row(Sz);
                __syncthreads();
                // (4) check weights: their gradient, and the gradient of the
                // raw check outputs
                {
                    const VarSlots vs = var_slots_l(bbase);
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        float pw[QL];
#pragma unroll
                        for (int cc = 0; cc < QL; ++cc) {
                            const float raw = ldsf(lslot(vs, s, cc));
                            const float mag = fabsf(raw);
                            const float cl = mag < kLmaxQ ? mag : kLmaxQ;
                            const float out = raw < 0.0f ? -cl : cl;
                            const float wc = w[s * QNB + r * QL + cc];
                            const float ge = G[s][r][cc];
                            pw[cc] = ge * out;
                            stf(lslot(vs, s, cc), mag > kLmaxQ ? 0.0f : wc * ge);
                        }
                        const float sum = reduce12(pw, static_cast<uint32_t>(lane));
                        if (wslot >= 0) atomicAdd(&gw_it[s * QNB + r * QL + wslot], sum);
                    }
                }
                __syncthreads();
                // (5) check adjoint
                {
                    int x = t;
                    asm volatile("" : "+v"(x));
#pragma unroll 1
                    for (int s = 0; s < 2; ++s) {
                        const uint32_t inf = s == 0 ? info0 : info1;
                        uint32_t ad[QL];
#pragma unroll
                        for (int i = 0; i < QL; ++i) ad[i] = lcheck_slot(bbase, g.qc_tab, s, r, i, x);
                        const int i1 = static_cast<int>(inf & 15u), i2 = static_cast<int>((inf >> 4) & 15u);
                        float to1 = 0.0f, to2 = 0.0f;
#pragma unroll
                        for (int i = 0; i < QL; ++i) {
                            const float ge = ldsf(ad[i]);
                            const float se = ((inf >> (20 + i)) & 1u) ? -ge : ge;
                            if (i == i1) to2 += se;
                            else to1 += se;
                        }
#pragma unroll
                        for (int i = 0; i < QL; ++i) {
                            float o = 0.0f;
                            if (i == i1) o = ((inf >> (8 + i)) & 1u) ? -to1 : to1;
                            else if (i == i2) o = ((inf >> (8 + i)) & 1u) ? -to2 : to2;
                            stf(ad[i], o);
                        }
                    }
                }
                // row r's own messages before the step, for a (arriving during
                // the barrier)
                float pl[2][QL];
                {
                    const uint32_t prow = tape_row(it - 1);
#pragma unroll
                    for (int s = 0; s < 2; ++s)
#pragma unroll
                        for (int cc = 0; cc < QL; ++cc) pl[s][cc] = tape_ld(tape, voff, prow + edge_off(s, r, cc));
                }
                __syncthreads();
                // (6) the variable side: wv gradients; the gradient of the state
                // before the step (row r's messages were replaced: their
                // gradient restarts from the sums' shares less their own share
                // of a; the other rows gain the sums' shares)
                {
                    const VarSlots vs = var_slots_l(bbase);
                    // wv gradients, one side at a time
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        float pv[QL];
#pragma unroll
                        for (int cc = 0; cc < QL; ++cc) {
                            const float gm = ldsf(lslot(vs, s, cc));
                            const float f = s == 0 ? fterm(Sz[cc], q.lx) : fterm(Sx[cc], q.lz);
                            const float a = ((s == 0 ? Sx[cc] : Sz[cc]) - pl[s][cc]) + f;
                            pv[cc] = gm * a;
                        }
                        const float sum = reduce12(pv, static_cast<uint32_t>(lane));
                        if (wslot >= 0) atomicAdd(&gw_it[2 * QNB + s * QNB + r * QL + wslot], sum);
                    }
                    // the gradient of the state before the step
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) {
                        const float gmx = ldsf(lslot(vs, 0, cc)), gmz = ldsf(lslot(vs, 1, cc));
                        const float gax = w[2 * QNB + r * QL + cc] * gmx, gaz = w[3 * QNB + r * QL + cc] * gmz;
                        const float dsx = gax + gaz * fterm_d(Sx[cc], q.lz);  // d S_x: a_x, and a_z through F_z
                        const float dsz = gaz + gax * fterm_d(Sz[cc], q.lx);
#pragma unroll
                        for (int rr = 0; rr < QJ; ++rr) {
                            G[0][rr][cc] = rr == r ? dsx - gax : G[0][rr][cc] + dsx;
                            G[1][rr][cc] = rr == r ? dsz - gaz : G[1][rr][cc] + dsz;
                        }
                    }
                }
                pin_side(G[0]);
                pin_side(G[1]);
                if (r == 0) {
#pragma unroll
                    for (int cc = 0; cc < QL; ++cc) {
                        sxc[cc] = Sx[cc];
                        szc[cc] = Sz[cc];
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    for (int i = t; i < QW * T; i += QP) {
        const int kk = i / T, j = i - kk * T;
        const float v = s_gw[j * QW + kk];
        if (v != 0.0f) atomicAdd(&grad[i], v);
    }
}

}  // namespace

bool learned_qc_train_supported(const DevGraph &g, const LearnedDevParams &lp) {
    const bool kind_ok = lp.kind == kLearnedMS || (lp.kind == kLearnedLMS && lp.lpb == QL && !lp.use_mlp);
    return kind_ok && g.qc_tab && g.qc_P == QP && g.qc_J == QJ && g.qc_L == QL && lp.nb == QNB && lp.T >= 1 &&
           lp.T <= kMaxLearnedT;
}

size_t learned_qc_chk_bytes(int batch, int T) {
    return static_cast<size_t>(batch) * static_cast<size_t>(T) * kChkIterBytes;
}

void launch_learned_train_qc(const DevGraph &g, const LearnedDevParams &lp, const float *params, TrialPlanes tp,
                             int batch, float *act, uint32_t *chk, float *tape_d, float *tape_u,
                             const double *prior, float *loss, float *grad, int *conv, int *counters, bool backward,
                             hipStream_t st) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = std::max(1, std::min(batch, cus > 0 ? cus : 256));
    const size_t fwd_lds = static_cast<size_t>(QNB + QL) * QP * sizeof(float);
    const size_t bwd_lds = static_cast<size_t>(QNB) * QP * sizeof(float) +
                           static_cast<size_t>(4) * QNB * lp.T * sizeof(float);
    launch_fill_i32(counters, 0, 2, st);
    if (lp.kind == kLearnedLMS) {
        const size_t lfwd_lds = static_cast<size_t>(2) * 2 * QL * QP * sizeof(float);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(learned_qc_lms_fwd_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lfwd_lds));
        hipLaunchKernelGGL(learned_qc_lms_fwd_kernel, dim3(grid), dim3(QP), lfwd_lds, st, g, params, lp.T, tp, act,
                           chk, prior, loss, conv, batch, counters);
        if (!backward) return;
        const size_t lbwd_lds = static_cast<size_t>(2) * QL * QP * sizeof(float) +
                                static_cast<size_t>(QW) * lp.T * sizeof(float);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(learned_qc_lms_bwd_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lbwd_lds));
        hipLaunchKernelGGL(learned_qc_lms_bwd_kernel, dim3(grid), dim3(QP), lbwd_lds, st, g, params, lp.T, tp, act,
                           chk, prior, grad, batch, counters + 1);
        return;
    }
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(learned_qc_fwd_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(fwd_lds));
    hipLaunchKernelGGL(learned_qc_fwd_kernel, dim3(grid), dim3(QP), fwd_lds, st, g, params, lp.T, lp.use_mlp, tp,
                       act, chk, tape_d, tape_u, prior, loss, conv, batch, counters);
    if (!backward) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(learned_qc_bwd_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bwd_lds));
    hipLaunchKernelGGL(learned_qc_bwd_kernel, dim3(grid), dim3(QP), bwd_lds, st, g, params, lp.T, lp.use_mlp, tp,
                       act, chk, tape_d, tape_u, prior, grad, batch, counters + 1);
}

}  // namespace jbp
