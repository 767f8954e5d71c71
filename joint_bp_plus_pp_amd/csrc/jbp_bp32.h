// FP32 joint-BP arithmetic shared by the LDS-resident permutation-block
// decoder (jbp_decode_f32.hip) and the generic HBM-resident decoder
// (jbp_decode_gen.hip).  Messages are binary log-likelihood ratios in base 2
// (bits); the variable and check updates they implement are derived in the
// header comment of jbp_decode_f32.hip.
#pragma once

#include "jbp_common.h"

namespace jbp {

constexpr float kLlrMax = 64.0f;  // clamp of |lambda| in bits

struct PriorF {
    float lp[4];  // log2 prior (I, X, Z, Y)
    float fx[4];  // (pi_I, pi_Z, pi_X, pi_Y) for the X-side extrinsic
    float fz[4];  // (pi_I, pi_X, pi_Z, pi_Y) for the Z-side extrinsic
};

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }

// log2((a 2^L + b) / (c 2^L + d)) without overflow for any L.
__device__ __forceinline__ float extrinsic_term(float L, const float *k) {
    float t = fast_exp2(-fabsf(L));
    float num, den;
    if (L >= 0.0f) {
        num = k[0] + k[1] * t;
        den = k[2] + k[3] * t;
    } else {
        num = k[0] * t + k[1];
        den = k[2] * t + k[3];
    }
    return fast_log2(num) - fast_log2(den);
}

// Hard decision over (I, X, Z, Y) from the summed check messages: strict >,
// ties to the lower state (jointbp_ets.cpp:2475-2481).
__device__ __forceinline__ int hard_decision(float lx, float lz, const PriorF &pr) {
    float l0 = pr.lp[0] + lx + lz, l1 = pr.lp[1] + lz, l2 = pr.lp[2] + lx, l3 = pr.lp[3];
    int best = 0;
    float bv = l0;
    if (l1 > bv) { best = 1; bv = l1; }
    if (l2 > bv) { best = 2; bv = l2; }
    if (l3 > bv) { best = 3; }
    return best;
}

}  // namespace jbp
