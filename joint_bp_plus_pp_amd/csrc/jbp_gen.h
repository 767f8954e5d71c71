// Generic FP32 decoder state (jbp_decode_gen.hip): S = 64*groups resident
// trial slots, every array stored [group][item][64 lanes] or as one 64-bit
// word per (group, item).
#pragma once

#include "jbp_common.h"

namespace jbp {

// Words per group of the unsatisfied-slot mask (spread to limit contention
// of the check pass's atomics; one per lane of the book pass).
constexpr int kBadWays = 64;

struct GenState {
    int groups = 0;
    float *lam[2] = {nullptr, nullptr};      // [groups][E_side + 1][64] check messages (edge E: 0)
    float *gb[2] = {nullptr, nullptr};       // [groups][n][64] G_x, G_z per variable
    uint64_t *estw[2] = {nullptr, nullptr};  // [groups][n] hard-decision bits
    uint64_t *synw[2] = {nullptr, nullptr};  // [groups][m_side] syndrome bits
    uint64_t *bad = nullptr;     // [groups][kBadWays] slots with an unsatisfied check (OR of the ways)
    uint64_t *active = nullptr;  // [groups] slots holding a trial (bad..refill contiguous)
    uint64_t *fresh = nullptr;   // [groups] slots whose messages read as 0 this step
    uint64_t *refill = nullptr;  // [groups] slots given a trial by the last book pass
    int *slot_trial = nullptr;   // [S] chunk index of the slot's trial
    int *slot_iter = nullptr;    // [S] iteration index of the slot's trial
    float *prior32 = nullptr;    // [groups][kPrior32][64]
    int2 *fin = nullptr;         // [S] (trial, slot) stopped this step
    int *ctrl = nullptr;         // [0] next trial, [1] trials finished, [2] fin entries
    // damping (the reference's scaled_mix): each side's previous
    // variable-to-check messages, [groups][E_side + 1][64]; null when off
    float *mprev[2] = {nullptr, nullptr};
    float damping = 0.0f;
};

bool gen_supported(const DevGraph &g);
size_t gen_carve_bytes(const DevGraph &g, int groups, bool damped = false);
void gen_carve(const DevGraph &g, int groups, void *base, GenState &s, bool damped = false);

// Decode trials 0..count-1 of tp (syndrome planes in place) to est planes,
// iterations and syn_ok.  h_ctrl: pinned host ints [2]; evs: two events.
// Step batches are replayed from a captured graph when use_graph.
// Returns 0 or a hipError_t value, -1 when the step bound is exceeded.
int gen_decode(const DevGraph &g, const GenState &s, TrialPlanes tp, int count, const NoiseTab &nt,
               int max_iter, int warmup, int interval, int *h_ctrl, hipEvent_t *evs,
               hipStream_t st, int64_t *steps_out, bool use_graph);

// One code's share of a mixed engine: its graph, its own resident groups and
// state, its trials and their channel table.  gen_decode_mix fills the
// launch geometry (group0 ..).
struct GenCode {
    DevGraph g;
    GenState s;
    TrialPlanes tp;
    NoiseTab nt;
    int count = 0, max_iter = 0, warmup = 0, interval = 1;
    int group0 = 0;                 // first engine group of this code
    int vx = 0;                     // variable-pass blocks per group
    int var_b0 = 0, chk_b0 = 0;     // first variable-pass and check-pass block
    int cvar = 0, cpw = 4;          // check-pass variant (mixed kernels), checks per wavefront
    int nch0 = 0, nch = 0;          // check chunks of side 0, of both sides
    int ex = 0, lx = 0, post0 = 0;  // post-pass emit and load blocks, first block
};
constexpr int kMaxMixCodes = 2;
struct GenMix {
    GenCode c[kMaxMixCodes];
    int n = 0;
};

// Decode every code's trials 0..count-1 in one engine: each step's kernels
// cover all codes (launched, not graph-replayed).  h_ctrl: pinned host ints
// [2][kMaxMixCodes]; evs: two events.  Returns 0, a hipError_t value, -1
// when the step bound is exceeded, or -2 for an empty or oversized mix.
int gen_decode_mix(GenMix &m, int *h_ctrl, hipEvent_t *evs, hipStream_t st, int64_t *steps_out);

}  // namespace jbp
