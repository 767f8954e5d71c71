// Private host-side definitions of libjbp shared by jbp_api.hip and
// jbp_learned_api.hip: the context object and chunk-buffer helpers.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/jbp.h"
#include "jbp_common.h"
#include "jbp_gen.h"
#include "jbp_learned.h"
#include "jbp_pp.h"

namespace jbp {

// Largest dense basis (bytes per side) built for the device residual test.
constexpr size_t kMaxBasisBytes = size_t(256) << 20;
// Largest host elimination (rows x rows x 64-bit words) done at create.
constexpr double kMaxBasisWork = 2e10;
constexpr int kDefaultChunk = 65536;
// Trials per persistent launch of the LDS-resident decoder when a run is
// pipelined.
constexpr int kPipeChunk = 8192;
// Chunks in flight in a pipelined run (buffer sets).
constexpr int kPipeDepth = 8;
// Iterations after which a pipelined plain-BP launch hands a trial over to
// the finishing stream (rounded up to a scheduled syndrome test).
constexpr int kStragglerIter = 32;
// Fewest chunks for which a pipelined plain-BP run offloads stragglers.
constexpr int kOffloadChunks = 4;

// One chunk's buffer set in the pipeline: its own trial planes, decoder
// counter, events and pinned outcome arrays.
struct PipeBuf {
    int cap = 0;
    TrialPlanes tp;
    std::vector<void *> allocs;
    int *counter = nullptr;
    // [0] sampling start, [1] decode start, [2] decode end (fallback
    // included), [3] outcomes copied, [4] first launch of the decode ended,
    // [5] sampled and ordered
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int *h_iters = nullptr;   // pinned
    uint8_t *h_flags = nullptr;  // pinned: ok | success | exact, cap each
    int *keys = nullptr;      // decode-order scratch
    int *order = nullptr;
    int *fb = nullptr;        // min-sum decoder's fallback list (cap + 4 ints)
};
constexpr int kMaxPList = JBP_MAX_P_LIST;

// Device buffers of the post-processing stage (jbp_pp_stage.hip).
struct PPBuf {
    int cap = 0;            // trial slots
    int rec_cap = 0;        // iteration records per slot
    size_t stride = 0;      // bytes of FP64 state per slot
    char *state = nullptr;
    int *list = nullptr;        // slots advanced by a launch
    int *slot_trial = nullptr;  // chunk index of each slot's trial
    int *rec = nullptr;         // [slot][iteration][4]
    uint8_t *out = nullptr;     // [slot][2][n] estimate, previous estimate
    int *idx = nullptr;         // slots whose marginals are gathered
    void *stage = nullptr;      // gather / write-back staging
    size_t stage_bytes = 0;
    // pinned host copies (fast transfers): records, estimates, marginals,
    // staging
    int *h_rec = nullptr;
    uint8_t *h_out = nullptr;
    double *h_tot = nullptr;
    void *h_stage = nullptr;
    void *d_hstage = nullptr;  // h_stage's device address (kernels read it directly)
    // pinned index lists copied to the device (slots whose marginals are
    // gathered, slots a step advances, trial of each slot): 3 x cap ints.
    // Copies from pinned memory read their source when they run, so a list
    // stays here (owned by the buffer set, not by a local vector) until the
    // stream has passed the copy.
    int *h_list = nullptr;
    // marginals reach the host in up to kTotChunks pieces, one event each
    static constexpr int kTotChunks = 16;
    hipEvent_t tot_ev[kTotChunks] = {};
    hipEvent_t ready_ev = nullptr;  // a batch's first device phase and its copies done
};

// What the post-processing stage decided for one trial.
struct PPOutcome {
    bool bp_ok = false;  // BP itself matched the syndrome
    bool ok = false;     // syndrome matched in the end (BP or a repair)
    int iterations = 0;
    // repair inside the decode loop
    bool pp_success = false, pp_ets = false, pp_flip = false;
    bool used_any = false, used6 = false, used12 = false;
    // repair after the decode
    bool post_run = false;
    pp::PostOut post;
};

void set_err(char *err, size_t len, const char *fmt, ...);

// Host-side fault injection for CPU-only tests of the caller's fallback:
// JBP_FAULT_INJECT=create|decode|simulate makes that entry point fail.
bool fault(const char *what);

struct HostSide {
    int m = 0, E = 0;
    std::vector<int> chk_off, edge_var, var_off, var_edge;
};

}  // namespace jbp

struct jbp_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n = 0;
    jbp::HostSide hs[2];
    jbp::DevGraph dg;
    std::vector<void *> allocs;  // graph arrays, basis, tables
    bool qc = false;
    jbp::QcParams qcp;
    int *d_qc_tables = nullptr;
    bool have_basis = false;
    jbp::DevBasis basis[2];
    // logical operators (jbp_set_logicals); used instead of the basis when
    // both sides are set
    bool have_logicals = false;
    bool lg_set[2] = {false, false};
    jbp::DevLogicals lg[2];
    void *lg_mem[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    // chunk buffers
    int chunk_cap = 0;
    jbp::TrialPlanes tp;
    std::vector<void *> chunk_allocs;
    int *h_iters = nullptr;     // pinned [chunk_cap]: outcome copies of a chunk
    uint8_t *h_flags = nullptr; // pinned [3][chunk_cap]: syn_ok, success, exact
    hipEvent_t evt[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // sample / decode start, end; [4] learned-phase end
    double *f64_scratch = nullptr;
    int f64_slots = 0;
    int f64_block = 256;
    int *d_counter = nullptr;
    int *d_keys = nullptr;    // decode order of the chunk (chunk_cap each)
    uint64_t *d_mt = nullptr; // sampler generator states, mt_cap x 312 words
    int mt_cap = 0;
    int *d_order = nullptr;
    int *d_fb = nullptr;      // min-sum decoder's fallback list (chunk_cap + 4)
    int *d_states = nullptr;  // unpack buffer
    size_t d_states_cap = 0;
    int *d_ints = nullptr;    // pack input buffer
    size_t d_ints_cap = 0;
    // channel tables (set_noise)
    jbp::NoiseTab noise;
    int noise_cap = 0;
    jbp::DrawThresholds *d_th = nullptr;
    double *d_prior = nullptr;
    float *d_prior32 = nullptr;
    std::vector<double> noise_key;  // what the tables hold (skip re-uploads)
    // generic (HBM) FP32 decoder
    jbp::GenState gen;
    void *gen_mem = nullptr;
    size_t gen_cap = 0;
    int gen_slots = 0;        // default resident trials (0: not chosen yet)
    int *h_ctrl = nullptr;    // pinned [2]
    int *h_ctrl_mix = nullptr;  // pinned [2][kMaxMixCodes] (mixed runs led by this context)
    hipEvent_t gen_ev[2] = {nullptr, nullptr};
    int64_t gen_steps = 0;    // engine steps of the last generic decode
    // single-trial decode replayed from a captured graph (use_graph)
    hipGraphExec_t g1_exec = nullptr;
    // what the captured graph baked in: parameters and every device pointer
    // it uses (buffers that a later call may reallocate)
    std::vector<uintptr_t> g1_key;
    uint32_t *h_stage = nullptr;   // pinned: syndrome words in, outcome out
    jbp::PipeBuf pipe[jbp::kPipeDepth];
    // pipelined runs: sampling goes on `stream`, decoder launches on
    // pipe_st[0]; fallback launches, the residual test and outcome copies on
    // pipe_st[1]
    hipStream_t pipe_st[2] = {nullptr, nullptr};
    // post-processing: host graph, trapping-set tables, device buffers
    jbp::pp::Graph pp_graph;
    jbp::pp::Tables pp_tables;
    // two buffer sets and a second stream: one batch's host repairs overlap
    // the next batch's first device phase
    jbp::PPBuf ppb[2];
    hipStream_t pp_stream = nullptr;
    int *pp_idx = nullptr;  // the stage's slot list 0..batch-1, then its trial indices
    size_t pp_idx_cap = 0;
    double last_ms[5] = {0, 0, 0, 0, 0};
    int64_t launches[5] = {0, 0, 0, 0, 0};
};

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            jbp::set_err(err, errlen, "%s failed: %s", #call, hipGetErrorString(e_));   \
            return JBP_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

namespace jbp {

void free_chunk(jbp_ctx *ctx);
int ensure_chunk(jbp_ctx *ctx, int count, char *err, size_t errlen);
// the sampler's generator-state scratch (ctx->d_mt), for count trials or
// the largest chunk, whichever is smaller
int ensure_mt(jbp_ctx *ctx, int count, char *err, size_t errlen);
int ensure_states(jbp_ctx *ctx, size_t elems, char *err, size_t errlen);
int ensure_ints(jbp_ctx *ctx, size_t elems, char *err, size_t errlen);
Prior make_prior(const double p4[4]);
Prior prior_from_p(double p);
// Upload n channel entries into ctx->noise (phase 0): priors[i], and the
// sampler thresholds of p[i] when p is not NULL.
int set_noise(jbp_ctx *ctx, const Prior *priors, const double *p, int n, char *err, size_t errlen);
void free_noise(jbp_ctx *ctx);
// Decode `count` trials whose syndrome planes are already in ctx->tp, slot i
// with channel entry nt.entry(i); records the decoder launch time (HIP
// events) in ctx->last_ms[1].  defer_timing: record the events
// ctx->evt[2], evt[3] around the launch and return without waiting; the
// caller reads them after its own synchronisation.
int run_decode(jbp_ctx *ctx, int count, const NoiseTab &nt, int max_iter, int warmup, int interval,
               double damping, int precision, char *err, size_t errlen,
               const LearnedInfer *learned = nullptr, int engine = JBP_ENGINE_AUTO,
               int slots = 0, bool use_graph = false, bool defer_timing = false,
               bool freeze = false);
void free_gen(jbp_ctx *ctx);
void free_logicals(jbp_ctx *ctx);
void free_pipe(jbp_ctx *ctx);
void free_pp(jbp_ctx *ctx);
// Post-processing stage over the chunk trials `trials` (their syndrome
// planes in ctx->tp), which the first decode left unconverged after
// min(99, max_iter - 1) iterations: finishes their decode with the
// reference's PP and writes estimate planes, iterations and syndrome flags
// back into ctx->tp; outs[i] describes trials[i].
int run_pp(jbp_ctx *ctx, const NoiseTab &nt, int max_iter, double damping, int hist_window,
           const std::vector<int> &trials, std::vector<PPOutcome> &outs, char *err, size_t errlen);

// Looks up this device's address of the allocation-free decode's scratch
// array (jbp_noalloc.hip), once per device; the device must be current.
int noalloc_setup(int device, char *err, size_t errlen);

int simulate_impl(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_trial_out *pt, jbp_run_stats *stats,
                  const LearnedInfer *learned, char *err, size_t errlen);

}  // namespace jbp
