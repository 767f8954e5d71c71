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
#include "jbp_learned.h"

namespace jbp {

// Largest dense basis (bytes per side) built for the device residual test.
constexpr size_t kMaxBasisBytes = size_t(256) << 20;
constexpr int kDefaultChunk = 65536;

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
    // chunk buffers
    int chunk_cap = 0;
    jbp::TrialPlanes tp;
    std::vector<void *> chunk_allocs;
    double *f64_scratch = nullptr;
    int f64_slots = 0;
    int f64_block = 256;
    int *d_counter = nullptr;
    int *d_states = nullptr;  // unpack buffer
    size_t d_states_cap = 0;
    int *d_ints = nullptr;    // pack input buffer
    size_t d_ints_cap = 0;
    double last_ms[3] = {0, 0, 0};
    int64_t launches[3] = {0, 0, 0};
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
int ensure_states(jbp_ctx *ctx, size_t elems, char *err, size_t errlen);
int ensure_ints(jbp_ctx *ctx, size_t elems, char *err, size_t errlen);
Prior make_prior(const double p4[4]);
Prior prior_from_p(double p);
// Decode `count` trials whose syndrome planes are already in ctx->tp; records
// the decoder launch time (HIP events) in ctx->last_ms[1].
int run_decode(jbp_ctx *ctx, int count, const Prior &prior, int max_iter, int warmup, int interval,
               double damping, int precision, char *err, size_t errlen,
               const LearnedInfer *learned = nullptr);

int simulate_impl(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_trial_out *pt, jbp_run_stats *stats,
                  const LearnedInfer *learned, char *err, size_t errlen);

}  // namespace jbp
