// The reference program's GPU backend (the three functions declared in its
// jointbp_cuda.h:48-63, called at jointbp_ets.cpp:5420, 5499, 5957) on top
// of libjbp's C ABI (include/jbp.h).  A maintainer keeps jointbp_ets.cpp and
// jointbp_cuda.h as they are, drops this file in place of jointbp_cuda.cu and
// builds with -DUSE_CUDA, linking libjbp.so:
//
//   g++ -O2 -std=c++17 -DUSE_CUDA -I<reference> -I<repo>/include \
//       <reference>/jointbp_ets.cpp jointbp_cuda_jbp.cpp \
//       -L<repo>/joint_bp_plus_pp_amd -ljbp -o jointbp_ets
//
// Decodes run in FP64, which reproduces the reference's CPU decoder trial
// for trial (the reference's own CUDA build also defaults to double
// messages); -DUSE_CUDA_FP32, the reference's switch for float messages,
// selects libjbp's FP32 decoder instead.  Errors come back as the caller
// expects them: create returns null with a message (the caller falls back
// to its CPU decoder), decode returns false for what the device does not
// run (the caller falls back for that trial).
#include <string>
#include <vector>

#include "jbp.h"
#include "jointbp_cuda.h"

struct CudaBPContext {
    jbp_ctx *ctx = nullptr;
    int nvars = 0;
};

CudaBPContext *cuda_bp_create(const CudaBPGraph &g, int device_id, std::string *error_out) {
    jbp_graph_desc d = {g.nvars,
                        g.mX,
                        g.mZ,
                        g.x_edges,
                        g.z_edges,
                        g.x_check_offsets.data(),
                        g.x_check_edges.data(),
                        g.x_edge_var.data(),
                        g.x_edge_check.data(),
                        g.x_edge_pos.data(),
                        g.x_var_offsets.data(),
                        g.x_var_edges.data(),
                        g.z_check_offsets.data(),
                        g.z_check_edges.data(),
                        g.z_edge_var.data(),
                        g.z_edge_check.data(),
                        g.z_edge_pos.data(),
                        g.z_var_offsets.data(),
                        g.z_var_edges.data()};
    char err[512] = {0};
    jbp_ctx *c = nullptr;
    if (jbp_create(&d, device_id, &c, err, sizeof err) != JBP_OK) {
        if (error_out) *error_out = err;
        return nullptr;
    }
    auto *out = new CudaBPContext;
    out->ctx = c;
    out->nvars = g.nvars;
    return out;
}

void cuda_bp_destroy(CudaBPContext *c) {
    if (!c) return;
    jbp_destroy(c->ctx);
    delete c;
}

bool cuda_bp_decode(CudaBPContext *c, const std::vector<int> &sx, const std::vector<int> &sz,
                    const CudaMsg &prior, int max_iter, int check_warmup, int check_interval,
                    bool measure_costs, bool use_graph, bool freeze_syn, double damping,
                    CudaBPResult &out, std::string *error_out) {
#ifdef USE_CUDA_FP32
    const int precision = JBP_FP32;
#else
    const int precision = JBP_FP64;
#endif
    jbp_decode_result r;
    char err[512] = {0};
    out.est.assign(c->nvars, 0);
    const int rc = jbp_decode(c->ctx, sx.data(), sz.data(), prior.v, max_iter, check_warmup,
                              check_interval, measure_costs ? 1 : 0, use_graph ? 1 : 0,
                              freeze_syn ? 1 : 0, damping, precision, &r, out.est.data(), err,
                              sizeof err);
    if (rc != JBP_OK) {
        if (error_out) *error_out = err;
        return false;
    }
    out.iterations = r.iterations;
    out.check_count = r.check_count;
    out.syndrome_match = r.syndrome_match != 0;
    out.kernel_ms = r.kernel_ms;
    out.memcpy_ms = r.memcpy_ms;
    out.host_ms = r.host_ms;
    return true;
}
