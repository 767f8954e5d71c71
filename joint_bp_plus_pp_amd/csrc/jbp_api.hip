// Host side of libjbp: the C ABI declared in include/jbp.h.
//
// Replaces the reference's CUDA backend host code (jointbp_cuda.cu:518-1104:
// context, uploads, per-trial decode loop) with a batched design: a call
// hands the device a whole range of trials; sampling, syndromes, decoding
// and the stabilizer-residual test all run there, and only per-trial
// outcomes (iterations, flags) come back.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "jbp_ctx.h"

using namespace jbp;

namespace jbp {

void set_err(char *err, size_t len, const char *fmt, ...) {
    if (!err || !len) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
}

bool fault(const char *what) {
    const char *v = std::getenv("JBP_FAULT_INJECT");
    return v && std::strstr(v, what) != nullptr;
}

}  // namespace jbp

namespace {


template <typename T>
int dev_upload(jbp_ctx *ctx, const std::vector<T> &h, const T **out, char *err, size_t errlen) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
    HIPCHK(hipMalloc(&p, bytes));
    ctx->allocs.push_back(p);
    if (!h.empty()) HIPCHK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(p);
    return JBP_OK;
}

int validate_side(const char *name, int nvars, int m, int E, const int *chk_off, const int *edge_var,
                  const int *var_off, const int *var_edge, HostSide &hs, char *err,
                  size_t errlen) {
    if (m < 0 || E < 0 || !chk_off || !edge_var || !var_off || !var_edge) {
        set_err(err, errlen, "%s graph arrays missing", name);
        return JBP_ERR_ARG;
    }
    hs.m = m;
    hs.E = E;
    hs.chk_off.assign(chk_off, chk_off + m + 1);
    hs.edge_var.assign(edge_var, edge_var + E);
    hs.var_off.assign(var_off, var_off + nvars + 1);
    hs.var_edge.assign(var_edge, var_edge + E);
    if (hs.chk_off[0] != 0 || hs.chk_off[m] != E || hs.var_off[0] != 0 || hs.var_off[nvars] != E) {
        set_err(err, errlen, "%s offsets do not span the %d edges", name, E);
        return JBP_ERR_ARG;
    }
    for (int c = 0; c < m; ++c)
        if (hs.chk_off[c + 1] < hs.chk_off[c]) {
            set_err(err, errlen, "%s check offsets not monotone at %d", name, c);
            return JBP_ERR_ARG;
        }
    std::vector<int> seen(E, 0);
    for (int v = 0; v < nvars; ++v) {
        if (hs.var_off[v + 1] < hs.var_off[v]) {
            set_err(err, errlen, "%s var offsets not monotone at %d", name, v);
            return JBP_ERR_ARG;
        }
        for (int k = hs.var_off[v]; k < hs.var_off[v + 1]; ++k) {
            int e = hs.var_edge[k];
            if (e < 0 || e >= E || hs.edge_var[e] != v || seen[e]++) {
                set_err(err, errlen, "%s var_edges inconsistent at var %d", name, v);
                return JBP_ERR_ARG;
            }
        }
    }
    for (int e = 0; e < E; ++e)
        if (hs.edge_var[e] < 0 || hs.edge_var[e] >= nvars) {
            set_err(err, errlen, "%s edge %d has variable out of range", name, e);
            return JBP_ERR_ARG;
        }
    return JBP_OK;
}

long long gcdll(long long a, long long b) {
    while (b) { long long t = a % b; a = b; b = t; }
    return a < 0 ? -a : a;
}

long long inv_mod(long long a, long long m) {
    long long g = m, x = 0, x1 = 1, aa = ((a % m) + m) % m;
    while (aa) {
        long long q = g / aa;
        long long t = g - q * aa; g = aa; aa = t;
        t = x - q * x1; x = x1; x1 = t;
    }
    return g == 1 ? ((x % m) + m) % m : -1;
}

// Detect the affine permutation-block structure of both sides from the CSR.
bool detect_qc(const HostSide hs[2], int n, QcParams &qc) {
    if (hs[0].m == 0 || hs[0].m != hs[1].m) return false;
    int L = hs[0].chk_off[1] - hs[0].chk_off[0];
    if (L <= 0 || n % L != 0) return false;
    int P = n / L;
    if (P < 2 || hs[0].m % P != 0) return false;
    int J = hs[0].m / P;
    if (J * L > kQcMaxBlocks) return false;
    for (int s = 0; s < 2; ++s) {
        const HostSide &h = hs[s];
        for (int c = 0; c < h.m; ++c)
            if (h.chk_off[c + 1] - h.chk_off[c] != L) return false;
        for (int r = 0; r < J; ++r)
            for (int cc = 0; cc < L; ++cc) {
                auto y_of = [&](int x) {
                    int v = h.edge_var[h.chk_off[r * P + x] + cc];
                    return v / P == cc ? v % P : -1;
                };
                int y0 = y_of(0), y1 = y_of(1);
                if (y0 < 0 || y1 < 0) return false;
                long long a = ((y1 - y0) % P + P) % P;
                if (gcdll(a, P) != 1) return false;
                for (int x = 0; x < P; ++x)
                    if (y_of(x) != static_cast<int>((a * x + y0) % P)) return false;
                qc.a[s][r * L + cc] = static_cast<int>(a);
                qc.b[s][r * L + cc] = y0;
                qc.ainv[s][r * L + cc] = static_cast<int>(inv_mod(a, P));
            }
    }
    qc.P = P;
    qc.J = J;
    qc.L = L;
    return true;
}

// Reduced row-echelon basis of the rows of one side over GF(2).
void build_rref(const HostSide &h, int n, std::vector<uint32_t> &rows32, std::vector<int> &pivots,
                int &rank) {
    const int W = (n + 63) / 64;
    std::vector<std::vector<uint64_t>> rows(h.m, std::vector<uint64_t>(W, 0));
    for (int c = 0; c < h.m; ++c)
        for (int e = h.chk_off[c]; e < h.chk_off[c + 1]; ++e) {
            int v = h.edge_var[e];
            rows[c][v >> 6] ^= 1ull << (v & 63);
        }
    rank = 0;
    pivots.clear();
    for (int col = 0; col < n && rank < h.m; ++col) {
        const int w = col >> 6;
        const uint64_t bit = 1ull << (col & 63);
        int sel = -1;
        for (int r = rank; r < h.m; ++r)
            if (rows[r][w] & bit) { sel = r; break; }
        if (sel < 0) continue;
        std::swap(rows[rank], rows[sel]);
        const std::vector<uint64_t> &pr = rows[rank];
        for (int r = 0; r < h.m; ++r) {
            if (r == rank || !(rows[r][w] & bit)) continue;
            uint64_t *dst = rows[r].data();
            for (int k = w; k < W; ++k) dst[k] ^= pr[k];
            // words below w are zero in pr: columns < col are pivots or eliminated
        }
        pivots.push_back(col);
        ++rank;
    }
    const int W32 = (n + 31) / 32;
    rows32.assign(static_cast<size_t>(rank) * W32, 0u);
    for (int r = 0; r < rank; ++r)
        for (int k = 0; k < W32; ++k)
            rows32[static_cast<size_t>(r) * W32 + k] =
                static_cast<uint32_t>(rows[r][k >> 1] >> (32 * (k & 1)));
}

}  // namespace

namespace jbp {

void free_chunk(jbp_ctx *ctx) {
    for (void *p : ctx->chunk_allocs) (void)hipFree(p);
    ctx->chunk_allocs.clear();
    if (ctx->h_iters) (void)hipHostFree(ctx->h_iters);
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    ctx->h_iters = nullptr;
    ctx->h_flags = nullptr;
    ctx->chunk_cap = 0;
    ctx->tp = TrialPlanes();
    ctx->d_keys = nullptr;
    ctx->d_order = nullptr;
    ctx->d_fb = nullptr;
}

int ensure_chunk(jbp_ctx *ctx, int count, char *err, size_t errlen) {
    if (count <= ctx->chunk_cap) return JBP_OK;
    free_chunk(ctx);
    const size_t c = static_cast<size_t>(count);
    auto alloc = [&](size_t bytes, void **p) -> int {
        HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 4)));
        ctx->chunk_allocs.push_back(*p);
        return JBP_OK;
    };
    const DevGraph &g = ctx->dg;
    int rc = JBP_OK;
    void *p = nullptr;
    if ((rc = alloc(c * g.wn * 4, &p))) return rc; ctx->tp.err_x = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * g.wn * 4, &p))) return rc; ctx->tp.err_z = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * g.wn * 4, &p))) return rc; ctx->tp.est_x = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * g.wn * 4, &p))) return rc; ctx->tp.est_z = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * g.wm[0] * 4, &p))) return rc; ctx->tp.syn_x = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * g.wm[1] * 4, &p))) return rc; ctx->tp.syn_z = static_cast<uint32_t *>(p);
    if ((rc = alloc(c * 4, &p))) return rc; ctx->tp.iters = static_cast<int *>(p);
    // the three outcome flags in one block, so one copy brings them back
    if ((rc = alloc(3 * c, &p))) return rc;
    ctx->tp.syn_ok = static_cast<uint8_t *>(p);
    ctx->tp.success = ctx->tp.syn_ok + c;
    ctx->tp.exact = ctx->tp.syn_ok + 2 * c;
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_iters), c * 4));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_flags), 3 * c));
    if ((rc = alloc(c * 4, &p))) return rc; ctx->d_keys = static_cast<int *>(p);
    if ((rc = alloc(c * 4, &p))) return rc; ctx->d_order = static_cast<int *>(p);
    if ((rc = alloc((c + 4) * 4, &p))) return rc; ctx->d_fb = static_cast<int *>(p);
    ctx->chunk_cap = count;
    return JBP_OK;
}

int ensure_mt(jbp_ctx *ctx, int count, char *err, size_t errlen) {
    // generator states for up to one largest chunk at a time (longer
    // sampling calls run in pieces)
    const int want = std::max(1, std::min(count, kDefaultChunk));
    if (want <= ctx->mt_cap) return JBP_OK;
    if (ctx->d_mt) (void)hipFree(ctx->d_mt);
    ctx->d_mt = nullptr;
    ctx->mt_cap = 0;
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&ctx->d_mt),
                     static_cast<size_t>(want) * kMtWords * sizeof(uint64_t)));
    ctx->mt_cap = want;
    return JBP_OK;
}

int ensure_states(jbp_ctx *ctx, size_t elems, char *err, size_t errlen) {
    if (elems <= ctx->d_states_cap) return JBP_OK;
    if (ctx->d_states) (void)hipFree(ctx->d_states);
    ctx->d_states = nullptr;
    ctx->d_states_cap = 0;
    HIPCHK(hipMalloc(&ctx->d_states, elems * sizeof(int)));
    ctx->d_states_cap = elems;
    return JBP_OK;
}

int ensure_ints(jbp_ctx *ctx, size_t elems, char *err, size_t errlen) {
    if (elems <= ctx->d_ints_cap) return JBP_OK;
    if (ctx->d_ints) (void)hipFree(ctx->d_ints);
    ctx->d_ints = nullptr;
    ctx->d_ints_cap = 0;
    HIPCHK(hipMalloc(&ctx->d_ints, elems * sizeof(int)));
    ctx->d_ints_cap = elems;
    return JBP_OK;
}

int ensure_f64(jbp_ctx *ctx, char *err, size_t errlen) {
    if (ctx->f64_scratch) return JBP_OK;
    // resident 256-thread blocks per compute unit: 3 for degree 12 / 3
    const int per_cu = ctx->dg.max_cdeg <= 12 && ctx->dg.max_vdeg <= 3 ? 3 : 2;
    int slots = ctx->f64_slots > 0 ? ctx->f64_slots : per_cu * std::max(1, ctx->cus);
    size_t bytes = decode_f64_scratch_bytes(ctx->dg, slots);
    HIPCHK(hipMalloc(&ctx->f64_scratch, bytes));
    ctx->f64_slots = slots;
    return JBP_OK;
}

Prior make_prior(const double p4[4]) {
    Prior pr;
    for (int i = 0; i < 4; ++i) pr.p[i] = p4[i];
    return pr;
}

// normalize_msg of {1-p, p/3, p/3, p/3} (jointbp_ets.cpp:5484-5485, 650-657)
Prior prior_from_p(double p) {
    double m[4] = {1.0 - p, p / 3.0, p / 3.0, p / 3.0};
    double s = m[0] + m[1] + m[2] + m[3];
    Prior pr;
    for (int i = 0; i < 4; ++i) pr.p[i] = s <= 0.0 ? 0.25 : m[i] / s;
    return pr;
}

void free_noise(jbp_ctx *ctx) {
    if (ctx->d_th) (void)hipFree(ctx->d_th);
    if (ctx->d_prior) (void)hipFree(ctx->d_prior);
    if (ctx->d_prior32) (void)hipFree(ctx->d_prior32);
    ctx->d_th = nullptr;
    ctx->d_prior = nullptr;
    ctx->d_prior32 = nullptr;
    ctx->noise_cap = 0;
    ctx->noise = NoiseTab();
    ctx->noise_key.clear();
}

int set_noise(jbp_ctx *ctx, const Prior *priors, const double *p, int n, char *err, size_t errlen) {
    if (n <= 0) {
        set_err(err, errlen, "empty channel table");
        return JBP_ERR_ARG;
    }
    std::vector<double> key;
    key.reserve(5 * static_cast<size_t>(n) + 1);
    key.push_back(p ? 1.0 : 0.0);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) key.push_back(priors[i].p[k]);
        key.push_back(p ? p[i] : 0.0);
    }
    if (key == ctx->noise_key) {
        ctx->noise.phase = 0;
        return JBP_OK;
    }
    if (n > ctx->noise_cap) {
        free_noise(ctx);
        HIPCHK(hipMalloc(&ctx->d_th, n * sizeof(DrawThresholds)));
        HIPCHK(hipMalloc(&ctx->d_prior, n * 4 * sizeof(double)));
        HIPCHK(hipMalloc(&ctx->d_prior32, n * kPrior32 * sizeof(float)));
        ctx->noise_cap = n;
    }
    std::vector<DrawThresholds> th(n);
    std::vector<double> pr(4 * static_cast<size_t>(n));
    std::vector<float> p32(kPrior32 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        if (p) th[i] = draw_thresholds(p[i]);
        else std::memset(&th[i], 0, sizeof(DrawThresholds));
        for (int k = 0; k < 4; ++k) pr[4 * i + k] = priors[i].p[k];
        prior32_from(priors[i], &p32[kPrior32 * i]);
    }
    // The previous call's kernels have finished (every entry point ends with
    // a synchronization), so the tables can be overwritten.
    HIPCHK(hipMemcpy(ctx->d_th, th.data(), n * sizeof(DrawThresholds), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_prior, pr.data(), pr.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_prior32, p32.data(), p32.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->noise.th = ctx->d_th;
    ctx->noise.prior = ctx->d_prior;
    ctx->noise.prior32 = ctx->d_prior32;
    ctx->noise.n = n;
    ctx->noise.phase = 0;
    ctx->noise_key.swap(key);
    return JBP_OK;
}

void free_logicals(jbp_ctx *ctx) {
    for (int k = 0; k < 2; ++k) {
        for (void *&p : ctx->lg_mem[k]) {
            if (p) (void)hipFree(p);
            p = nullptr;
        }
        ctx->lg[k] = DevLogicals();
    }
    ctx->have_logicals = false;
}

void free_gen(jbp_ctx *ctx) {
    if (ctx->gen_mem) (void)hipFree(ctx->gen_mem);
    ctx->gen_mem = nullptr;
    ctx->gen_cap = 0;
    ctx->gen = GenState();
    if (ctx->h_ctrl) (void)hipHostFree(ctx->h_ctrl);
    ctx->h_ctrl = nullptr;
    if (ctx->h_ctrl_mix) (void)hipHostFree(ctx->h_ctrl_mix);
    ctx->h_ctrl_mix = nullptr;
    for (auto &e : ctx->gen_ev)
        if (e) (void)hipEventDestroy(e);
    ctx->gen_ev[0] = ctx->gen_ev[1] = nullptr;
}

namespace {

// Resident trials of the generic decoder: enough wavefronts to fill the
// device several times over in the check pass, capped by memory.
int default_gen_slots(const jbp_ctx *ctx) {
    if (ctx->gen_slots > 0) return ctx->gen_slots;
    const DevGraph &g = ctx->dg;
    const int64_t waves_per_group = (g.side[0].m + g.side[1].m + 3) / 4;
    int64_t groups = (64LL * std::max(1, ctx->cus) * 8 + waves_per_group - 1) / waves_per_group;
    groups = std::max<int64_t>(groups, 8);
    while (groups > 1 && gen_carve_bytes(g, static_cast<int>(groups)) > (size_t(8) << 30)) groups /= 2;
    return static_cast<int>(std::min<int64_t>(groups, 256) * 64);
}

int run_gen(jbp_ctx *ctx, int count, const NoiseTab &nt, int max_iter, int warmup, int interval,
            int slots, bool use_graph, double damping, char *err, size_t errlen) {
    const DevGraph &g = ctx->dg;
    if (!gen_supported(g)) {
        set_err(err, errlen, "generic FP32 decoder supports check degree <= 32");
        return JBP_ERR_UNSUPPORTED;
    }
    int S = slots > 0 ? slots : default_gen_slots(ctx);
    S = std::min(S, ((count + 63) / 64) * 64);
    const int groups = std::max(1, (S + 63) / 64);
    const bool damped = damping > 0.0;
    const size_t bytes = gen_carve_bytes(g, groups, damped);
    if (bytes > ctx->gen_cap) {
        if (ctx->gen_mem) (void)hipFree(ctx->gen_mem);
        ctx->gen_mem = nullptr;
        ctx->gen_cap = 0;
        HIPCHK(hipMalloc(&ctx->gen_mem, bytes));
        ctx->gen_cap = bytes;
    }
    if (!ctx->h_ctrl) HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_ctrl), 2 * sizeof(int)));
    for (auto &e : ctx->gen_ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    gen_carve(g, groups, ctx->gen_mem, ctx->gen, damped);
    ctx->gen.damping = static_cast<float>(damped ? damping : 0.0);
    int64_t steps = 0;
    const int rc = gen_decode(g, ctx->gen, ctx->tp, count, nt, max_iter, warmup, interval,
                              ctx->h_ctrl, ctx->gen_ev, ctx->stream, &steps, use_graph);
    ctx->gen_steps = steps;
    if (rc == -1) {
        set_err(err, errlen, "generic decoder exceeded its step bound");
        return JBP_ERR_DEVICE;
    }
    if (rc != 0) {
        set_err(err, errlen, "generic decoder: %s", hipGetErrorString(static_cast<hipError_t>(rc)));
        return JBP_ERR_DEVICE;
    }
    return JBP_OK;
}

}  // namespace

// The learned min-sum decoder runs as its own launch (timed apart).
static bool learned_ms(const LearnedInfer *li) {
    return li && li->T > 0 && (li->kind == kLearnedMS || li->kind == kLearnedLMS);
}

// Decode `count` trials whose syndrome planes are already in ctx->tp.
int run_decode(jbp_ctx *ctx, int count, const NoiseTab &nt, int max_iter, int warmup, int interval,
               double damping, int precision, char *err, size_t errlen,
               const LearnedInfer *learned, int engine, int slots, bool use_graph,
               bool defer_timing,
               bool freeze) {
    if (precision != JBP_FP32 && precision != JBP_FP64) {
        set_err(err, errlen, "precision must be 32 or 64");
        return JBP_ERR_ARG;
    }
    if (engine < JBP_ENGINE_AUTO || engine > JBP_ENGINE_HBM) {
        set_err(err, errlen, "unknown engine %d", engine);
        return JBP_ERR_ARG;
    }
    launch_fill_i32(ctx->d_counter, 0, 1, ctx->stream);
    HIPCHK(hipGetLastError());
    hipEvent_t e0 = defer_timing ? ctx->evt[2] : ctx->ev0, e1 = defer_timing ? ctx->evt[3] : ctx->ev1;
    HIPCHK(hipEventRecord(e0, ctx->stream));
    if (freeze && precision != JBP_FP64) {
        set_err(err, errlen, "freeze-syn runs on the FP64 decoder (precision 64)");
        return JBP_ERR_UNSUPPORTED;
    }
    if (precision == JBP_FP32) {
        // damping mixes each variable-to-check message with the previous
        // one, which the HBM engine keeps and the LDS engine has no room for
        if (damping > 0.0 && (engine == JBP_ENGINE_LDS || learned)) {
            set_err(err, errlen, "FP32 damping runs on the HBM-resident decoder (engine AUTO or HBM)");
            return JBP_ERR_UNSUPPORTED;
        }
        const bool lds = engine == JBP_ENGINE_LDS ||
                         (engine == JBP_ENGINE_AUTO && damping <= 0.0 && (ctx->qc || learned));
        if (lds && !ctx->qc) {
            set_err(err, errlen, "the LDS-resident FP32 decoder needs a 768 x (3 x 12) permutation-block code");
            return JBP_ERR_UNSUPPORTED;
        }
        if (!lds && learned) {
            set_err(err, errlen, "learned decoding runs on the LDS-resident decoder only");
            return JBP_ERR_UNSUPPORTED;
        }
        if (lds) {
            launch_decode_qc_f32(ctx->dg, ctx->qcp, ctx->d_qc_tables, ctx->tp, count, nt, max_iter,
                                 warmup, interval, std::max(1, ctx->cus), ctx->d_counter, ctx->stream,
                                 learned, ctx->d_fb, learned_ms(learned) ? ctx->evt[4] : nullptr);
        } else {
            int rc = run_gen(ctx, count, nt, max_iter, warmup, interval, slots, use_graph, damping, err,
                             errlen);
            if (rc) return rc;
        }
    } else {
        int rc = ensure_f64(ctx, err, errlen);
        if (rc) return rc;
        launch_decode_f64(ctx->dg, ctx->tp, count, nt, max_iter, damping, ctx->f64_slots,
                          ctx->f64_block, ctx->f64_scratch, ctx->d_counter, ctx->stream, freeze);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, ctx->stream));
    if (defer_timing) return JBP_OK;
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms[1] = ms;
    ctx->launches[1]++;
    if (learned_ms(learned) && precision == JBP_FP32) {
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->evt[4]));
        ctx->last_ms[4] = ms;
        ctx->launches[4]++;
    }
    return JBP_OK;
}

void free_pipe(jbp_ctx *ctx) {
    for (hipStream_t &st : ctx->pipe_st)
        if (st) (void)hipStreamSynchronize(st);
    for (PipeBuf &b : ctx->pipe) {
        for (void *p : b.allocs) (void)hipFree(p);
        b.allocs.clear();
        if (b.h_iters) (void)hipHostFree(b.h_iters);
        if (b.h_flags) (void)hipHostFree(b.h_flags);
        for (auto &e : b.ev)
            if (e) (void)hipEventDestroy(e);
        b = PipeBuf();
    }
    for (hipStream_t &st : ctx->pipe_st) {
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
    }
}

namespace {

int ensure_pipe(jbp_ctx *ctx, int cap, char *err, size_t errlen) {
    if (ctx->pipe[0].cap >= cap) return JBP_OK;
    free_pipe(ctx);
    const DevGraph &g = ctx->dg;
    const size_t c = static_cast<size_t>(cap);
    for (PipeBuf &b : ctx->pipe) {
        auto alloc = [&](size_t bytes, void **p) -> int {
            HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 4)));
            b.allocs.push_back(*p);
            return JBP_OK;
        };
        int rc = JBP_OK;
        void *p = nullptr;
        if ((rc = alloc(c * g.wn * 4, &p))) return rc; b.tp.err_x = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * g.wn * 4, &p))) return rc; b.tp.err_z = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * g.wn * 4, &p))) return rc; b.tp.est_x = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * g.wn * 4, &p))) return rc; b.tp.est_z = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * g.wm[0] * 4, &p))) return rc; b.tp.syn_x = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * g.wm[1] * 4, &p))) return rc; b.tp.syn_z = static_cast<uint32_t *>(p);
        if ((rc = alloc(c * 4, &p))) return rc; b.tp.iters = static_cast<int *>(p);
        if ((rc = alloc(3 * c, &p))) return rc;
        b.tp.syn_ok = static_cast<uint8_t *>(p);
        b.tp.success = b.tp.syn_ok + c;
        b.tp.exact = b.tp.syn_ok + 2 * c;
        if ((rc = alloc(4, &p))) return rc; b.counter = static_cast<int *>(p);
        if ((rc = alloc(c * 4, &p))) return rc; b.keys = static_cast<int *>(p);
        if ((rc = alloc(c * 4, &p))) return rc; b.order = static_cast<int *>(p);
        if ((rc = alloc((c + 4) * 4, &p))) return rc; b.fb = static_cast<int *>(p);
        for (auto &e : b.ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&b.h_iters), c * 4));
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&b.h_flags), 3 * c));
        b.cap = cap;
    }
    for (hipStream_t &st : ctx->pipe_st) HIPCHK(hipStreamCreate(&st));
    return JBP_OK;
}

// Runs of the LDS-resident FP32 decoder longer than one launch, as chunks of
// C trials over kPipeDepth buffer sets and three streams (three hardware
// queues of the four a process gets by default, the fourth left to the
// runtime's own copies):
//   sampling   (ctx->stream): seed, sample, weigh and order chunk j; these
//              one-wavefront workgroups run beside the decoder workgroup
//              of the chunk before
//   decode     (pipe_st[0]): after chunk j is sampled, its decoder launch
//   finishing  (pipe_st[1]): after chunk j's decoder launch, the second
//              launch of a two-phase decode (the min-sum decoder's fallback,
//              or plain BP's stragglers), the residual test and the outcome
//              copies
// (Plain BP over fewer than kOffloadChunks chunks instead decodes each chunk
// in one launch, the decode and finishing streams taking turns.)
// The host reduces outcomes in chunk order (min_fails cuts at the same index
// as the serial loop) and queues chunk j+kPipeDepth once chunk j's outcomes
// are read, so a chunk whose slowest trial runs long (the fallback of a
// trial that reaches max-iter holds one compute unit for tens of ms) delays
// nothing but its own outcome: the decode stream keeps up to kPipeDepth - 1
// later chunks going meanwhile.
int simulate_pipelined(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_run_stats &st, int np,
                       const LearnedInfer *learned, jbp_trial_out *pt, char *err, size_t errlen) {
    const DevGraph &g = ctx->dg;
    const int64_t total = cfg->n_trials;
    const int C = cfg->chunk > 0 ? cfg->chunk : kPipeChunk;
    int rc = ensure_pipe(ctx, C, err, errlen);
    if (rc) return rc;
    if ((rc = ensure_mt(ctx, C, err, errlen))) return rc;
    const int64_t nch = (total + C - 1) / C;
    const bool judged = ctx->have_logicals || ctx->have_basis;
    const int interval = std::max(1, cfg->check_interval);
    const int warmup = std::max(0, cfg->check_warmup);
    const bool ms_kind = learned_ms(learned);
    hipStream_t sst = ctx->stream, dst = ctx->pipe_st[0], fst = ctx->pipe_st[1];
    // plain BP: the decoder launch stops at the first scheduled syndrome test
    // at or after max(kStragglerIter, 3 x the mean iterations of the chunks
    // reduced so far), and the finishing stream decodes the trials still
    // unconverged there again from the prior up to max_iter (so a trial that
    // runs to max_iter holds up only its own chunk's outcome, not the decode
    // stream, and few trials are decoded twice at any p)
    auto cap_for = [&](double mean_iters) {
        int c = std::max({kStragglerIter, warmup + interval, static_cast<int>(std::ceil(3.0 * mean_iters))});
        if (c > warmup) c = warmup + interval * ((c - warmup + interval - 1) / interval);
        return c >= cfg->max_iter ? 0 : c;
    };
    // A run of few chunks has little for the decode stream to go on with
    // while the finishing stream decodes stragglers again; there plain BP
    // decodes each chunk in one launch, on the two streams in turn (chunk
    // j+1's workgroups take over compute units as chunk j's last trials
    // finish).
    const bool offload = !learned && nch >= kOffloadChunks && cap_for(0.0) > 0;
    const bool alternate = !learned && !offload;
    int cap = offload ? cap_for(0.0) : 0;
    // the previous call's work on the context stream (e.g. table uploads) first
    HIPCHK(hipStreamSynchronize(sst));
    auto enqueue = [&](int64_t j) -> int {
        PipeBuf &b = ctx->pipe[j % kPipeDepth];
        const int cnt = static_cast<int>(std::min<int64_t>(C, total - j * C));
        const int64_t first = cfg->first_trial + j * C;
        NoiseTab nt = ctx->noise;
        nt.phase = static_cast<int>(((first % np) + np) % np);
        HIPCHK(hipEventRecord(b.ev[0], sst));
        launch_sample(g, b.tp, cfg->base_seed, first, cnt, nt, sst, ctx->d_mt, ctx->mt_cap);
        TrialPlanes tpo = b.tp;
        if (cfg->schedule == 0) {
            launch_order_by_weight(g, b.tp, cnt, b.keys, b.order, sst);
            tpo.order = b.order;
        }
        launch_fill_i32(b.counter, 0, 1, sst);
        HIPCHK(hipEventRecord(b.ev[5], sst));
        // decode and finishing streams of this chunk: with two streams taking
        // turns, each chunk finishes on its own decode stream
        hipStream_t ds = dst, fs = fst;
        if (alternate) ds = fs = (j & 1) ? fst : dst;
        HIPCHK(hipStreamWaitEvent(ds, b.ev[5], 0));
        HIPCHK(hipEventRecord(b.ev[1], ds));
        launch_decode_qc_f32(g, ctx->qcp, ctx->d_qc_tables, tpo, cnt, nt, cfg->max_iter, warmup,
                             interval, std::max(1, ctx->cus), b.counter, ds, learned, b.fb,
                             b.ev[4], fs, learned ? 0 : cap);
        if (!ms_kind) {
            HIPCHK(hipEventRecord(b.ev[4], ds));
            HIPCHK(hipStreamWaitEvent(fs, b.ev[4], 0));
        }
        HIPCHK(hipEventRecord(b.ev[2], fs));
        if (ctx->have_logicals) launch_logical(g, ctx->lg[0], ctx->lg[1], b.tp, cnt, fs);
        else if (ctx->have_basis) launch_residual(g, ctx->basis[0], ctx->basis[1], b.tp, cnt, fs);
        else launch_logical(g, DevLogicals(), DevLogicals(), b.tp, cnt, fs);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(b.h_iters, b.tp.iters, cnt * 4, hipMemcpyDeviceToHost, fs));
        // the three flag arrays sit b.cap apart (the buffers may be larger than C)
        HIPCHK(hipMemcpyAsync(b.h_flags, b.tp.syn_ok, 3 * static_cast<size_t>(b.cap), hipMemcpyDeviceToHost, fs));
        HIPCHK(hipEventRecord(b.ev[3], fs));
        return JBP_OK;
    };
    double dev_ms = 0.0;
    bool stop = false;
    int64_t done = 0;
    for (int64_t j = 0; j < std::min<int64_t>(kPipeDepth, nch); ++j)
        if ((rc = enqueue(j))) return rc;
    for (int64_t j = 0; j < nch && !stop; ++j) {
        PipeBuf &b = ctx->pipe[j % kPipeDepth];
        const int cnt = static_cast<int>(std::min<int64_t>(C, total - j * C));
        HIPCHK(hipEventSynchronize(b.ev[3]));
        float sms = 0.f, dms = 0.f;
        HIPCHK(hipEventElapsedTime(&sms, b.ev[0], b.ev[5]));
        HIPCHK(hipEventElapsedTime(&dms, b.ev[1], b.ev[2]));
        ctx->last_ms[0] = sms;
        ctx->last_ms[1] = dms;
        ctx->launches[0]++;
        ctx->launches[1]++;
        if (judged) {
            float rms = 0.f;
            HIPCHK(hipEventElapsedTime(&rms, b.ev[2], b.ev[3]));
            ctx->last_ms[2] = rms;
        }
        ctx->launches[2]++;
        if (ms_kind) {
            float lms = 0.f;
            HIPCHK(hipEventElapsedTime(&lms, b.ev[1], b.ev[4]));
            ctx->last_ms[4] = lms;
            ctx->launches[4]++;
        }
        dev_ms += dms;
        const uint8_t *ok = b.h_flags, *succ = b.h_flags + b.cap, *exact = b.h_flags + 2 * b.cap;
        int take = cnt;
        for (int i = 0; i < cnt; ++i) {
            const uint8_t s = judged ? succ[i] : static_cast<uint8_t>(ok[i] && exact[i]);
            st.total_iters += b.h_iters[i];
            st.failures += s ? 0 : 1;
            st.bp_failures += ok[i] ? 0 : 1;
            st.exact += exact[i] ? 1 : 0;
            st.stab_success += (s && !exact[i] && judged) ? 1 : 0;
            if (cfg->min_fails > 0 && st.failures >= cfg->min_fails) {
                take = i + 1;
                stop = true;
                break;
            }
        }
        if (pt) {
            for (int i = 0; i < take; ++i) {
                const uint8_t s = judged ? succ[i] : static_cast<uint8_t>(ok[i] && exact[i]);
                if (pt->pp_flags) pt->pp_flags[done + i] = ok[i] ? JBP_PPF_BP_OK : 0;
                if (pt->iterations) pt->iterations[done + i] = b.h_iters[i];
                if (pt->syn_match) pt->syn_match[done + i] = ok[i];
                if (pt->success) pt->success[done + i] = s;
                if (pt->exact) pt->exact[done + i] = exact[i];
            }
        }
        done += take;
        if (offload && done > 0) cap = cap_for(static_cast<double>(st.total_iters) / static_cast<double>(done));
        if (!stop && j + kPipeDepth < nch && (rc = enqueue(j + kPipeDepth))) return rc;
    }
    // chunks still in flight after an early stop finish before the buffers
    // are reused
    HIPCHK(hipStreamSynchronize(sst));
    for (hipStream_t s : ctx->pipe_st) HIPCHK(hipStreamSynchronize(s));
    st.done = done;
    st.device_ms = dev_ms;
    return JBP_OK;
}

}  // namespace

int simulate_impl(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_trial_out *pt, jbp_run_stats *stats,
                  const LearnedInfer *learned, char *err, size_t errlen) {
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !cfg || cfg->n_trials < 0 || cfg->max_iter < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (fault("simulate")) {
        set_err(err, errlen, "fault injected (JBP_FAULT_INJECT=simulate)");
        return JBP_ERR_DEVICE;
    }
    HIPCHK(hipSetDevice(ctx->device));
    jbp_run_stats st;
    std::memset(&st, 0, sizeof(st));
    const DevGraph &g = ctx->dg;
    if (cfg->n_p < 0 || (cfg->n_p > 0 && !cfg->p_list) || cfg->n_p > kMaxPList) {
        set_err(err, errlen, "p_list needs 1..%d entries", kMaxPList);
        return JBP_ERR_ARG;
    }
    const int np = cfg->n_p > 0 ? cfg->n_p : 1;
    const double *plist = cfg->n_p > 0 ? cfg->p_list : &cfg->p;
    for (int i = 0; i < np; ++i)
        if (!(plist[i] >= 0.0 && plist[i] <= 1.0)) {
            set_err(err, errlen, "p = %g is outside [0, 1]", plist[i]);
            return JBP_ERR_ARG;
        }
    {
        std::vector<Prior> pri(np);
        for (int i = 0; i < np; ++i) pri[i] = prior_from_p(plist[i]);
        int rc0 = set_noise(ctx, pri.data(), plist, np, err, errlen);
        if (rc0) return rc0;
    }
    if (cfg->chunk < 0 || cfg->chunk > kDefaultChunk) {
        set_err(err, errlen, "chunk must be 0..%d (got %d)", kDefaultChunk, cfg->chunk);
        return JBP_ERR_ARG;
    }
    if (cfg->block > 0) {
        if (cfg->block > 512 || cfg->block % 64) {
            set_err(err, errlen, "block must be a multiple of 64 up to 512 (got %d)", cfg->block);
            return JBP_ERR_ARG;
        }
        ctx->f64_block = cfg->block;
    }
    if (cfg->slots > 0 && cfg->precision == JBP_FP64 && cfg->slots != ctx->f64_slots) {
        if (ctx->f64_scratch) (void)hipFree(ctx->f64_scratch);
        ctx->f64_scratch = nullptr;
        ctx->f64_slots = cfg->slots;
    }
    const int64_t total = cfg->n_trials;
    const bool pp_on = cfg->pp != 0;
    if (pp_on && learned) {
        set_err(err, errlen, "post-processing is not combined with the learned decoder");
        return JBP_ERR_UNSUPPORTED;
    }
    if (cfg->freeze_syn && (pp_on || learned || cfg->precision != JBP_FP64)) {
        set_err(err, errlen, "freeze-syn runs on the FP64 decoder without post-processing");
        return JBP_ERR_UNSUPPORTED;
    }
    if (pp_on && cfg->flip_hist < 0) {
        set_err(err, errlen, "flip_hist must be 0 or more");
        return JBP_ERR_ARG;
    }
    // decode iterations before the first post-processing trigger can fire
    const int pp_h = std::min(99, cfg->max_iter - 1);
    const bool lds = cfg->precision == JBP_FP32 && cfg->damping <= 0.0 && ctx->qc &&
                     (cfg->engine == JBP_ENGINE_LDS ||
                      (cfg->engine == JBP_ENGINE_AUTO) || learned);
    if (lds && total > (cfg->chunk > 0 ? cfg->chunk : kPipeChunk) && !pp_on && !(pt && (pt->err || pt->est))) {
        int rc0 = simulate_pipelined(ctx, cfg, st, np, learned, pt, err, errlen);
        if (rc0) return rc0;
        st.elapsed_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return JBP_OK;
    }
    const int chunk = static_cast<int>(std::min<int64_t>(std::max<int64_t>(total, 1), kDefaultChunk));
    int rc = ensure_chunk(ctx, chunk, err, errlen);
    if (rc) return rc;
    if ((rc = ensure_mt(ctx, chunk, err, errlen))) return rc;
    // outcome copies land in pinned memory: iterations, and the three flag
    // arrays (one block on the device) in one copy
    const int cap = ctx->chunk_cap;
    int *iters = ctx->h_iters;
    uint8_t *ok = ctx->h_flags, *succ = ctx->h_flags + cap, *exact = ctx->h_flags + 2 * cap;
    double dev_ms = 0.0;
    bool stop = false;
    for (int64_t done = 0; done < total && !stop;) {
        const int cnt = static_cast<int>(std::min<int64_t>(chunk, total - done));
        const int64_t first = cfg->first_trial + done;
        // trial index idx uses channel entry idx mod n_p
        NoiseTab nt = ctx->noise;
        nt.phase = static_cast<int>(((first % np) + np) % np);
        // (no host wait between the launches of a chunk: the sample and
        // decode timings are read after the chunk's one synchronisation)
        HIPCHK(hipEventRecord(ctx->evt[0], ctx->stream));
        launch_sample(g, ctx->tp, cfg->base_seed, first, cnt, nt, ctx->stream, ctx->d_mt, ctx->mt_cap);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->evt[1], ctx->stream));
        if (cfg->schedule == 0) {
            launch_order_by_weight(g, ctx->tp, cnt, ctx->d_keys, ctx->d_order, ctx->stream);
            HIPCHK(hipGetLastError());
            ctx->tp.order = ctx->d_order;
        }
        const bool decoded = !pp_on || pp_h > 0;
        if (decoded) {
            rc = run_decode(ctx, cnt, nt, pp_on ? pp_h : cfg->max_iter, cfg->check_warmup,
                            cfg->check_interval, cfg->damping, cfg->precision, err, errlen, learned,
                            cfg->engine, cfg->precision == JBP_FP32 ? cfg->slots : 0, cfg->graph != 0,
                            true, cfg->freeze_syn != 0);
            ctx->tp.order = nullptr;
            if (rc) return rc;
        }
        ctx->tp.order = nullptr;
        // post-processing: the trials still unconverged after pp_h iterations
        std::vector<int> pp_list;
        std::vector<PPOutcome> ppo;
        std::vector<int> pp_of;
        if (pp_on) {
            if (pp_h > 0) {
                HIPCHK(hipMemcpyAsync(ok, ctx->tp.syn_ok, cnt, hipMemcpyDeviceToHost, ctx->stream));
                HIPCHK(hipStreamSynchronize(ctx->stream));
            }
            for (int i = 0; i < cnt; ++i)
                if (pp_h <= 0 || !ok[i]) pp_list.push_back(i);
            auto p0 = std::chrono::steady_clock::now();
            rc = run_pp(ctx, nt, cfg->max_iter, cfg->damping, cfg->flip_hist, pp_list, ppo, err, errlen);
            if (rc) return rc;
            st.pp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p0).count();
            pp_of.assign(cnt, -1);
            for (size_t k = 0; k < pp_list.size(); ++k) pp_of[pp_list[k]] = static_cast<int>(k);
        }
        if (ctx->have_logicals) {
            HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
            launch_logical(g, ctx->lg[0], ctx->lg[1], ctx->tp, cnt, ctx->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        } else if (ctx->have_basis) {
            HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
            launch_residual(g, ctx->basis[0], ctx->basis[1], ctx->tp, cnt, ctx->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        } else {
            // no test of the residual: the kernel with no operators still
            // sets the exact-match flags, and success becomes exact below
            launch_logical(g, DevLogicals(), DevLogicals(), ctx->tp, cnt, ctx->stream);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpyAsync(iters, ctx->tp.iters, cnt * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(ok, ctx->tp.syn_ok, 3 * static_cast<size_t>(cap), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        {
            float sms = 0.f;
            HIPCHK(hipEventElapsedTime(&sms, ctx->evt[0], ctx->evt[1]));
            ctx->last_ms[0] = sms;
            ctx->launches[0]++;
            if (decoded) {
                float dms = 0.f;
                HIPCHK(hipEventElapsedTime(&dms, ctx->evt[2], ctx->evt[3]));
                ctx->last_ms[1] = dms;
                ctx->launches[1]++;
                if (learned_ms(learned) && cfg->precision == JBP_FP32) {
                    float lms = 0.f;
                    HIPCHK(hipEventElapsedTime(&lms, ctx->evt[2], ctx->evt[4]));
                    ctx->last_ms[4] = lms;
                    ctx->launches[4]++;
                }
                dev_ms += dms;
            }
        }
        const bool judged = ctx->have_logicals || ctx->have_basis;
        if (judged) {
            float rms = 0.f;
            HIPCHK(hipEventElapsedTime(&rms, ctx->ev0, ctx->ev1));
            ctx->last_ms[2] = rms;
            ctx->launches[2]++;
        }
        if (pt && (pt->err || pt->est)) {
            if ((rc = ensure_states(ctx, static_cast<size_t>(cnt) * g.n, err, errlen))) return rc;
            for (int which = 0; which < 2; ++which) {
                int *dst = which == 0 ? pt->err : pt->est;
                if (!dst) continue;
                const uint32_t *xb = which == 0 ? ctx->tp.err_x : ctx->tp.est_x;
                const uint32_t *zb = which == 0 ? ctx->tp.err_z : ctx->tp.est_z;
                launch_unpack_states(xb, zb, g.wn, g.n, cnt, ctx->d_states, ctx->stream);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpy(dst + static_cast<size_t>(done) * g.n, ctx->d_states,
                                 static_cast<size_t>(cnt) * g.n * 4, hipMemcpyDeviceToHost));
            }
        }
        int take = cnt;
        std::vector<uint8_t> flags(cnt, 0);
        for (int i = 0; i < cnt; ++i) {
            if (!judged) succ[i] = (ok[i] && exact[i]) ? 1 : 0;
            const PPOutcome *po = pp_on && pp_of[i] >= 0 ? &ppo[pp_of[i]] : nullptr;
            const bool bp_ok = po ? po->bp_ok : ok[i] != 0;
            const bool success = succ[i] != 0;
            st.total_iters += iters[i];
            st.failures += success ? 0 : 1;
            st.bp_failures += bp_ok ? 0 : 1;
            st.exact += exact[i] ? 1 : 0;
            st.stab_success += (success && !exact[i] && judged) ? 1 : 0;
            uint8_t f = bp_ok ? JBP_PPF_BP_OK : 0;
            if (po) {
                // the reference's counting (jointbp_ets.cpp:6113-6228)
                bool used_any = po->used_any || po->used6 || po->used12;
                bool used6 = po->used6, used12 = po->used12;
                if (po->pp_success && used_any) {
                    st.ets_used++;
                    st.ets6_used += used6 ? 1 : 0;
                    st.ets12_used += used12 ? 1 : 0;
                }
                bool ets_ok = po->pp_ets, flip_ok = po->pp_flip, osd_ok = false;
                if (po->post_run) {
                    used_any = used_any || po->post.ets_applied;
                    used6 = used6 || po->post.used6;
                    used12 = used12 || po->post.used12;
                    st.ets_used += po->post.ets_applied ? 1 : 0;
                    st.ets6_used += used6 ? 1 : 0;
                    st.ets12_used += used12 ? 1 : 0;
                    ets_ok = ets_ok || po->post.ets_ok;
                    flip_ok = flip_ok || po->post.flip_ok;
                    osd_ok = po->post.osd_ok;
                }
                if (success && ets_ok) {
                    st.ets_saved++;
                    st.ets6_saved += used6 ? 1 : 0;
                    st.ets12_saved += used12 ? 1 : 0;
                    f |= JBP_PPF_ETS_SAVED;
                }
                if (success && flip_ok) {
                    st.flip_pp_success++;
                    f |= JBP_PPF_FLIP_SAVED;
                }
                const bool osd_success = osd_ok || (po->pp_success && !po->pp_ets && !po->pp_flip);
                if (success && (po->pp_success || ets_ok || flip_ok || osd_ok)) {
                    st.pp_success++;
                    f |= JBP_PPF_PP_SUCCESS;
                }
                if (success && osd_success) {
                    st.osd_pp_success++;
                    f |= JBP_PPF_OSD_SAVED;
                }
                f |= (used_any ? JBP_PPF_ETS_USED : 0) | (used6 ? JBP_PPF_ETS6 : 0) |
                     (used12 ? JBP_PPF_ETS12 : 0);
            }
            flags[i] = f;
            if (cfg->min_fails > 0 && st.failures >= cfg->min_fails) {
                take = i + 1;
                stop = true;
                break;
            }
        }
        st.pp_trials += static_cast<int64_t>(std::count_if(
            pp_list.begin(), pp_list.end(), [&](int i) { return i < take; }));
        if (pt) {
            for (int i = 0; i < take; ++i) {
                if (pt->iterations) pt->iterations[done + i] = iters[i];
                if (pt->syn_match) pt->syn_match[done + i] = ok[i];
                if (pt->success) pt->success[done + i] = succ[i];
                if (pt->exact) pt->exact[done + i] = exact[i];
                if (pt->pp_flags) pt->pp_flags[done + i] = flags[i];
            }
        }
        done += take;
        st.done = done;
    }
    st.device_ms = dev_ms;
    st.elapsed_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return JBP_OK;
}


}  // namespace jbp

namespace jbp {
namespace {

// One trial from syndrome planes bx / bz through a graph captured on first
// use: H2D of the syndrome words from pinned staging, the counter reset (a
// fill kernel), the decoder launch, and D2H of iterations, match flag and
// estimate planes.  Re-captured when the decode parameters change; the
// prior lives in the context's channel table, which the graph reads.
int decode_graph(jbp_ctx *ctx, const std::vector<uint32_t> &bx, const std::vector<uint32_t> &bz,
                 int max_iter, int warmup, int interval, double damping, int precision,
                 jbp_decode_result *out, int *est_out,
                 std::chrono::steady_clock::time_point t0, char *err, size_t errlen) {
    const DevGraph &g = ctx->dg;
    const size_t in_words = static_cast<size_t>(g.wm[0]) + g.wm[1];
    const size_t out_words = 2 + 2 * static_cast<size_t>(g.wn);
    if (!ctx->h_stage)
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_stage), (in_words + out_words) * 4));
    uint32_t *hin = ctx->h_stage;
    uint32_t *hout = ctx->h_stage + in_words;
    if (precision == JBP_FP64) {
        int rc = ensure_f64(ctx, err, errlen);
        if (rc) return rc;
    }
    uint64_t dbits = 0;
    std::memcpy(&dbits, &damping, sizeof(dbits));
    const TrialPlanes &tp = ctx->tp;
    const std::vector<uintptr_t> key = {
        static_cast<uintptr_t>(precision), static_cast<uintptr_t>(max_iter),
        static_cast<uintptr_t>(warmup), static_cast<uintptr_t>(interval),
        static_cast<uintptr_t>(dbits), reinterpret_cast<uintptr_t>(tp.syn_x),
        reinterpret_cast<uintptr_t>(tp.syn_z), reinterpret_cast<uintptr_t>(tp.iters),
        reinterpret_cast<uintptr_t>(tp.syn_ok), reinterpret_cast<uintptr_t>(tp.est_x),
        reinterpret_cast<uintptr_t>(tp.est_z), reinterpret_cast<uintptr_t>(ctx->noise.prior),
        reinterpret_cast<uintptr_t>(ctx->noise.prior32), reinterpret_cast<uintptr_t>(ctx->f64_scratch),
        static_cast<uintptr_t>(ctx->f64_slots), static_cast<uintptr_t>(ctx->f64_block)};
    const bool same = ctx->g1_exec && key == ctx->g1_key;
    if (!same) {
        if (ctx->g1_exec) (void)hipGraphExecDestroy(ctx->g1_exec);
        ctx->g1_exec = nullptr;
        hipStream_t st = ctx->stream;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        (void)hipMemcpyAsync(ctx->tp.syn_x, hin, g.wm[0] * 4, hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(ctx->tp.syn_z, hin + g.wm[0], g.wm[1] * 4, hipMemcpyHostToDevice, st);
        launch_fill_i32(ctx->d_counter, 0, 1, st);
        if (precision == JBP_FP32)
            launch_decode_qc_f32(g, ctx->qcp, ctx->d_qc_tables, ctx->tp, 1, ctx->noise, max_iter,
                                 warmup, interval, std::max(1, ctx->cus), ctx->d_counter, st, nullptr);
        else
            launch_decode_f64(g, ctx->tp, 1, ctx->noise, max_iter, damping, ctx->f64_slots,
                              ctx->f64_block, ctx->f64_scratch, ctx->d_counter, st);
        (void)hipMemcpyAsync(hout, ctx->tp.iters, 4, hipMemcpyDeviceToHost, st);
        (void)hipMemcpyAsync(hout + 1, ctx->tp.syn_ok, 1, hipMemcpyDeviceToHost, st);
        (void)hipMemcpyAsync(hout + 2, ctx->tp.est_x, g.wn * 4, hipMemcpyDeviceToHost, st);
        (void)hipMemcpyAsync(hout + 2 + g.wn, ctx->tp.est_z, g.wn * 4, hipMemcpyDeviceToHost, st);
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamEndCapture(st, &graph));
        hipError_t e = hipGraphInstantiate(&ctx->g1_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            ctx->g1_exec = nullptr;
            set_err(err, errlen, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            return JBP_ERR_DEVICE;
        }
        ctx->g1_key = key;
    }
    std::copy(bx.begin(), bx.end(), hin);
    std::copy(bz.begin(), bz.end(), hin + g.wm[0]);
    hout[1] = 0;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(hipGraphLaunch(ctx->g1_exec, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms[1] = ms;
    ctx->launches[1]++;
    const uint32_t *ex = hout + 2;
    const uint32_t *ez = hout + 2 + g.wn;
    if (est_out)
        for (int v = 0; v < g.n; ++v)
            est_out[v] = static_cast<int>(((ex[v >> 5] >> (v & 31)) & 1u) |
                                          (((ez[v >> 5] >> (v & 31)) & 1u) << 1));
    out->iterations = static_cast<int>(hout[0]);
    out->syndrome_match = static_cast<int>(hout[1] & 0xffu);
    out->check_count = out->iterations;
    out->kernel_ms = ms;  // the whole graph: copies and decode
    out->memcpy_ms = 0.0;
    out->host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return JBP_OK;
}

}  // namespace
}  // namespace jbp

extern "C" {

int jbp_create(const jbp_graph_desc *g, int device_id, jbp_ctx **out, char *err, size_t errlen) {
    if (!out) {
        set_err(err, errlen, "out is NULL");
        return JBP_ERR_ARG;
    }
    *out = nullptr;
    if (fault("create")) {
        set_err(err, errlen, "fault injected (JBP_FAULT_INJECT=create)");
        return JBP_ERR_DEVICE;
    }
    if (!g || g->nvars <= 0) {
        set_err(err, errlen, "graph has no variables");
        return JBP_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err(err, errlen, "no HIP device available");
        return JBP_ERR_DEVICE;
    }
    if (device_id < 0 || device_id >= ndev) {
        set_err(err, errlen, "invalid device id %d (have %d)", device_id, ndev);
        return JBP_ERR_ARG;
    }
    HIPCHK(hipSetDevice(device_id));
    {
        int rc0 = noalloc_setup(device_id, err, errlen);
        if (rc0) return rc0;
    }
    auto *ctx = new jbp_ctx();
    ctx->device = device_id;
    ctx->n = g->nvars;
    int rc = validate_side("X", g->nvars, g->mX, g->x_edges, g->x_check_offsets, g->x_edge_var,
                           g->x_var_offsets, g->x_var_edges, ctx->hs[0], err, errlen);
    if (!rc)
        rc = validate_side("Z", g->nvars, g->mZ, g->z_edges, g->z_check_offsets, g->z_edge_var,
                           g->z_var_offsets, g->z_var_edges, ctx->hs[1], err, errlen);
    if (rc) {
        delete ctx;
        return rc;
    }
    {
        const int m[2] = {ctx->hs[0].m, ctx->hs[1].m};
        const int *off[2] = {ctx->hs[0].chk_off.data(), ctx->hs[1].chk_off.data()};
        const int *var[2] = {ctx->hs[0].edge_var.data(), ctx->hs[1].edge_var.data()};
        ctx->pp_graph = pp::graph_from_csr(ctx->n, m, off, var);
    }
    auto fail = [&](int code) {
        jbp_destroy(ctx);
        return code;
    };
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_id) != hipSuccess) {
        set_err(err, errlen, "hipGetDeviceProperties failed");
        return fail(JBP_ERR_DEVICE);
    }
    ctx->cus = prop.multiProcessorCount;
    if (hipStreamCreate(&ctx->stream) != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
        hipEventCreate(&ctx->ev1) != hipSuccess || hipEventCreate(&ctx->evt[0]) != hipSuccess ||
        hipEventCreate(&ctx->evt[1]) != hipSuccess || hipEventCreate(&ctx->evt[2]) != hipSuccess ||
        hipEventCreate(&ctx->evt[3]) != hipSuccess || hipEventCreate(&ctx->evt[4]) != hipSuccess) {
        set_err(err, errlen, "stream/event creation failed");
        return fail(JBP_ERR_DEVICE);
    }
    DevGraph &dg = ctx->dg;
    dg.n = ctx->n;
    dg.wn = (ctx->n + 31) / 32;
    for (int s = 0; s < 2; ++s) {
        HostSide &h = ctx->hs[s];
        DevSide &d = dg.side[s];
        d.m = h.m;
        d.E = h.E;
        dg.wm[s] = std::max(1, (h.m + 31) / 32);
        if ((rc = dev_upload(ctx, h.chk_off, &d.chk_off, err, errlen))) return fail(rc);
        if ((rc = dev_upload(ctx, h.edge_var, &d.edge_var, err, errlen))) return fail(rc);
        if ((rc = dev_upload(ctx, h.var_off, &d.var_off, err, errlen))) return fail(rc);
        if ((rc = dev_upload(ctx, h.var_edge, &d.var_edge, err, errlen))) return fail(rc);
        for (int c = 0; c < h.m; ++c)
            dg.max_cdeg = std::max(dg.max_cdeg, h.chk_off[c + 1] - h.chk_off[c]);
        for (int v = 0; v < ctx->n; ++v)
            dg.max_vdeg = std::max(dg.max_vdeg, h.var_off[v + 1] - h.var_off[v]);
    }
    if (dg.max_cdeg > 32 || dg.max_vdeg > 16) {
        set_err(err, errlen, "degrees above 32 (check) / 16 (variable) are not supported");
        return fail(JBP_ERR_UNSUPPORTED);
    }
    dg.min_cdeg = dg.max_cdeg;
    for (int s = 0; s < 2; ++s) {
        // FP64 message rows (position-major)
        const HostSide &h = ctx->hs[s];
        DevSide &d = dg.side[s];
        d.cap = std::max(1, dg.max_cdeg * h.m);
        std::vector<int> row(std::max(h.E, 1), 0), vrow(std::max(h.E, 1), 0);
        for (int c = 0; c < h.m; ++c)
            for (int i = h.chk_off[c]; i < h.chk_off[c + 1]; ++i) row[i] = (i - h.chk_off[c]) * h.m + c;
        for (int k = 0; k < h.E; ++k) vrow[k] = row[h.var_edge[k]];
        if ((rc = dev_upload(ctx, vrow, &d.vrow, err, errlen))) return fail(rc);
    }
    dg.vpad_w = dg.max_vdeg <= 4 ? 4 : dg.max_vdeg <= 8 ? 8 : 16;
    dg.cpad_w = dg.max_cdeg <= 8 ? 8 : dg.max_cdeg <= 12 ? 12 : dg.max_cdeg <= 16 ? 16 : 32;
    for (int s = 0; s < 2; ++s) {
        const HostSide &h = ctx->hs[s];
        const int W = dg.cpad_w;
        std::vector<int> pad(std::max<size_t>(static_cast<size_t>(h.m) * W, 1), ctx->n);  // n: a zero bit
        for (int c = 0; c < h.m; ++c)
            for (int i = h.chk_off[c]; i < h.chk_off[c + 1]; ++i)
                pad[static_cast<size_t>(c) * W + (i - h.chk_off[c])] = h.edge_var[i];
        if ((rc = dev_upload(ctx, pad, &dg.cpad[s], err, errlen))) return fail(rc);
    }
    for (int s = 0; s < 2; ++s) {
        const HostSide &h = ctx->hs[s];
        for (int c = 0; c < h.m; ++c) dg.min_cdeg = std::min(dg.min_cdeg, h.chk_off[c + 1] - h.chk_off[c]);
        const int W = dg.vpad_w;
        std::vector<int> pad(static_cast<size_t>(ctx->n) * W, h.E);  // E: the zero message
        for (int v = 0; v < ctx->n; ++v)
            for (int i = h.var_off[v]; i < h.var_off[v + 1]; ++i)
                pad[static_cast<size_t>(v) * W + (i - h.var_off[v])] = h.var_edge[i];
        if ((rc = dev_upload(ctx, pad, &dg.vpad[s], err, errlen))) return fail(rc);
    }
    const bool is_qc = detect_qc(ctx->hs, ctx->n, ctx->qcp);
    ctx->qc = is_qc && qc_f32_supported(ctx->qcp, device_id, nullptr, 0);
    if (is_qc) {
        const int JL = ctx->qcp.J * ctx->qcp.L;
        std::vector<int> tab(6 * JL);
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < JL; ++i) {
                tab[(3 * s + 0) * JL + i] = ctx->qcp.a[s][i];
                tab[(3 * s + 1) * JL + i] = ctx->qcp.b[s][i];
                tab[(3 * s + 2) * JL + i] = ctx->qcp.ainv[s][i];
            }
        if ((rc = dev_upload(ctx, tab, &dg.qc_tab, err, errlen))) return fail(rc);
        dg.qc_P = ctx->qcp.P;
        dg.qc_J = ctx->qcp.J;
        dg.qc_L = ctx->qcp.L;
    }
    if (ctx->qc) {
        const int JL = ctx->qcp.J * ctx->qcp.L;
        std::vector<int> tab(4 * JL);
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < JL; ++i) {
                tab[s * JL + i] = ctx->qcp.ainv[s][i];
                tab[2 * JL + s * JL + i] = ctx->qcp.b[s][i];
            }
        const int *dt = nullptr;
        if ((rc = dev_upload(ctx, tab, &dt, err, errlen))) return fail(rc);
        ctx->d_qc_tables = const_cast<int *>(dt);
    }
    // Residual-test bases: row spaces of Hx (side 0) and Hz (side 1).
    // Skipped when the dense elimination would be too slow on the host (the
    // caller can then install logical operators with jbp_set_logicals).
    const int mmax = std::max(ctx->hs[0].m, ctx->hs[1].m);
    size_t basis_bytes = static_cast<size_t>(mmax) * dg.wn * 4;
    const double elim_work = static_cast<double>(mmax) * mmax * ((ctx->n + 63) / 64);
    if (basis_bytes <= kMaxBasisBytes && elim_work <= kMaxBasisWork) {
        for (int s = 0; s < 2; ++s) {
            std::vector<uint32_t> rows;
            std::vector<int> piv;
            int rank = 0;
            build_rref(ctx->hs[s], ctx->n, rows, piv, rank);
            DevBasis &b = ctx->basis[s];
            b.rank = rank;
            b.words = dg.wn;
            if ((rc = dev_upload(ctx, rows, &b.rows, err, errlen))) return fail(rc);
            if ((rc = dev_upload(ctx, piv, &b.pivot, err, errlen))) return fail(rc);
        }
        ctx->have_basis = true;
    }
    if (hipMalloc(&ctx->d_counter, sizeof(int)) != hipSuccess) {
        set_err(err, errlen, "hipMalloc counter failed");
        return fail(JBP_ERR_NOMEM);
    }
    *out = ctx;
    return JBP_OK;
}

void jbp_destroy(jbp_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->g1_exec) (void)hipGraphExecDestroy(ctx->g1_exec);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    free_pipe(ctx);
    free_pp(ctx);
    free_chunk(ctx);
    free_noise(ctx);
    free_gen(ctx);
    free_logicals(ctx);
    for (void *p : ctx->allocs) (void)hipFree(p);
    if (ctx->f64_scratch) (void)hipFree(ctx->f64_scratch);
    if (ctx->d_counter) (void)hipFree(ctx->d_counter);
    if (ctx->d_states) (void)hipFree(ctx->d_states);
    if (ctx->d_mt) (void)hipFree(ctx->d_mt);
    if (ctx->d_ints) (void)hipFree(ctx->d_ints);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipEvent_t e : ctx->evt)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int jbp_info(jbp_ctx *ctx, int64_t out[8]) {
    if (!ctx || !out) return JBP_ERR_ARG;
    out[0] = ctx->qc ? 1 : 0;
    out[1] = ctx->qcp.P;
    out[2] = ctx->qcp.J;
    out[3] = ctx->qcp.L;
    out[4] = ctx->device;
    out[5] = ctx->cus;
    out[6] = ctx->f64_block;
    out[7] = ctx->have_logicals ? 2 : ctx->have_basis ? 1 : 0;
    return JBP_OK;
}

double jbp_last_kernel_ms(jbp_ctx *ctx, int which, int64_t *launches) {
    if (!ctx || which < 0 || which > 4) return -1.0;
    if (which == 3) {  // generic decoder: engine steps of its last decode call
        if (launches) *launches = ctx->gen_steps;
        return ctx->last_ms[1];
    }
    if (launches) *launches = ctx->launches[which];
    return ctx->last_ms[which];
}

int jbp_sample(jbp_ctx *ctx, uint64_t base_seed, int64_t first_trial, int count, double p,
               int *err_out, int *sx_out, int *sz_out, char *err, size_t errlen) {
    if (!ctx || count < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (count == 0) return JBP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, count, err, errlen);
    if (rc) return rc;
    if ((rc = ensure_mt(ctx, count, err, errlen))) return rc;
    const DevGraph &g = ctx->dg;
    {
        const Prior pri = prior_from_p(p);
        if ((rc = set_noise(ctx, &pri, &p, 1, err, errlen))) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    launch_sample(g, ctx->tp, base_seed, first_trial, count, ctx->noise, ctx->stream, ctx->d_mt,
                  ctx->mt_cap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    if (err_out) {
        if ((rc = ensure_states(ctx, static_cast<size_t>(count) * g.n, err, errlen))) return rc;
        launch_unpack_states(ctx->tp.err_x, ctx->tp.err_z, g.wn, g.n, count, ctx->d_states,
                             ctx->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(err_out, ctx->d_states, static_cast<size_t>(count) * g.n * 4,
                              hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms[0] = ms;
    ctx->launches[0]++;
    for (int s = 0; s < 2; ++s) {
        int *dst = s == 0 ? sx_out : sz_out;
        if (!dst) continue;
        const int m = ctx->hs[s].m, wm = g.wm[s];
        std::vector<uint32_t> bits(static_cast<size_t>(count) * wm);
        HIPCHK(hipMemcpy(bits.data(), s == 0 ? ctx->tp.syn_x : ctx->tp.syn_z, bits.size() * 4,
                         hipMemcpyDeviceToHost));
        for (int t = 0; t < count; ++t)
            for (int c = 0; c < m; ++c)
                dst[static_cast<size_t>(t) * m + c] =
                    (bits[static_cast<size_t>(t) * wm + (c >> 5)] >> (c & 31)) & 1;
    }
    return JBP_OK;
}

int jbp_decode_batch(jbp_ctx *ctx, int count, const uint32_t *sx_bits, const uint32_t *sz_bits,
                     const double prior[4], int max_iter, double damping, int precision,
                     int *iterations_out, uint8_t *syn_match_out, int *est_out, char *err,
                     size_t errlen) {
    if (!ctx || count < 0 || !sx_bits || !sz_bits || !prior || max_iter < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (fault("decode")) {
        set_err(err, errlen, "fault injected (JBP_FAULT_INJECT=decode)");
        return JBP_ERR_DEVICE;
    }
    if (count == 0) return JBP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, count, err, errlen);
    if (rc) return rc;
    const DevGraph &g = ctx->dg;
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_x, sx_bits, static_cast<size_t>(count) * g.wm[0] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->tp.syn_z, sz_bits, static_cast<size_t>(count) * g.wm[1] * 4,
                          hipMemcpyHostToDevice, ctx->stream));
    {
        const Prior pri = make_prior(prior);
        if ((rc = set_noise(ctx, &pri, nullptr, 1, err, errlen))) return rc;
    }
    rc = run_decode(ctx, count, ctx->noise, max_iter, 0, 1, damping, precision, err, errlen);
    if (rc) return rc;
    if (iterations_out)
        HIPCHK(hipMemcpy(iterations_out, ctx->tp.iters, static_cast<size_t>(count) * 4,
                         hipMemcpyDeviceToHost));
    if (syn_match_out)
        HIPCHK(hipMemcpy(syn_match_out, ctx->tp.syn_ok, static_cast<size_t>(count),
                         hipMemcpyDeviceToHost));
    if (est_out) {
        if ((rc = ensure_states(ctx, static_cast<size_t>(count) * g.n, err, errlen))) return rc;
        launch_unpack_states(ctx->tp.est_x, ctx->tp.est_z, g.wn, g.n, count, ctx->d_states,
                             ctx->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(est_out, ctx->d_states, static_cast<size_t>(count) * g.n * 4,
                              hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return JBP_OK;
}

int jbp_decode(jbp_ctx *ctx, const int *sx, const int *sz, const double prior[4], int max_iter,
               int check_warmup, int check_interval, int measure_costs, int use_graph,
               int freeze_syn, double damping, int precision, jbp_decode_result *out,
               int *est_out, char *err, size_t errlen) {
    (void)measure_costs;
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !sx || !sz || !prior || !out || max_iter < 0) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (freeze_syn && precision != JBP_FP64) {
        set_err(err, errlen, "freeze-syn runs on the FP64 decoder (precision 64)");
        return JBP_ERR_UNSUPPORTED;
    }
    if (fault("decode")) {
        set_err(err, errlen, "fault injected (JBP_FAULT_INJECT=decode)");
        return JBP_ERR_DEVICE;
    }
    std::memset(out, 0, sizeof(*out));
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, 1, err, errlen);
    if (rc) return rc;
    const DevGraph &g = ctx->dg;
    std::vector<uint32_t> bx(g.wm[0], 0u), bz(g.wm[1], 0u);
    for (int c = 0; c < ctx->hs[0].m; ++c)
        if (sx[c]) bx[c >> 5] |= 1u << (c & 31);
    for (int c = 0; c < ctx->hs[1].m; ++c)
        if (sz[c]) bz[c >> 5] |= 1u << (c & 31);
    int warmup = precision == JBP_FP32 ? std::max(0, check_warmup) : 0;
    int interval = precision == JBP_FP32 ? std::max(1, check_interval) : 1;
    {
        const Prior pri = make_prior(prior);
        if ((rc = set_noise(ctx, &pri, nullptr, 1, err, errlen))) return rc;
    }
    // Graph replay applies to the single-launch decoders (the LDS-resident
    // FP32 decoder and the FP64 decoder); the generic FP32 decoder's host
    // loop cannot be captured, so it runs eagerly.
    if (use_graph && !freeze_syn && (precision == JBP_FP64 || (ctx->qc && damping <= 0.0)) &&
        (precision == JBP_FP32 || precision == JBP_FP64)) {
        rc = decode_graph(ctx, bx, bz, max_iter, warmup, interval, damping, precision, out, est_out,
                          t0, err, errlen);
        return rc;
    }
    auto m0 = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpy(ctx->tp.syn_x, bx.data(), bx.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tp.syn_z, bz.data(), bz.size() * 4, hipMemcpyHostToDevice));
    auto m1 = std::chrono::steady_clock::now();
    rc = run_decode(ctx, 1, ctx->noise, max_iter, warmup, interval, damping, precision, err,
                    errlen, nullptr, JBP_ENGINE_AUTO, 0, false, false, freeze_syn != 0);
    if (rc) return rc;
    out->kernel_ms = ctx->last_ms[1];
    auto m2 = std::chrono::steady_clock::now();
    int iters = 0;
    uint8_t ok = 0;
    HIPCHK(hipMemcpy(&iters, ctx->tp.iters, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&ok, ctx->tp.syn_ok, 1, hipMemcpyDeviceToHost));
    std::vector<uint32_t> ex(g.wn), ez(g.wn);
    HIPCHK(hipMemcpy(ex.data(), ctx->tp.est_x, g.wn * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ez.data(), ctx->tp.est_z, g.wn * 4, hipMemcpyDeviceToHost));
    auto m3 = std::chrono::steady_clock::now();
    if (est_out)
        for (int v = 0; v < g.n; ++v)
            est_out[v] = static_cast<int>(((ex[v >> 5] >> (v & 31)) & 1u) |
                                          (((ez[v >> 5] >> (v & 31)) & 1u) << 1));
    out->iterations = iters;
    out->syndrome_match = ok;
    out->check_count = iters;
    out->memcpy_ms = std::chrono::duration<double, std::milli>((m1 - m0) + (m3 - m2)).count();
    out->host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return JBP_OK;
}

int jbp_in_row_space(jbp_ctx *ctx, int side, int count, const int *bits, uint8_t *out, char *err,
                     size_t errlen) {
    if (!ctx || (side != 0 && side != 1) || count < 0 || (count && (!bits || !out))) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (!ctx->have_basis) {
        set_err(err, errlen, "no dense row basis for this code size");
        return JBP_ERR_UNSUPPORTED;
    }
    if (count == 0) return JBP_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, count, err, errlen);
    if (rc) return rc;
    const DevGraph &g = ctx->dg;
    const size_t words = static_cast<size_t>(count) * g.wn;
    if ((rc = ensure_ints(ctx, static_cast<size_t>(count) * g.n, err, errlen))) return rc;
    HIPCHK(hipMemcpy(ctx->d_ints, bits, static_cast<size_t>(count) * g.n * 4,
                     hipMemcpyHostToDevice));
    // The residual kernel tests err_x ^ est_x against rowspace(Hz) and
    // err_z ^ est_z against rowspace(Hx); put the vector in the plane that
    // meets the requested basis and zeros elsewhere.
    uint32_t *target = side == 1 ? ctx->tp.err_x : ctx->tp.err_z;
    uint32_t *other = side == 1 ? ctx->tp.err_z : ctx->tp.err_x;
    launch_pack_bits(ctx->d_ints, g.n, g.wn, count, target, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(other, 0, words * 4, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->tp.est_x, 0, words * 4, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->tp.est_z, 0, words * 4, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->tp.syn_ok, 1, static_cast<size_t>(count), ctx->stream));
    launch_residual(g, ctx->basis[0], ctx->basis[1], ctx->tp, count, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->tp.success, static_cast<size_t>(count), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return JBP_OK;
}

int jbp_simulate(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_trial_out *pt, jbp_run_stats *stats,
                 char *err, size_t errlen) {
    return jbp::simulate_impl(ctx, cfg, pt, stats, nullptr, err, errlen);
}

int jbp_simulate_mixed(jbp_ctx *const *ctxs, const jbp_sim_cfg *cfgs, int n_codes,
                       jbp_trial_out *per_trial, jbp_run_stats *stats, char *err, size_t errlen) {
    auto t0 = std::chrono::steady_clock::now();
    if (!ctxs || !cfgs || !stats || n_codes < 1 || n_codes > kMaxMixCodes) {
        set_err(err, errlen, "bad arguments (1..%d codes)", kMaxMixCodes);
        return JBP_ERR_ARG;
    }
    for (int c = 0; c < n_codes; ++c) {
        const jbp_sim_cfg &cf = cfgs[c];
        if (!ctxs[c] || cf.n_trials < 0 || cf.max_iter < 0 || cf.damping < 0.0) {
            set_err(err, errlen, "bad arguments for code %d", c);
            return JBP_ERR_ARG;
        }
        if (ctxs[c]->device != ctxs[0]->device) {
            set_err(err, errlen, "mixed codes must share one device");
            return JBP_ERR_ARG;
        }
        for (int d = 0; d < c; ++d)
            if (ctxs[d] == ctxs[c]) {
                set_err(err, errlen, "each code of a mixed run needs its own context");
                return JBP_ERR_ARG;
            }
        if (cf.precision != JBP_FP32 || cf.pp || cf.freeze_syn || cf.min_fails > 0) {
            set_err(err, errlen, "a mixed run is FP32 BP without post-processing, freeze-syn or min_fails");
            return JBP_ERR_UNSUPPORTED;
        }
        if (cf.n_trials > kDefaultChunk) {
            set_err(err, errlen, "at most %d trials per code in a mixed run", kDefaultChunk);
            return JBP_ERR_ARG;
        }
        if (cf.n_p < 0 || (cf.n_p > 0 && !cf.p_list) || cf.n_p > kMaxPList) {
            set_err(err, errlen, "p_list needs 1..%d entries", kMaxPList);
            return JBP_ERR_ARG;
        }
        if (!gen_supported(ctxs[c]->dg)) {
            set_err(err, errlen, "generic FP32 decoder supports check degree <= 32");
            return JBP_ERR_UNSUPPORTED;
        }
    }
    jbp_ctx *c0 = ctxs[0];
    HIPCHK(hipSetDevice(c0->device));
    hipStream_t st = c0->stream;
    for (int c = 0; c < n_codes; ++c) HIPCHK(hipStreamSynchronize(ctxs[c]->stream));
    GenMix mix;
    int map[kMaxMixCodes];  // mix entry of code c, -1: no trials
    for (int c = 0; c < n_codes; ++c) {
        jbp_ctx *ctx = ctxs[c];
        const jbp_sim_cfg &cf = cfgs[c];
        map[c] = -1;
        const int np = cf.n_p > 0 ? cf.n_p : 1;
        const double *plist = cf.n_p > 0 ? cf.p_list : &cf.p;
        std::vector<Prior> pri(np);
        for (int i = 0; i < np; ++i) {
            if (!(plist[i] >= 0.0 && plist[i] <= 1.0)) {
                set_err(err, errlen, "p = %g is outside [0, 1]", plist[i]);
                return JBP_ERR_ARG;
            }
            pri[i] = prior_from_p(plist[i]);
        }
        int rc = set_noise(ctx, pri.data(), plist, np, err, errlen);
        if (rc) return rc;
        const int cnt = static_cast<int>(cf.n_trials);
        if (cnt == 0) continue;
        if ((rc = ensure_chunk(ctx, cnt, err, errlen))) return rc;
        if ((rc = ensure_mt(ctx, cnt, err, errlen))) return rc;
        NoiseTab nt = ctx->noise;
        nt.phase = static_cast<int>(((cf.first_trial % np) + np) % np);
        launch_sample(ctx->dg, ctx->tp, cf.base_seed, cf.first_trial, cnt, nt, st, ctx->d_mt, ctx->mt_cap);
        TrialPlanes tpo = ctx->tp;
        if (cf.schedule == 0) {
            launch_order_by_weight(ctx->dg, ctx->tp, cnt, ctx->d_keys, ctx->d_order, st);
            tpo.order = ctx->d_order;
        }
        HIPCHK(hipGetLastError());
        // this code's resident slots and state, as its own engine would have them
        int S = cf.slots > 0 ? cf.slots : default_gen_slots(ctx);
        S = std::min(S, ((cnt + 63) / 64) * 64);
        const int groups = std::max(1, (S + 63) / 64);
        const bool damped = cf.damping > 0.0;
        const size_t bytes = gen_carve_bytes(ctx->dg, groups, damped);
        if (bytes > ctx->gen_cap) {
            if (ctx->gen_mem) (void)hipFree(ctx->gen_mem);
            ctx->gen_mem = nullptr;
            ctx->gen_cap = 0;
            HIPCHK(hipMalloc(&ctx->gen_mem, bytes));
            ctx->gen_cap = bytes;
        }
        gen_carve(ctx->dg, groups, ctx->gen_mem, ctx->gen, damped);
        ctx->gen.damping = static_cast<float>(damped ? cf.damping : 0.0);
        GenCode &k = mix.c[mix.n];
        k.g = ctx->dg;
        k.s = ctx->gen;
        k.tp = tpo;
        k.nt = nt;
        k.count = cnt;
        k.max_iter = cf.max_iter;
        k.warmup = cf.check_warmup;
        k.interval = cf.check_interval;
        map[c] = mix.n++;
    }
    double dev_ms = 0.0;
    if (mix.n > 0) {
        if (!c0->h_ctrl_mix)
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&c0->h_ctrl_mix), 2 * kMaxMixCodes * sizeof(int)));
        for (auto &e : c0->gen_ev)
            if (!e) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(c0->evt[2], st));
        int64_t steps = 0;
        const int rc = gen_decode_mix(mix, c0->h_ctrl_mix, c0->gen_ev, st, &steps);
        if (rc == -1) {
            set_err(err, errlen, "mixed decoder exceeded its step bound");
            return JBP_ERR_DEVICE;
        }
        if (rc != 0) {
            set_err(err, errlen, "mixed decoder: %s",
                    rc == -2 ? "empty mix" : hipGetErrorString(static_cast<hipError_t>(rc)));
            return JBP_ERR_DEVICE;
        }
        HIPCHK(hipEventRecord(c0->evt[3], st));
        HIPCHK(hipEventSynchronize(c0->evt[3]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c0->evt[2], c0->evt[3]));
        dev_ms = ms;
        c0->last_ms[1] = ms;
        c0->launches[1]++;
        c0->gen_steps = steps;
    }
    // each code's success test and outcomes
    for (int c = 0; c < n_codes; ++c) {
        jbp_ctx *ctx = ctxs[c];
        jbp_run_stats &rs = stats[c];
        std::memset(&rs, 0, sizeof(rs));
        rs.device_ms = dev_ms;
        if (map[c] < 0) continue;
        const int cnt = static_cast<int>(cfgs[c].n_trials);
        const DevGraph &g = ctx->dg;
        const bool judged = ctx->have_logicals || ctx->have_basis;
        if (ctx->have_logicals) launch_logical(g, ctx->lg[0], ctx->lg[1], ctx->tp, cnt, st);
        else if (ctx->have_basis) launch_residual(g, ctx->basis[0], ctx->basis[1], ctx->tp, cnt, st);
        else launch_logical(g, DevLogicals(), DevLogicals(), ctx->tp, cnt, st);
        HIPCHK(hipGetLastError());
        const int cap = ctx->chunk_cap;
        HIPCHK(hipMemcpyAsync(ctx->h_iters, ctx->tp.iters, cnt * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(ctx->h_flags, ctx->tp.syn_ok, 3 * static_cast<size_t>(cap),
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        jbp_trial_out *pt = per_trial ? &per_trial[c] : nullptr;
        if (pt && (pt->err || pt->est)) {
            int rc = ensure_states(ctx, static_cast<size_t>(cnt) * g.n, err, errlen);
            if (rc) return rc;
            for (int which = 0; which < 2; ++which) {
                int *dst = which == 0 ? pt->err : pt->est;
                if (!dst) continue;
                const uint32_t *xb = which == 0 ? ctx->tp.err_x : ctx->tp.est_x;
                const uint32_t *zb = which == 0 ? ctx->tp.err_z : ctx->tp.est_z;
                launch_unpack_states(xb, zb, g.wn, g.n, cnt, ctx->d_states, st);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(dst, ctx->d_states, static_cast<size_t>(cnt) * g.n * 4,
                                      hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
            }
        }
        const uint8_t *ok = ctx->h_flags, *succ = ctx->h_flags + cap, *exact = ctx->h_flags + 2 * cap;
        for (int i = 0; i < cnt; ++i) {
            const uint8_t sc = judged ? succ[i] : static_cast<uint8_t>(ok[i] && exact[i]);
            rs.total_iters += ctx->h_iters[i];
            rs.failures += sc ? 0 : 1;
            rs.bp_failures += ok[i] ? 0 : 1;
            rs.exact += exact[i] ? 1 : 0;
            rs.stab_success += (sc && !exact[i] && judged) ? 1 : 0;
            if (pt) {
                if (pt->pp_flags) pt->pp_flags[i] = ok[i] ? JBP_PPF_BP_OK : 0;
                if (pt->iterations) pt->iterations[i] = ctx->h_iters[i];
                if (pt->syn_match) pt->syn_match[i] = ok[i];
                if (pt->success) pt->success[i] = sc;
                if (pt->exact) pt->exact[i] = exact[i];
            }
        }
        rs.done = cnt;
    }
    const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int c = 0; c < n_codes; ++c) stats[c].elapsed_ms = el;
    return JBP_OK;
}

int jbp_set_logicals(jbp_ctx *ctx, int side, int64_t n_rows, const int64_t *row_off,
                     const int *qubits, char *err, size_t errlen) {
    if (!ctx || side < -1 || side > 1 || n_rows < 0 || (n_rows > 0 && (!row_off || !qubits))) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    HIPCHK(hipSetDevice(ctx->device));
    if (side == -1) {
        free_logicals(ctx);
        return JBP_OK;
    }
    const int64_t total = n_rows > 0 ? row_off[n_rows] : 0;
    if (n_rows > 0 && row_off[0] != 0) {
        set_err(err, errlen, "row_off[0] must be 0");
        return JBP_ERR_ARG;
    }
    for (int64_t r = 0; r < n_rows; ++r)
        if (row_off[r + 1] < row_off[r]) {
            set_err(err, errlen, "row offsets decrease at row %lld", static_cast<long long>(r));
            return JBP_ERR_ARG;
        }
    for (int64_t e = 0; e < total; ++e)
        if (qubits[e] < 0 || qubits[e] >= ctx->n) {
            set_err(err, errlen, "qubit %d out of range", qubits[e]);
            return JBP_ERR_ARG;
        }
    for (void *&p : ctx->lg_mem[side]) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    ctx->lg[side] = DevLogicals();
    ctx->lg_set[side] = false;
    ctx->have_logicals = false;
    HIPCHK(hipMalloc(&ctx->lg_mem[side][0], (n_rows + 1) * sizeof(int64_t)));
    HIPCHK(hipMalloc(&ctx->lg_mem[side][1], std::max<int64_t>(total, 1) * sizeof(int)));
    std::vector<int64_t> off(row_off ? row_off : nullptr, row_off ? row_off + n_rows + 1 : nullptr);
    if (off.empty()) off.push_back(0);
    HIPCHK(hipMemcpy(ctx->lg_mem[side][0], off.data(), off.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    if (total > 0)
        HIPCHK(hipMemcpy(ctx->lg_mem[side][1], qubits, total * sizeof(int), hipMemcpyHostToDevice));
    ctx->lg[side].rows = n_rows;
    ctx->lg[side].off = static_cast<const int64_t *>(ctx->lg_mem[side][0]);
    ctx->lg[side].qubit = static_cast<const int *>(ctx->lg_mem[side][1]);
    ctx->lg_set[side] = true;
    ctx->have_logicals = ctx->lg_set[0] && ctx->lg_set[1];
    return JBP_OK;
}

int jbp_tune(jbp_ctx *ctx, const jbp_sim_cfg *cfg, int best[2], double *cand_ms, int *n_cand,
             char *err, size_t errlen) {
    if (!ctx || !cfg || !best) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    const bool fp32 = cfg->precision == JBP_FP32;
    const bool hbm = fp32 && (cfg->engine == JBP_ENGINE_HBM ||
                              (cfg->engine == JBP_ENGINE_AUTO && !ctx->qc));
    std::vector<std::pair<int, int>> cands;  // (slots, block)
    if (hbm) {
        const int cap = std::max(64, static_cast<int>(((cfg->n_trials + 63) / 64) * 64));
        for (int S = 64; S <= std::min(cap, 64 * 512); S *= 2)
            if (gen_carve_bytes(ctx->dg, S / 64) <= (size_t(8) << 30)) cands.push_back({S, 0});
    } else if (!fp32) {
        const int cus = std::max(1, ctx->cus);
        for (int mult : {1, 2, 4})
            if (decode_f64_scratch_bytes(ctx->dg, mult * cus) <= (size_t(8) << 30))
                for (int blk : {64, 128, 256, 512}) cands.push_back({mult * cus, blk});
    }
    if (cands.empty()) {  // the LDS decoder: one 768-thread workgroup per compute unit
        best[0] = 0;
        best[1] = 0;
        if (n_cand) *n_cand = 0;
        return JBP_OK;
    }
    if (n_cand && *n_cand < static_cast<int>(cands.size())) {
        set_err(err, errlen, "cand_ms needs %zu entries", cands.size());
        return JBP_ERR_ARG;
    }
    double best_ms = 0.0;
    int bi = -1;
    for (size_t i = 0; i < cands.size(); ++i) {
        jbp_sim_cfg c = *cfg;
        c.slots = cands[i].first;
        c.block = cands[i].second;
        c.min_fails = 0;
        jbp_run_stats st;
        int rc = jbp::simulate_impl(ctx, &c, nullptr, &st, nullptr, err, errlen);
        if (rc) return rc;
        if (cand_ms) cand_ms[i] = st.device_ms;
        if (bi < 0 || st.device_ms < best_ms) {
            best_ms = st.device_ms;
            bi = static_cast<int>(i);
        }
    }
    if (n_cand) *n_cand = static_cast<int>(cands.size());
    best[0] = cands[bi].first;
    best[1] = cands[bi].second;
    if (hbm) {
        ctx->gen_slots = best[0];
    } else {
        if (ctx->f64_slots != best[0] && ctx->f64_scratch) {
            (void)hipFree(ctx->f64_scratch);
            ctx->f64_scratch = nullptr;
        }
        ctx->f64_slots = best[0];
        ctx->f64_block = best[1];
    }
    return JBP_OK;
}

}  // extern "C"
