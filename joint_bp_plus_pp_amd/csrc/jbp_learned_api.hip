// Host side of the learned decoder: model parameters, the loss / gradient
// entry point, device-sampled training steps (forward + adjoint kernels and
// Adam on the device), and simulation with learned weights followed by the
// plain-BP fallback.  No counterpart in the reference program.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "jbp_ctx.h"

using namespace jbp;

struct jbp_learned {
    jbp_ctx *ctx = nullptr;
    int T = 0, nb = 0, use_mlp = 1, count = 0;
    int kind = kLearnedSP;
    uint32_t *d_ms_rec = nullptr;      // min-sum decoder: FP16 weight records
    std::vector<double> params;        // host copy of the master parameters
    double *d_p64 = nullptr;           // device master (double)
    float *d_p32 = nullptr;            // device copy read by the FP32 training kernels
    float *d_infer = nullptr;          // iteration-major copy read by the decoder
    double *d_m = nullptr, *d_v = nullptr;
    int step = 0;
    int *d_blk[2] = {nullptr, nullptr};
    // training buffers (grown on demand)
    int batch_cap = 0, prec_cap = 0;
    std::vector<void *> bufs;
    void *d_act = nullptr, *d_scratch = nullptr, *d_tape_d = nullptr, *d_tape_u = nullptr;
    void *d_chk = nullptr;             // check records of the on-chip FP32 trainer
    void *d_loss = nullptr, *d_grad = nullptr;
    double *d_grad64 = nullptr, *d_prior = nullptr;
    DrawThresholds *d_th = nullptr;
    int *d_conv = nullptr, *d_counters = nullptr;
    int slots = 0;
};

namespace {

uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

void prior4(double p, double *out) {
    double m[4] = {1.0 - p, p / 3.0, p / 3.0, p / 3.0};
    double s = m[0] + m[1] + m[2] + m[3];
    for (int i = 0; i < 4; ++i) out[i] = s <= 0.0 ? 0.25 : m[i] / s;
}

void free_bufs(jbp_learned *m) {
    for (void *p : m->bufs) (void)hipFree(p);
    m->bufs.clear();
    m->batch_cap = 0;
    m->prec_cap = 0;
}

LearnedDevParams dev_params(const jbp_learned *m);

int ensure_train(jbp_learned *m, int batch, int precision, char *err, size_t errlen) {
    if (batch <= m->batch_cap && precision == m->prec_cap) return JBP_OK;
    free_bufs(m);
    jbp_ctx *ctx = m->ctx;
    const size_t R = precision == JBP_FP64 ? sizeof(double) : sizeof(float);
    const size_t Et = static_cast<size_t>(ctx->dg.side[0].E) + ctx->dg.side[1].E;
    const size_t b = static_cast<size_t>(batch);
    m->slots = std::min(batch, std::max(1, 2 * ctx->cus));
    auto alloc = [&](size_t bytes, void **p) -> int {
        HIPCHK(hipMalloc(p, std::max<size_t>(bytes, 8)));
        m->bufs.push_back(*p);
        return JBP_OK;
    };
    int rc = JBP_OK;
    void *p = nullptr;
    if ((rc = alloc(b * (m->T + 1) * Et * R, &m->d_act))) return rc;
    if (precision != JBP_FP64 && learned_qc_train_supported(ctx->dg, dev_params(m))) {
        // the on-chip trainer keeps no per-slot scratch, and records its
        // check rows for the adjoint
        m->d_scratch = nullptr;
        if ((rc = alloc(learned_qc_chk_bytes(batch, m->T), &m->d_chk))) return rc;
    } else {
        m->d_chk = nullptr;
        if ((rc = alloc(static_cast<size_t>(m->slots) * learned_slot_elems(ctx->dg) * R, &m->d_scratch)))
            return rc;
    }
    if ((rc = alloc(b * m->T * R, &m->d_tape_d))) return rc;
    if ((rc = alloc(b * m->T * R, &m->d_tape_u))) return rc;
    if ((rc = alloc(R, &m->d_loss))) return rc;
    if ((rc = alloc(m->count * R, &m->d_grad))) return rc;
    if ((rc = alloc(m->count * sizeof(double), &p))) return rc;
    m->d_grad64 = static_cast<double *>(p);
    if ((rc = alloc(b * 4 * sizeof(double), &p))) return rc;
    m->d_prior = static_cast<double *>(p);
    if ((rc = alloc(b * sizeof(DrawThresholds), &p))) return rc;
    m->d_th = static_cast<DrawThresholds *>(p);
    if ((rc = alloc(b * sizeof(int), &p))) return rc;
    m->d_conv = static_cast<int *>(p);
    if ((rc = alloc(2 * sizeof(int), &p))) return rc;
    m->d_counters = static_cast<int *>(p);
    m->batch_cap = batch;
    m->prec_cap = precision;
    return JBP_OK;
}

int upload_params(jbp_learned *m, char *err, size_t errlen) {
    std::vector<float> pf(m->params.begin(), m->params.end());
    HIPCHK(hipMemcpy(m->d_p64, m->params.data(), m->count * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->d_p32, pf.data(), m->count * sizeof(float), hipMemcpyHostToDevice));
    return JBP_OK;
}

LearnedDevParams dev_params(const jbp_learned *m) {
    LearnedDevParams lp;
    lp.nb = m->nb;
    lp.T = m->T;
    lp.use_mlp = m->use_mlp;
    lp.kind = m->kind;
    lp.lpb = m->ctx->qcp.L;
    lp.blk0 = m->d_blk[0];
    lp.blk1 = m->d_blk[1];
    return lp;
}

const float *params_of(const jbp_learned *m, float *) { return m->d_p32; }
const double *params_of(const jbp_learned *m, double *) { return m->d_p64; }

// Forward (+ adjoint) over the trials whose planes and priors are on the
// device; loss and gradient land in d_loss / d_grad64.
template <typename R>
int run_train(jbp_learned *m, int batch, bool backward, char *err, size_t errlen) {
    jbp_ctx *ctx = m->ctx;
    const R *params = params_of(m, static_cast<R *>(nullptr));
    HIPCHK(hipMemsetAsync(m->d_loss, 0, sizeof(R), ctx->stream));
    HIPCHK(hipMemsetAsync(m->d_grad, 0, m->count * sizeof(R), ctx->stream));
    const LearnedDevParams lp = dev_params(m);
    if (sizeof(R) == sizeof(float) && learned_qc_train_supported(ctx->dg, lp)) {
        launch_learned_train_qc(ctx->dg, lp, reinterpret_cast<const float *>(params), ctx->tp, batch,
                                static_cast<float *>(m->d_act), static_cast<uint32_t *>(m->d_chk),
                                static_cast<float *>(m->d_tape_d),
                                static_cast<float *>(m->d_tape_u), m->d_prior, static_cast<float *>(m->d_loss),
                                static_cast<float *>(m->d_grad), m->d_conv, m->d_counters, backward,
                                ctx->stream);
        HIPCHK(hipGetLastError());
        return JBP_OK;
    }
    launch_learned_train<R>(ctx->dg, lp, params, ctx->tp, batch, static_cast<R *>(m->d_act),
                            static_cast<R *>(m->d_scratch), m->slots, static_cast<R *>(m->d_tape_d),
                            static_cast<R *>(m->d_tape_u), m->d_prior, static_cast<R *>(m->d_loss),
                            static_cast<R *>(m->d_grad), m->d_conv, m->d_counters, backward,
                            ctx->stream);
    HIPCHK(hipGetLastError());
    return JBP_OK;
}

}  // namespace

extern "C" {

int jbp_learned_create(jbp_ctx *ctx, int T, jbp_learned **out, char *err, size_t errlen) {
    return jbp_learned_create_kind(ctx, T, JBP_LEARNED_SUM_PRODUCT, out, err, errlen);
}

int jbp_learned_create_kind(jbp_ctx *ctx, int T, int kind, jbp_learned **out, char *err,
                            size_t errlen) {
    if (kind != JBP_LEARNED_SUM_PRODUCT && kind != JBP_LEARNED_MIN_SUM &&
        kind != JBP_LEARNED_LAYERED_MIN_SUM) {
        set_err(err, errlen, "unknown learned-decoder kind %d", kind);
        return JBP_ERR_ARG;
    }
    if (!ctx || !out || T < 1 || T > kMaxLearnedT) {
        set_err(err, errlen, "learned model needs a context and 1 <= T <= %d", kMaxLearnedT);
        return JBP_ERR_ARG;
    }
    *out = nullptr;
    if (!ctx->qc) {
        set_err(err, errlen, "learned BP needs a permutation-block code");
        return JBP_ERR_UNSUPPORTED;
    }
    const int P = ctx->qcp.P, L = ctx->qcp.L, nb = ctx->qcp.J * ctx->qcp.L;
    if (nb > kMaxLearnedBlocks || ctx->dg.max_cdeg > 16) {
        set_err(err, errlen, "learned BP supports at most %d blocks and check degree 16",
                kMaxLearnedBlocks);
        return JBP_ERR_UNSUPPORTED;
    }
    HIPCHK(hipSetDevice(ctx->device));
    auto *m = new jbp_learned();
    m->ctx = ctx;
    m->kind = kind == JBP_LEARNED_MIN_SUM ? kLearnedMS
              : kind == JBP_LEARNED_LAYERED_MIN_SUM ? kLearnedLMS : kLearnedSP;
    if (m->kind == kLearnedLMS) m->use_mlp = 0;
    m->T = T;
    m->nb = nb;
    m->count = learned_param_count(nb, T);
    m->params.assign(m->count, 0.0);
    for (int i = 0; i < 4 * nb * T; ++i) m->params[i] = 1.0;
    // MLP: small deterministic first layer, damping ~ sigmoid(-4) at start
    double *mlp = m->params.data() + 4 * nb * T;
    for (int i = 0; i < 2 * kMlpHidden; ++i)
        mlp[i] = 0.5 * (static_cast<double>(mix64(1234 + i) >> 11) * 0x1p-53 - 0.5);
    for (int i = 0; i < kMlpHidden; ++i)
        mlp[3 * kMlpHidden + i] = 0.1 * (static_cast<double>(mix64(5678 + i) >> 11) * 0x1p-53 - 0.5);
    mlp[4 * kMlpHidden] = -4.0;
    auto fail = [&](int code) {
        jbp_learned_destroy(m);
        return code;
    };
    void *p = nullptr;
    if (hipMalloc(&p, m->count * sizeof(double)) != hipSuccess) return fail(JBP_ERR_NOMEM);
    m->d_p64 = static_cast<double *>(p);
    if (hipMalloc(&p, m->count * sizeof(float)) != hipSuccess) return fail(JBP_ERR_NOMEM);
    m->d_p32 = static_cast<float *>(p);
    if (hipMalloc(&p, m->count * sizeof(float)) != hipSuccess) return fail(JBP_ERR_NOMEM);
    m->d_infer = static_cast<float *>(p);
    if (hipMalloc(&p, static_cast<size_t>(T + 1) * kMsRec * sizeof(uint32_t)) != hipSuccess)
        return fail(JBP_ERR_NOMEM);
    m->d_ms_rec = static_cast<uint32_t *>(p);
    if (hipMalloc(&p, m->count * sizeof(double)) != hipSuccess) return fail(JBP_ERR_NOMEM);
    m->d_m = static_cast<double *>(p);
    if (hipMalloc(&p, m->count * sizeof(double)) != hipSuccess) return fail(JBP_ERR_NOMEM);
    m->d_v = static_cast<double *>(p);
    if (hipMemset(m->d_m, 0, m->count * sizeof(double)) != hipSuccess ||
        hipMemset(m->d_v, 0, m->count * sizeof(double)) != hipSuccess)
        return fail(JBP_ERR_DEVICE);
    for (int s = 0; s < 2; ++s) {
        const HostSide &h = ctx->hs[s];
        std::vector<int> blk(h.E);
        for (int c = 0; c < h.m; ++c)
            for (int e = h.chk_off[c]; e < h.chk_off[c + 1]; ++e) blk[e] = (c / P) * L + (e - h.chk_off[c]);
        if (hipMalloc(&p, std::max<size_t>(1, blk.size()) * sizeof(int)) != hipSuccess) return fail(JBP_ERR_NOMEM);
        m->d_blk[s] = static_cast<int *>(p);
        if (hipMemcpy(p, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return fail(JBP_ERR_DEVICE);
    }
    int rc = upload_params(m, err, errlen);
    if (rc) return fail(rc);
    *out = m;
    return JBP_OK;
}

void jbp_learned_destroy(jbp_learned *m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    free_bufs(m);
    for (void *p : {static_cast<void *>(m->d_p64), static_cast<void *>(m->d_p32),
                    static_cast<void *>(m->d_m), static_cast<void *>(m->d_v),
                    static_cast<void *>(m->d_infer), static_cast<void *>(m->d_ms_rec),
                    static_cast<void *>(m->d_blk[0]), static_cast<void *>(m->d_blk[1])})
        if (p) (void)hipFree(p);
    delete m;
}

int jbp_learned_param_count(const jbp_learned *m) { return m ? m->count : -1; }

int jbp_learned_kind(const jbp_learned *m) {
    return !m ? -1
           : m->kind == kLearnedMS  ? JBP_LEARNED_MIN_SUM
           : m->kind == kLearnedLMS ? JBP_LEARNED_LAYERED_MIN_SUM
                                    : JBP_LEARNED_SUM_PRODUCT;
}

int jbp_learned_set_params(jbp_learned *m, const double *params, int use_mlp, char *err,
                           size_t errlen) {
    if (!m || !params) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (use_mlp && m->kind == kLearnedLMS) {
        set_err(err, errlen, "the layered min-sum model has no damping (use_mlp must be 0)");
        return JBP_ERR_UNSUPPORTED;
    }
    HIPCHK(hipSetDevice(m->ctx->device));
    m->params.assign(params, params + m->count);
    m->use_mlp = use_mlp ? 1 : 0;
    return upload_params(m, err, errlen);
}

int jbp_learned_get_params(const jbp_learned *m, double *params) {
    if (!m || !params) return JBP_ERR_ARG;
    std::copy(m->params.begin(), m->params.end(), params);
    return JBP_OK;
}

int jbp_learned_loss_grad(jbp_learned *m, int precision, int batch, const int *err_states,
                          const double *p, double *loss, double *grad, int *conv_out, char *err,
                          size_t errlen) {
    if (!m || batch <= 0 || !err_states || !p || !loss ||
        (precision != JBP_FP32 && precision != JBP_FP64)) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    jbp_ctx *ctx = m->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, batch, err, errlen);
    if (!rc) rc = ensure_train(m, batch, precision, err, errlen);
    const int n = ctx->n;
    if (!rc) rc = ensure_ints(ctx, static_cast<size_t>(batch) * n, err, errlen);
    if (rc) return rc;
    std::vector<int> bits(static_cast<size_t>(batch) * n);
    for (int plane = 0; plane < 2; ++plane) {
        for (size_t i = 0; i < bits.size(); ++i) {
            int st = err_states[i];
            if (st < 0 || st > 3) {
                set_err(err, errlen, "error states must be 0..3");
                return JBP_ERR_ARG;
            }
            bits[i] = plane == 0 ? (st & 1) : (st >> 1);
        }
        HIPCHK(hipMemcpy(ctx->d_ints, bits.data(), bits.size() * sizeof(int), hipMemcpyHostToDevice));
        launch_pack_bits(ctx->d_ints, n, ctx->dg.wn, batch, plane == 0 ? ctx->tp.err_x : ctx->tp.err_z,
                         ctx->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    launch_syndromes(ctx->dg, ctx->tp, batch, ctx->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> pr(4 * static_cast<size_t>(batch));
    for (int b = 0; b < batch; ++b) prior4(p[b], &pr[4 * b]);
    HIPCHK(hipMemcpy(m->d_prior, pr.data(), pr.size() * sizeof(double), hipMemcpyHostToDevice));
    rc = precision == JBP_FP64 ? run_train<double>(m, batch, grad != nullptr, err, errlen)
                               : run_train<float>(m, batch, grad != nullptr, err, errlen);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (precision == JBP_FP64) {
        HIPCHK(hipMemcpy(loss, m->d_loss, sizeof(double), hipMemcpyDeviceToHost));
        if (grad) HIPCHK(hipMemcpy(grad, m->d_grad, m->count * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        float lf = 0.f;
        HIPCHK(hipMemcpy(&lf, m->d_loss, sizeof(float), hipMemcpyDeviceToHost));
        *loss = lf;
        if (grad) {
            std::vector<float> gf(m->count);
            HIPCHK(hipMemcpy(gf.data(), m->d_grad, m->count * sizeof(float), hipMemcpyDeviceToHost));
            std::copy(gf.begin(), gf.end(), grad);
        }
    }
    if (conv_out) HIPCHK(hipMemcpy(conv_out, m->d_conv, batch * sizeof(int), hipMemcpyDeviceToHost));
    return JBP_OK;
}

int jbp_learned_train_step(jbp_learned *m, uint64_t base_seed, int64_t first_trial, int batch,
                           double p_lo, double p_hi, double lr, double *loss, char *err,
                           size_t errlen) {
    if (!m || batch <= 0 || !(p_lo > 0.0) || p_hi < p_lo || p_hi >= 0.75) {
        set_err(err, errlen, "bad arguments (need 0 < p_lo <= p_hi < 0.75)");
        return JBP_ERR_ARG;
    }
    jbp_ctx *ctx = m->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_chunk(ctx, batch, err, errlen);
    if (!rc) rc = ensure_train(m, batch, JBP_FP32, err, errlen);
    if (rc) return rc;
    // per-trial p from its own seeded stream, independent of the error draws
    std::vector<double> pr(4 * static_cast<size_t>(batch));
    std::vector<DrawThresholds> th(batch);
    for (int b = 0; b < batch; ++b) {
        uint64_t idx = static_cast<uint64_t>(first_trial + b);
        double u = static_cast<double>(mix64(base_seed ^ 0xd1b54a32d192ed03ull ^ (idx * 0x9e3779b97f4a7c15ull)) >> 11) * 0x1p-53;
        double pb = p_lo + (p_hi - p_lo) * u;
        th[b] = draw_thresholds(pb);
        prior4(pb, &pr[4 * b]);
    }
    HIPCHK(hipMemcpy(m->d_th, th.data(), th.size() * sizeof(DrawThresholds), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->d_prior, pr.data(), pr.size() * sizeof(double), hipMemcpyHostToDevice));
    NoiseTab nt;
    nt.th = m->d_th;
    nt.n = batch;
    if ((rc = ensure_mt(ctx, batch, err, errlen))) return rc;
    launch_sample(ctx->dg, ctx->tp, base_seed, first_trial, batch, nt, ctx->stream, ctx->d_mt, ctx->mt_cap);
    HIPCHK(hipGetLastError());
    if ((rc = run_train<float>(m, batch, true, err, errlen))) return rc;
    launch_to_double(static_cast<const float *>(m->d_grad), m->d_grad64, m->count, ctx->stream);
    HIPCHK(hipGetLastError());
    m->step += 1;
    launch_adam(m->d_p64, m->d_grad64, m->d_m, m->d_v, m->d_p32, m->count, lr, m->step, ctx->stream);
    HIPCHK(hipGetLastError());
    float lf = 0.f;
    HIPCHK(hipMemcpyAsync(&lf, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(m->params.data(), m->d_p64, m->count * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = lf;
    return JBP_OK;
}

int jbp_simulate_learned(jbp_ctx *ctx, const jbp_learned *m, const jbp_sim_cfg *cfg,
                         jbp_trial_out *per_trial, jbp_run_stats *stats, char *err, size_t errlen) {
    if (!ctx || !m || !cfg || m->ctx != ctx) {
        set_err(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (cfg->precision != JBP_FP32 || cfg->damping > 0.0) {
        set_err(err, errlen, "the learned decoder runs FP32 with damping 0 (its own damping is learned)");
        return JBP_ERR_UNSUPPORTED;
    }
    // iteration-major copy of the current parameters for the decoder
    std::vector<float> inf(m->count);
    const int nb = m->nb, T = m->T;
    for (int t = 0; t < T; ++t)
        for (int part = 0; part < 4; ++part)
            for (int b = 0; b < nb; ++b)
                inf[(static_cast<size_t>(t) * 4 + part) * nb + b] =
                    static_cast<float>(m->params[(static_cast<size_t>(part) * nb + b) * T + t]);
    for (int i = 4 * nb * T; i < m->count; ++i) inf[i] = static_cast<float>(m->params[i]);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpy(m->d_infer, inf.data(), inf.size() * sizeof(float), hipMemcpyHostToDevice));
    LearnedInfer li;
    li.params = m->d_infer;
    li.T = m->T;
    li.use_mlp = m->use_mlp;
    li.kind = m->kind;
    if (m->kind == kLearnedMS || m->kind == kLearnedLMS) {
        // min-sum: variable pass k: (wc X, wc Z) of iteration k-1 per row
        // block, then (wv X, wv Z) of iteration k, FP16 pairs (1 where there
        // is none).  Layered: record k = iteration k+1's own weights, one
        // 32-word group per block row r: (wc X, wc Z) per column block, then
        // (wv X, wv Z) per column block, then padding.
        if (ctx->qcp.J != 3 || ctx->qcp.L != 12) {
            set_err(err, errlen, "the min-sum decoder is built for 3 x 12 blocks");
            return JBP_ERR_UNSUPPORTED;
        }
        const int J = ctx->qcp.J, L = ctx->qcp.L;
        auto h16 = [](double v) -> uint32_t {
            const _Float16 h = static_cast<_Float16>(static_cast<float>(v));
            return __builtin_bit_cast(uint16_t, h);
        };
        auto pair = [&](double x, double z) { return h16(x) | (h16(z) << 16); };
        std::vector<uint32_t> rec(static_cast<size_t>(T + 1) * kMsRec, 0u);
        const bool layered = m->kind == kLearnedLMS;
        auto par = [&](int part, int b, int t) {
            return m->params[(static_cast<size_t>(part) * nb + b) * T + t];
        };
        for (int k = 0; k <= T; ++k)
            for (int cc = 0; cc < L; ++cc)
                for (int r = 0; r < J; ++r) {
                    const int b = r * L + cc;
                    uint32_t *w = &rec[static_cast<size_t>(k) * kMsRec + cc * kMsRecCC];
                    if (layered) {
                        if (k == T) continue;
                        uint32_t *lw = &rec[static_cast<size_t>(k) * kMsRec + r * 32];
                        lw[cc] = pair(par(0, b, k), par(1, b, k));
                        lw[12 + cc] = pair(par(2, b, k), par(3, b, k));
                        continue;
                    }
                    w[r] = k == 0 ? pair(1.0, 1.0) : pair(par(0, b, k - 1), par(1, b, k - 1));
                    w[3 + r] = k == T ? pair(1.0, 1.0) : pair(par(2, b, k), par(3, b, k));
                }
        HIPCHK(hipMemcpy(m->d_ms_rec, rec.data(), rec.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        li.ms_rec = m->d_ms_rec;
    }
    return simulate_impl(ctx, cfg, per_trial, stats, &li, err, errlen);
}

}  // extern "C"
