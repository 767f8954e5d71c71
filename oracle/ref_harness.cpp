// Test infrastructure only: a C entry layer over the reference decoder, built
// from the reference's own source file where it lies (the Makefile in this
// directory passes its directory with -I and renames its main()).  Output goes
// to oracle/_ref/ only.  Nothing in the product links or loads this.
//
// Every function below calls the reference's own routines:
//   read_params / build_block_matrix_hx/hz / build_check_neighbors /
//   build_var_adjacency / build_row_basis   jointbp_ets.cpp:98-180, 628-648, 845-935
//   trial_seed_from_base + mt19937_64 + uniform_real_distribution sampling
//                                           jointbp_ets.cpp:4408-4411, 5931-5946
//   compute_syndrome                        jointbp_ets.cpp:1834-1849
//   joint_bp_decode (PP off)                jointbp_ets.cpp:1943-3079
//   fill_diff_bits / in_row_space           jointbp_ets.cpp:887-920, 6199-6207
//   compute_girth / orthogonality           jointbp_ets.cpp:1018-1171
//   read_ets_file                           jointbp_ets.cpp:3199-3258
//   attempt_ets_fix_side_multi / attempt_flip_pp_side / attempt_osd_pp_side
//                                           jointbp_ets.cpp:3644, 3726, 4078
//   joint_bp_decode (PP on)                 jointbp_ets.cpp:1943-3079
// The only code written here is the loop glue that main() has inline
// (jointbp_ets.cpp:5931-5948, 5993-6228), restated line for line.
#include "jointbp_ets.cpp"

#include <atomic>
#include <memory>
#include <thread>

namespace {

struct RefState {
    std::vector<std::vector<int>> x_checks, z_checks;
    std::vector<std::vector<EdgeRef>> var_to_x, var_to_z;
    RowBasis hx_basis, hz_basis;
    int nvars = 0;
    bool ready = false;
};

RefState g_ref;

// Trapping-set tables loaded by the reference's reader, in load order.
struct RefEtsGroup {
    int size = 0;
    std::string label;
    std::unique_ptr<ETSIndex> x, z;
};
std::vector<RefEtsGroup> g_ets;

std::vector<ETSIndexRef> ets_refs() {
    std::vector<ETSIndexRef> refs;
    for (const auto &g : g_ets) {
        ETSIndexRef r;
        r.size = g.size;
        r.label = g.label;
        r.x = g.x.get();
        r.z = g.z.get();
        if (r.x || r.z) refs.push_back(r);
    }
    return refs;
}

void finish_code(int nvars) {
    g_ref.nvars = nvars;
    g_ref.var_to_x = build_var_adjacency(g_ref.x_checks, nvars);
    g_ref.var_to_z = build_var_adjacency(g_ref.z_checks, nvars);
    g_ref.hx_basis = build_row_basis(g_ref.x_checks, nvars);
    g_ref.hz_basis = build_row_basis(g_ref.z_checks, nvars);
    g_ref.ready = true;
}

Msg make_prior(double p_err) {
    // jointbp_ets.cpp:5484-5485
    Msg prior{1.0 - p_err, p_err / 3.0, p_err / 3.0, p_err / 3.0};
    normalize_msg(prior);
    return prior;
}

void sample_trial(uint64_t base_seed, long long trial_idx, double p_err, std::vector<int> &err) {
    // jointbp_ets.cpp:5933-5946
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    uint64_t trial_seed = trial_seed_from_base(base_seed, static_cast<uint64_t>(trial_idx));
    std::mt19937_64 rng(trial_seed);
    err.assign(g_ref.nvars, 0);
    for (int i = 0; i < g_ref.nvars; ++i) {
        double u = dist(rng);
        if (u < (1.0 - p_err)) {
            err[i] = 0;
        } else {
            double v = (u - (1.0 - p_err)) / p_err;
            if (v < 1.0 / 3.0) err[i] = 1;
            else if (v < 2.0 / 3.0) err[i] = 2;
            else err[i] = 3;
        }
    }
}

JointBPResult decode_plain(const std::vector<int> &sx, const std::vector<int> &sz,
                           const Msg &prior, int max_iter, double damping, bool freeze_syn) {
    return joint_bp_decode(g_ref.x_checks, g_ref.z_checks, g_ref.var_to_x, g_ref.var_to_z,
                           sx, sz, prior, max_iter, /*flip_hist_window=*/5, freeze_syn, damping,
                           false, false, false, nullptr, nullptr,
                           &g_ref.hx_basis, &g_ref.hz_basis, nullptr, nullptr);
}

bool stab_equiv(const std::vector<int> &err, const std::vector<int> &est) {
    std::vector<uint64_t> dx(g_ref.hz_basis.words, 0), dz(g_ref.hx_basis.words, 0);
    fill_diff_bits(dx, err, est, true);
    fill_diff_bits(dz, err, est, false);
    return in_row_space(g_ref.hz_basis, dx) && in_row_space(g_ref.hx_basis, dz);
}

}  // namespace

extern "C" {

// Build the APM code from a params file with the reference's own constructors.
int ref_load_params(const char *path) {
    Params params;
    if (!read_params(path, params)) return -1;
    auto hx_blocks = build_block_matrix_hx(params.P, params.J, params.L2, params.a, params.b, params.c, params.d);
    auto hz_blocks = build_block_matrix_hz(params.P, params.J, params.L2, params.a, params.b, params.c, params.d);
    g_ref.x_checks = build_check_neighbors(hx_blocks, params.P);
    g_ref.z_checks = build_check_neighbors(hz_blocks, params.P);
    finish_code(params.L * static_cast<int>(params.P));
    return 0;
}

// Install an arbitrary code from flat check lists (row-major, `deg` entries
// per check, -1 = absent), e.g. a hypergraph-product code or a tiny code.
int ref_set_checks(int nvars, int mx, int xdeg, const int *xflat, int mz, int zdeg, const int *zflat) {
    g_ref.x_checks.assign(mx, std::vector<int>(xdeg));
    g_ref.z_checks.assign(mz, std::vector<int>(zdeg));
    for (int c = 0; c < mx; ++c)
        for (int i = 0; i < xdeg; ++i) g_ref.x_checks[c][i] = xflat[c * xdeg + i];
    for (int c = 0; c < mz; ++c)
        for (int i = 0; i < zdeg; ++i) g_ref.z_checks[c][i] = zflat[c * zdeg + i];
    finish_code(nvars);
    return 0;
}

// out[0..8] = n, mx, mz, rank_x, rank_z, edges_x, edges_z, orthogonal, odd_pairs
int ref_code_info(long long *out, int with_girth, int *girth_out) {
    if (!g_ref.ready) return -1;
    OrthCheckResult orth = check_hx_hz_orthogonality(g_ref.x_checks, g_ref.z_checks, g_ref.nvars);
    DegreeStats xs = compute_var_degree_stats(g_ref.var_to_x);
    DegreeStats zs = compute_var_degree_stats(g_ref.var_to_z);
    out[0] = g_ref.nvars;
    out[1] = static_cast<long long>(g_ref.x_checks.size());
    out[2] = static_cast<long long>(g_ref.z_checks.size());
    out[3] = static_cast<long long>(g_ref.hx_basis.rows.size());
    out[4] = static_cast<long long>(g_ref.hz_basis.rows.size());
    out[5] = xs.edges;
    out[6] = zs.edges;
    out[7] = orth.orthogonal ? 1 : 0;
    out[8] = orth.odd_pairs;
    if (with_girth && girth_out) {
        girth_out[0] = compute_girth(g_ref.x_checks, g_ref.var_to_x);
        girth_out[1] = compute_girth(g_ref.z_checks, g_ref.var_to_z);
    }
    return 0;
}

// Flat check lists of the loaded code (for comparing code construction).
int ref_get_checks(int side, int *out_flat) {
    const auto &ch = side == 0 ? g_ref.x_checks : g_ref.z_checks;
    size_t k = 0;
    for (const auto &row : ch)
        for (int v : row) out_flat[k++] = v;
    return static_cast<int>(k);
}

uint64_t ref_trial_seed(uint64_t base, uint64_t idx) { return trial_seed_from_base(base, idx); }

void ref_sample(uint64_t base, long long idx, double p, int *err_out) {
    std::vector<int> err;
    sample_trial(base, idx, p, err);
    std::copy(err.begin(), err.end(), err_out);
}

void ref_syndrome(const int *err_in, int *sx_out, int *sz_out) {
    std::vector<int> err(err_in, err_in + g_ref.nvars);
    auto sx = compute_syndrome(g_ref.x_checks, err, true);
    auto sz = compute_syndrome(g_ref.z_checks, err, false);
    std::copy(sx.begin(), sx.end(), sx_out);
    std::copy(sz.begin(), sz.end(), sz_out);
}

// Decode one syndrome pair with PP off. Returns iterations; writes est and match flag.
int ref_decode(const int *sx_in, const int *sz_in, const double *prior4, int max_iter,
               double damping, int freeze_syn, int *est_out, int *syn_match) {
    std::vector<int> sx(sx_in, sx_in + g_ref.x_checks.size());
    std::vector<int> sz(sz_in, sz_in + g_ref.z_checks.size());
    Msg prior{prior4[0], prior4[1], prior4[2], prior4[3]};
    JointBPResult res = decode_plain(sx, sz, prior, max_iter, damping, freeze_syn != 0);
    std::copy(res.est.begin(), res.est.end(), est_out);
    *syn_match = res.syndrome_match ? 1 : 0;
    return res.iterations;
}

int ref_stab_equiv(const int *err_in, const int *est_in) {
    std::vector<int> err(err_in, err_in + g_ref.nvars), est(est_in, est_in + g_ref.nvars);
    return stab_equiv(err, est) ? 1 : 0;
}

void ref_prior(double p, double *out4) {
    Msg m = make_prior(p);
    for (int i = 0; i < 4; ++i) out4[i] = m[i];
}

// The simulate loop (jointbp_ets.cpp:5931-6207, PP off) over trial indices
// first_idx .. first_idx+count-1, spread over `threads` std::threads (the
// reference runs them one after another; trials are independent).
// Per trial: iterations, syndrome_match, success (= match && stabilizer
// residual), exact (est == err).  Returns wall seconds of the loop.
double ref_simulate(uint64_t base, long long first_idx, long long count, double p, int max_iter,
                    double damping, int threads, int *iters, unsigned char *syn_ok,
                    unsigned char *success, unsigned char *exact) {
    Msg prior = make_prior(p);
    std::atomic<long long> next{0};
    auto worker = [&]() {
        std::vector<int> err;
        for (;;) {
            long long t = next.fetch_add(1);
            if (t >= count) break;
            sample_trial(base, first_idx + t, p, err);
            auto sx = compute_syndrome(g_ref.x_checks, err, true);
            auto sz = compute_syndrome(g_ref.z_checks, err, false);
            JointBPResult res = decode_plain(sx, sz, prior, max_iter, damping, false);
            bool ok = res.syndrome_match;
            bool eq = ok && stab_equiv(err, res.est);
            iters[t] = res.iterations;
            syn_ok[t] = ok ? 1 : 0;
            success[t] = (ok && eq) ? 1 : 0;
            exact[t] = (res.est == err) ? 1 : 0;
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One decode with PP off that keeps the |LLR| of the last iteration run
// (abs_llr_x / abs_llr_z, jointbp_ets.cpp:2483-2493): the per-iteration
// diagnostics path (truth vector + iteration log) collects them at every
// iteration, so a decode that converges early keeps its last iteration's.
// llr_out: 2 x nvars (X then Z).  Returns iterations.
int ref_decode_llr(const int *sx_in, const int *sz_in, const double *prior4, int max_iter,
                   double damping, int *est_out, int *syn_match, double *llr_out) {
    std::vector<int> sx(sx_in, sx_in + g_ref.x_checks.size());
    std::vector<int> sz(sz_in, sz_in + g_ref.z_checks.size());
    Msg prior{prior4[0], prior4[1], prior4[2], prior4[3]};
    std::vector<int> truth(g_ref.nvars, 0);
    std::vector<std::string> logs;
    JointBPResult res = joint_bp_decode(g_ref.x_checks, g_ref.z_checks, g_ref.var_to_x, g_ref.var_to_z,
                                        sx, sz, prior, max_iter, 5, false, damping, false, false, true,
                                        nullptr, &truth, &g_ref.hx_basis, &g_ref.hz_basis, &logs, nullptr);
    std::copy(res.est.begin(), res.est.end(), est_out);
    *syn_match = res.syndrome_match ? 1 : 0;
    const size_t n = static_cast<size_t>(g_ref.nvars);
    for (size_t v = 0; v < n; ++v) {
        llr_out[v] = res.abs_llr_x.size() == n ? res.abs_llr_x[v] : -1.0;
        llr_out[n + v] = res.abs_llr_z.size() == n ? res.abs_llr_z[v] : -1.0;
    }
    return res.iterations;
}

// Loads one table group with read_ets_file (jointbp_ets.cpp:5238-5307 keeps a
// side when its file reads and has entries).  counts: entries per side, 0
// empty, -1 missing.
int ref_ets_add(int size, const char *label, const char *path_x, const char *path_z, int P,
                int *counts) {
    RefEtsGroup grp;
    grp.size = size;
    grp.label = label;
    const char *paths[2] = {path_x, path_z};
    for (int s = 0; s < 2; ++s) {
        counts[s] = -1;
        if (!paths[s]) continue;
        auto idx = std::make_unique<ETSIndex>();
        if (!read_ets_file(paths[s], P, size, s == 0 ? 'X' : 'Z', *idx)) continue;
        counts[s] = static_cast<int>(idx->entries.size());
        if (idx->entries.empty()) continue;
        (s == 0 ? grp.x : grp.z) = std::move(idx);
    }
    g_ets.push_back(std::move(grp));
    return 0;
}

void ref_ets_clear() { g_ets.clear(); }

// One repair with the reference's own routine: method 0 ETS (all loaded
// groups), 1 flip set, 2 OSD.  Returns 1 when est was changed by a repair
// that reported success (ETS: the group's size), else 0.
int ref_pp_repair(int method, int side, const int *syndrome, const double *abs_llr,
                  const int *vars, int n_vars, int *est_io) {
    const bool xs = side == 0;
    const auto &checks = xs ? g_ref.x_checks : g_ref.z_checks;
    std::vector<int> target(syndrome, syndrome + checks.size());
    std::vector<int> est(est_io, est_io + g_ref.nvars);
    int applied = 0;
    if (method == 0) {
        ETSFixResult r = attempt_ets_fix_side_multi(xs ? "X" : "Z", ets_refs(), checks,
                                                    xs ? g_ref.var_to_x : g_ref.var_to_z, target,
                                                    xs, est, false);
        applied = r.applied ? r.size : 0;
    } else if (method == 1) {
        std::vector<int> set(vars, vars + n_vars);
        applied = attempt_flip_pp_side(xs ? "X" : "Z", set, checks, target, xs, est, false,
                                       nullptr, "") ? 1 : 0;
    } else {
        std::vector<double> llr(abs_llr, abs_llr + g_ref.nvars);
        applied = attempt_osd_pp_side(xs ? "X" : "Z", llr, checks, target, xs, est, false,
                                      nullptr) ? 1 : 0;
    }
    std::copy(est.begin(), est.end(), est_io);
    return applied;
}

// The simulate loop with PP on (jointbp_ets.cpp:5931-6228), --damping as
// given: the decode with
// in-loop PP, then the post-decode ETS / flip / OSD passes and the
// reference's counting.  Per trial: iterations, final syndrome match,
// success, exact, flags (bit 0 BP match, 1 PP success, 2 ETS saved, 3 flip
// saved, 4 OSD saved, 5 ETS used, 6 ETS6, 7 ETS12) and, when est_out is not
// NULL, the final estimate.  counters[10]: pp_success, flip_pp_success,
// osd_pp_success, stab_success, ets_used, ets_saved, ets6_used, ets6_saved,
// ets12_used, ets12_saved.  Returns wall seconds.
double ref_simulate_pp(uint64_t base, long long first_idx, long long count, double p, int max_iter,
                       double damping, int flip_hist, int threads, int *iters, unsigned char *syn_ok,
                       unsigned char *success, unsigned char *exact, unsigned char *flags,
                       int *est_out, long long *counters) {
    Msg prior = make_prior(p);
    const std::vector<ETSIndexRef> refs = ets_refs();
    PPEarlyContext pp_ctx{true, false, refs, nullptr};
    std::atomic<long long> next{0};
    std::vector<std::array<long long, 10>> part(std::max(threads, 1));
    for (auto &a : part) a.fill(0);
    auto worker = [&](int wid) {
        auto &cnt = part[wid];
        std::vector<int> err;
        for (;;) {
            long long t = next.fetch_add(1);
            if (t >= count) break;
            sample_trial(base, first_idx + t, p, err);
            auto sx = compute_syndrome(g_ref.x_checks, err, true);
            auto sz = compute_syndrome(g_ref.z_checks, err, false);
            JointBPResult res = joint_bp_decode(
                g_ref.x_checks, g_ref.z_checks, g_ref.var_to_x, g_ref.var_to_z, sx, sz, prior,
                max_iter, flip_hist, false, damping, false, false, true, &pp_ctx, &err,
                &g_ref.hx_basis, &g_ref.hz_basis, nullptr, nullptr);
            bool bp_ok = res.bp_syndrome_match;
            bool used6 = res.pp_used_ets6;
            bool used12 = res.pp_used_ets12;
            bool used_any = res.pp_used_ets_any || used6 || used12;
            bool ok = res.syndrome_match;
            bool ets_pp_ok = res.pp_success_ets;
            bool flip_pp_ok = res.pp_success_flip;
            bool osd_pp_ok = false;
            if (res.pp_success && used_any) {
                cnt[4]++;
                if (used6) cnt[6]++;
                if (used12) cnt[8]++;
            }
            if (!ok) {
                ETSFixResult ets_x = attempt_ets_fix_side_multi("X", refs, g_ref.x_checks,
                                                                g_ref.var_to_x, sx, true, res.est, false);
                ETSFixResult ets_z = attempt_ets_fix_side_multi("Z", refs, g_ref.z_checks,
                                                                g_ref.var_to_z, sz, false, res.est, false);
                bool applied = ets_x.applied || ets_z.applied;
                used_any = used_any || applied;
                used6 = used6 || ets_x.size == 6 || ets_z.size == 6;
                used12 = used12 || ets_x.size == 12 || ets_z.size == 12;
                if (applied) cnt[4]++;
                if (used6) cnt[6]++;
                if (used12) cnt[8]++;
                if (applied) {
                    ok = compute_syndrome(g_ref.x_checks, res.est, true) == sx &&
                         compute_syndrome(g_ref.z_checks, res.est, false) == sz;
                    if (ok) ets_pp_ok = true;
                }
            }
            if (!ok) {
                bool fx = false, fz = false;
                for (const auto &set : build_flip_sets(res.flip_x_history, res.flip_x_window, res.flip_x)) {
                    if (!set.vars || set.vars->empty()) continue;
                    if (attempt_flip_pp_side("X", *set.vars, g_ref.x_checks, sx, true, res.est, false, nullptr, set.tag)) {
                        fx = true;
                        break;
                    }
                }
                for (const auto &set : build_flip_sets(res.flip_z_history, res.flip_z_window, res.flip_z)) {
                    if (!set.vars || set.vars->empty()) continue;
                    if (attempt_flip_pp_side("Z", *set.vars, g_ref.z_checks, sz, false, res.est, false, nullptr, set.tag)) {
                        fz = true;
                        break;
                    }
                }
                if (fx || fz) {
                    ok = compute_syndrome(g_ref.x_checks, res.est, true) == sx &&
                         compute_syndrome(g_ref.z_checks, res.est, false) == sz;
                    if (ok) flip_pp_ok = true;
                }
            }
            if (!ok) {
                auto shx = compute_syndrome(g_ref.x_checks, res.est, true);
                auto shz = compute_syndrome(g_ref.z_checks, res.est, false);
                int us = std::max(unsatisfied_count(shx, sx), unsatisfied_count(shz, sz));
                if (us <= 20) {
                    bool ox = attempt_osd_pp_side("X", res.abs_llr_x, g_ref.x_checks, sx, true, res.est, false, &err);
                    bool oz = attempt_osd_pp_side("Z", res.abs_llr_z, g_ref.z_checks, sz, false, res.est, false, &err);
                    if (ox || oz) {
                        ok = compute_syndrome(g_ref.x_checks, res.est, true) == sx &&
                             compute_syndrome(g_ref.z_checks, res.est, false) == sz;
                        if (ok) osd_pp_ok = true;
                    }
                }
            }
            bool ex = res.est == err;
            bool stab = ok && stab_equiv(err, res.est);
            bool succ = ok && stab;
            unsigned char f = bp_ok ? 1 : 0;
            if (succ && !ex && stab) cnt[3]++;
            if (succ && ets_pp_ok) {
                cnt[5]++;
                if (used6) cnt[7]++;
                if (used12) cnt[9]++;
                f |= 4;
            }
            if (succ && flip_pp_ok) {
                cnt[1]++;
                f |= 8;
            }
            bool osd_success = osd_pp_ok || (res.pp_success && !res.pp_success_ets && !res.pp_success_flip);
            if (succ && (res.pp_success || ets_pp_ok || flip_pp_ok || osd_pp_ok)) {
                cnt[0]++;
                f |= 2;
            }
            if (succ && osd_success) {
                cnt[2]++;
                f |= 16;
            }
            f |= (used_any ? 32 : 0) | (used6 ? 64 : 0) | (used12 ? 128 : 0);
            iters[t] = res.iterations;
            syn_ok[t] = ok ? 1 : 0;
            success[t] = succ ? 1 : 0;
            exact[t] = ex ? 1 : 0;
            flags[t] = f;
            if (est_out) std::copy(res.est.begin(), res.est.end(), est_out + t * g_ref.nvars);
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
        worker(0);
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < threads; ++i) pool.emplace_back(worker, i);
        for (auto &th : pool) th.join();
    }
    for (int k = 0; k < 10; ++k) {
        counters[k] = 0;
        for (const auto &a : part) counters[k] += a[k];
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
