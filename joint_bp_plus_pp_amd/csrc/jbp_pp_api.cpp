// Host-only entry points of post-processing (include/jbp.h): trapping-set
// table sets and single repairs on one estimate.  No device is touched.
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "../../include/jbp.h"
#include "jbp_pp.h"

namespace {

void say(char *err, size_t len, const char *fmt, ...) {
    if (!err || !len) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
}

bool desc_ok(const jbp_graph_desc *g) {
    return g && g->nvars > 0 && g->mX >= 0 && g->mZ >= 0 && g->x_check_offsets && g->x_edge_var &&
           g->z_check_offsets && g->z_edge_var;
}

jbp::pp::Graph graph_of(const jbp_graph_desc *g) {
    const int m[2] = {g->mX, g->mZ};
    const int *off[2] = {g->x_check_offsets, g->z_check_offsets};
    const int *var[2] = {g->x_edge_var, g->z_edge_var};
    return jbp::pp::graph_from_csr(g->nvars, m, off, var);
}

}  // namespace

extern "C" {

int jbp_pp_tables_create(jbp_pp_tables **out) {
    if (!out) return JBP_ERR_ARG;
    *out = new (std::nothrow) jbp_pp_tables();
    return *out ? JBP_OK : JBP_ERR_NOMEM;
}

void jbp_pp_tables_free(jbp_pp_tables *t) { delete t; }

int jbp_pp_tables_add(jbp_pp_tables *t, const jbp_graph_desc *graph, int64_t P, int size,
                      const char *label, const char *path_x, const char *path_z, int counts[2],
                      char *err, size_t errlen) {
    if (!t || !desc_ok(graph) || P <= 0) {
        say(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    if (size <= 0 || size > 32) {
        say(err, errlen, "trapping sets of %d variables are not supported (1..32)", size);
        return JBP_ERR_ARG;
    }
    jbp::pp::EtsGroup grp;
    grp.size = size;
    grp.label = label ? label : std::to_string(size);
    const char *paths[2] = {path_x, path_z};
    const int m[2] = {graph->mX, graph->mZ};
    int got[2] = {-1, -1};
    for (int s = 0; s < 2; ++s) {
        if (!paths[s] || !*paths[s]) continue;
        got[s] = jbp::pp::load_ets_file(paths[s], P, graph->nvars, m[s], size, s == 0 ? 'X' : 'Z',
                                        grp.side[s]);
        grp.has[s] = got[s] > 0;
    }
    if (counts) {
        counts[0] = got[0];
        counts[1] = got[1];
    }
    if (grp.has[0] || grp.has[1]) t->t.groups.push_back(std::move(grp));
    return JBP_OK;
}

int jbp_pp_repair(const jbp_graph_desc *graph, const jbp_pp_tables *t, int method, int side,
                  const int *syndrome, const double *abs_llr, const int *vars, int n_vars,
                  int *est, int *applied, char *err, size_t errlen) {
    if (!desc_ok(graph) || !syndrome || !est || (side != 0 && side != 1) || method < 0 ||
        method > 3 || (method == 0 && !t) || (method == 1 && n_vars > 0 && !vars) ||
        (method >= 2 && !abs_llr)) {
        say(err, errlen, "bad arguments");
        return JBP_ERR_ARG;
    }
    const jbp::pp::Graph g = graph_of(graph);
    const int n = graph->nvars, m = g.m[side];
    std::vector<uint8_t> e(n), syn(m);
    for (int v = 0; v < n; ++v) {
        if (est[v] < 0 || est[v] > 3) {
            say(err, errlen, "estimate state %d at qubit %d is not 0..3", est[v], v);
            return JBP_ERR_ARG;
        }
        e[v] = static_cast<uint8_t>(est[v]);
    }
    for (int c = 0; c < m; ++c) syn[c] = syndrome[c] ? 1 : 0;
    int done = 0;
    if (method == 0) {
        const jbp::pp::EtsFix f = jbp::pp::ets_fix(g, t->t, side, syn.data(), e);
        done = f.applied ? f.size : 0;
    } else if (method == 1) {
        std::vector<int> set(vars, vars + (n_vars > 0 ? n_vars : 0));
        done = jbp::pp::flip_fix(g, side, set, syn.data(), e) ? 1 : 0;
    } else {
        const jbp::pp::LlrSide l =
            method == 2 ? jbp::pp::LlrSide::values(abs_llr) : jbp::pp::LlrSide::marginals(abs_llr);
        done = jbp::pp::osd_fix(g, side, l, syn.data(), e) ? 1 : 0;
    }
    if (done)
        for (int v = 0; v < n; ++v) est[v] = e[v];
    if (applied) *applied = done;
    return JBP_OK;
}

}  // extern "C"
