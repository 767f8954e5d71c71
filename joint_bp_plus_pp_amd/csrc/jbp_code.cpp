// Host-side code construction and startup diagnostics (include/jbp_code.h).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/jbp_code.h"

namespace {

void set_err(char *err, size_t len, const char *fmt, ...) {
    if (!err || !len) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, len, fmt, ap);
    va_end(ap);
}

long long pmod(long long x, long long m) {
    long long r = x % m;
    return r < 0 ? r + m : r;
}

// Modular inverse by extended Euclid; -1 when gcd(a, m) != 1.
long long modular_inverse(long long a, long long m) {
    if (m == 1) return 0;
    long long r0 = m, r1 = pmod(a, m), s0 = 0, s1 = 1;
    while (r1 != 0) {
        long long q = r0 / r1;
        long long t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    return r0 == 1 ? pmod(s0, m) : -1;
}

}  // namespace

struct jbp_code {
    int n = 0;
    long long P = 0;
    int J = 0, L = 0;
    std::vector<long long> abcd;          // a | b | c | d, L/2 each
    int deg[2] = {0, 0};                  // stored entries per check
    std::vector<int> checks[2];           // m x deg, -1 = absent
    // CSR in the reference's build_cuda_graph order
    std::vector<int> chk_off[2], chk_edge[2], edge_var[2], edge_chk[2], edge_pos[2];
    std::vector<int> var_off[2], var_edge[2];
    // hypergraph-product codes: the classical matrix (rows of column indices)
    int hgp_m = 0, hgp_n = 0;
    std::vector<std::vector<int>> hgp_H;
};

namespace {

int mcount(const jbp_code &c, int s) {
    return c.deg[s] ? static_cast<int>(c.checks[s].size()) / c.deg[s] : 0;
}

void build_csr(jbp_code &c) {
    for (int s = 0; s < 2; ++s) {
        const int m = mcount(c, s), d = c.deg[s];
        c.chk_off[s].assign(m + 1, 0);
        c.chk_edge[s].clear();
        c.edge_var[s].clear();
        c.edge_chk[s].clear();
        c.edge_pos[s].clear();
        std::vector<int> index(static_cast<size_t>(m) * d, -1);
        for (int ch = 0; ch < m; ++ch) {
            c.chk_off[s][ch] = static_cast<int>(c.edge_var[s].size());
            for (int i = 0; i < d; ++i) {
                int v = c.checks[s][static_cast<size_t>(ch) * d + i];
                if (v < 0 || v >= c.n) continue;
                index[static_cast<size_t>(ch) * d + i] = static_cast<int>(c.edge_var[s].size());
                c.chk_edge[s].push_back(static_cast<int>(c.edge_var[s].size()));
                c.edge_var[s].push_back(v);
                c.edge_chk[s].push_back(ch);
                c.edge_pos[s].push_back(i);
            }
        }
        c.chk_off[s][m] = static_cast<int>(c.edge_var[s].size());
        // variable adjacency in ascending (check, position) order
        std::vector<int> cnt(c.n + 1, 0);
        for (int v : c.edge_var[s]) cnt[v + 1]++;
        for (int v = 0; v < c.n; ++v) cnt[v + 1] += cnt[v];
        c.var_off[s] = cnt;
        c.var_edge[s].assign(c.edge_var[s].size(), 0);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t e = 0; e < c.edge_var[s].size(); ++e) c.var_edge[s][fill[c.edge_var[s][e]]++] = static_cast<int>(e);
    }
}

// GF(2) rank by incremental elimination on the highest set bit.
long long gf2_rank(const jbp_code &c, int s) {
    const int W = (c.n + 63) / 64;
    std::vector<std::vector<uint64_t>> basis;
    std::vector<int> pivot_row(c.n, -1);
    for (int ch = 0; ch < mcount(c, s); ++ch) {
        std::vector<uint64_t> row(W, 0);
        for (int e = c.chk_off[s][ch]; e < c.chk_off[s][ch + 1]; ++e) {
            int v = c.edge_var[s][e];
            row[v >> 6] ^= 1ull << (v & 63);
        }
        for (;;) {
            int piv = -1;
            for (int w = W - 1; w >= 0; --w)
                if (row[w]) { piv = w * 64 + 63 - __builtin_clzll(row[w]); break; }
            if (piv < 0) break;
            if (pivot_row[piv] < 0) {
                pivot_row[piv] = static_cast<int>(basis.size());
                basis.push_back(std::move(row));
                break;
            }
            const auto &b = basis[pivot_row[piv]];
            for (int w = 0; w < W; ++w) row[w] ^= b[w];
        }
    }
    return static_cast<long long>(basis.size());
}

// Reduced row echelon form over GF(2) of a dense matrix with `cols` columns
// (rows as 64-bit words); returns the pivot column of each nonzero row and
// leaves the reduced rows in `rows` (first rank entries).
std::vector<int> gf2_rref(std::vector<std::vector<uint64_t>> &rows, int cols) {
    std::vector<int> piv;
    int rank = 0;
    for (int col = 0; col < cols && rank < static_cast<int>(rows.size()); ++col) {
        const int w = col >> 6;
        const uint64_t bit = 1ull << (col & 63);
        int sel = -1;
        for (int r = rank; r < static_cast<int>(rows.size()); ++r)
            if (rows[r][w] & bit) { sel = r; break; }
        if (sel < 0) continue;
        std::swap(rows[rank], rows[sel]);
        for (int r = 0; r < static_cast<int>(rows.size()); ++r)
            if (r != rank && (rows[r][w] & bit))
                for (size_t k = 0; k < rows[r].size(); ++k) rows[r][k] ^= rows[rank][k];
        piv.push_back(col);
        ++rank;
    }
    rows.resize(rank);
    return piv;
}

// Null-space basis of a matrix (rows of column indices, `cols` columns) and
// the columns that are not pivots of its reduced row echelon form: unit
// vectors at those columns complete the row space to the whole space.
void gf2_kernel(const std::vector<std::vector<int>> &M, int cols,
                std::vector<std::vector<int>> &ker_support, std::vector<int> &free_cols) {
    const int W = (cols + 63) / 64;
    std::vector<std::vector<uint64_t>> rows(M.size(), std::vector<uint64_t>(W, 0));
    for (size_t r = 0; r < M.size(); ++r)
        for (int c : M[r]) rows[r][c >> 6] ^= 1ull << (c & 63);
    const std::vector<int> piv = gf2_rref(rows, cols);
    std::vector<char> is_piv(cols, 0);
    for (int c : piv) is_piv[c] = 1;
    free_cols.clear();
    ker_support.clear();
    for (int f = 0; f < cols; ++f) {
        if (is_piv[f]) continue;
        free_cols.push_back(f);
        // x[f] = 1, x[pivot_i] = rref[i][f], other free columns 0
        std::vector<int> sup;
        for (size_t i = 0; i < piv.size(); ++i)
            if ((rows[i][f >> 6] >> (f & 63)) & 1ull) sup.push_back(piv[i]);
        sup.push_back(f);
        std::sort(sup.begin(), sup.end());
        ker_support.push_back(std::move(sup));
    }
}

// Shortest cycle of the Tanner graph by breadth-first search from every
// variable node (a cycle through the root closes when two BFS branches meet).
int girth(const jbp_code &c, int s) {
    const int n = c.n, m = mcount(c, s), total = n + m;
    if (n == 0 || m == 0) return 0;
    int best = std::numeric_limits<int>::max();
    std::vector<int> dist(total, -1), par(total, -1), q(total);
    std::vector<int> touched;
    for (int root = 0; root < n && best > 4; ++root) {
        int qh = 0, qt = 0;
        q[qt++] = root;
        dist[root] = 0;
        par[root] = -1;
        touched.assign(1, root);
        while (qh < qt) {
            int u = q[qh++];
            if (2 * dist[u] + 1 >= best) break;
            auto visit = [&](int w) {
                if (dist[w] < 0) {
                    dist[w] = dist[u] + 1;
                    par[w] = u;
                    q[qt++] = w;
                    touched.push_back(w);
                } else if (par[u] != w) {
                    best = std::min(best, dist[u] + dist[w] + 1);
                }
            };
            if (u < n) {
                for (int k = c.var_off[s][u]; k < c.var_off[s][u + 1]; ++k)
                    visit(n + c.edge_chk[s][c.var_edge[s][k]]);
            } else {
                int ch = u - n;
                for (int e = c.chk_off[s][ch]; e < c.chk_off[s][ch + 1]; ++e) visit(c.edge_var[s][e]);
            }
        }
        for (int w : touched) dist[w] = -1;
    }
    return best == std::numeric_limits<int>::max() ? 0 : best;
}

// Number of (X check, Z check) pairs with odd overlap.
long long odd_pairs(const jbp_code &c) {
    const int mz = mcount(c, 1);
    std::vector<int> cnt(mz, 0), touched;
    long long odd = 0;
    for (int xc = 0; xc < mcount(c, 0); ++xc) {
        touched.clear();
        for (int e = c.chk_off[0][xc]; e < c.chk_off[0][xc + 1]; ++e) {
            int v = c.edge_var[0][e];
            for (int k = c.var_off[1][v]; k < c.var_off[1][v + 1]; ++k) {
                int zc = c.edge_chk[1][c.var_edge[1][k]];
                if (cnt[zc]++ == 0) touched.push_back(zc);
            }
        }
        for (int zc : touched) {
            odd += cnt[zc] & 1;
            cnt[zc] = 0;
        }
    }
    return odd;
}

// Seeded generator stream for the classical code (splitmix64 steps).
struct Stream {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t k) { return next() % k; }
};

// Random (col_w, row_w)-regular m x n matrix without repeated entries:
// socket matching with a seeded shuffle, redrawn until simple.
bool random_regular(int m, int n, int col_w, int row_w, uint64_t seed,
                    std::vector<std::vector<int>> &rows) {
    if (static_cast<long long>(m) * row_w != static_cast<long long>(n) * col_w) return false;
    Stream st{seed};
    std::vector<int> sockets(static_cast<size_t>(n) * col_w);
    for (int attempt = 0; attempt < 10000; ++attempt) {
        for (size_t i = 0; i < sockets.size(); ++i) sockets[i] = static_cast<int>(i / col_w);
        for (size_t i = sockets.size(); i > 1; --i) std::swap(sockets[i - 1], sockets[st.below(i)]);
        rows.assign(m, {});
        bool simple = true;
        for (int r = 0; r < m && simple; ++r) {
            for (int k = 0; k < row_w; ++k) {
                int col = sockets[static_cast<size_t>(r) * row_w + k];
                if (std::find(rows[r].begin(), rows[r].end(), col) != rows[r].end()) {
                    simple = false;
                    break;
                }
                rows[r].push_back(col);
            }
            std::sort(rows[r].begin(), rows[r].end());
        }
        if (simple) return true;
    }
    return false;
}

}  // namespace

extern "C" {

int jbp_code_from_params(const char *path, jbp_code **out, char *err, size_t errlen) {
    if (!out || !path) return JBP_ERR_ARG;
    *out = nullptr;
    std::ifstream in(path);
    if (!in) {
        set_err(err, errlen, "Failed to read params from %s", path);
        return JBP_ERR_ARG;
    }
    long long P = 0;
    int J = 0, L = 0;
    unsigned long long file_seed = 0;
    if (!(in >> P >> J >> L) || !(in >> file_seed) || L % 2 != 0 || P <= 0 || J <= 0 || L <= 0) {
        set_err(err, errlen, "Failed to read params from %s", path);
        return JBP_ERR_ARG;
    }
    const int L2 = L / 2;
    std::vector<long long> abcd(4 * static_cast<size_t>(L2));
    for (auto &v : abcd)
        if (!(in >> v)) {
            set_err(err, errlen, "Failed to read params from %s", path);
            return JBP_ERR_ARG;
        }
    auto *c = new jbp_code();
    c->P = P;
    c->J = J;
    c->L = L;
    c->n = L * static_cast<int>(P);
    c->abcd = abcd;
    const long long *a = &abcd[0], *b = &abcd[L2], *cc_ = &abcd[2 * L2], *d = &abcd[3 * L2];
    for (int s = 0; s < 2; ++s) {
        c->deg[s] = L;
        c->checks[s].assign(static_cast<size_t>(J) * P * L, -1);
    }
    for (int r = 0; r < J; ++r) {
        for (int col = 0; col < L; ++col) {
            const bool left = col < L2;
            const int k = static_cast<int>(pmod((left ? col : col - L2) - r, L2));
            // X block: (a, b) on the left half, (c, d) on the right, index (col - r) mod L2.
            long long xa = pmod(left ? a[k] : cc_[k], P), xb = pmod(left ? b[k] : d[k], P);
            // Z block: inverse map of the other half's pair at index -(col - r) mod L2.
            const int kz = static_cast<int>(pmod(-k, L2));
            long long inv = modular_inverse(pmod(left ? cc_[kz] : a[kz], P), P);
            long long za = inv, zb = inv >= 0 ? pmod(-inv * pmod(left ? d[kz] : b[kz], P), P) : -1;
            for (long long x = 0; x < P; ++x) {
                size_t slot = (static_cast<size_t>(r) * P + x) * L + col;
                c->checks[0][slot] = col * static_cast<int>(P) + static_cast<int>(pmod(xa * x + xb, P));
                if (inv >= 0)
                    c->checks[1][slot] = col * static_cast<int>(P) + static_cast<int>(pmod(za * x + zb, P));
            }
        }
    }
    build_csr(*c);
    *out = c;
    return JBP_OK;
}

int jbp_code_from_checks(int nvars, int mx, int xdeg, const int *xflat, int mz, int zdeg,
                         const int *zflat, jbp_code **out, char *err, size_t errlen) {
    if (!out || nvars <= 0 || mx < 0 || mz < 0 || xdeg < 0 || zdeg < 0 ||
        (mx && xdeg && !xflat) || (mz && zdeg && !zflat)) {
        set_err(err, errlen, "bad check lists");
        return JBP_ERR_ARG;
    }
    auto *c = new jbp_code();
    c->n = nvars;
    c->deg[0] = xdeg;
    c->deg[1] = zdeg;
    c->checks[0].assign(xflat, xflat + static_cast<size_t>(mx) * xdeg);
    c->checks[1].assign(zflat, zflat + static_cast<size_t>(mz) * zdeg);
    for (int s = 0; s < 2; ++s)
        for (int v : c->checks[s])
            if (v >= nvars) {
                delete c;
                set_err(err, errlen, "check entry out of range");
                return JBP_ERR_ARG;
            }
    build_csr(*c);
    *out = c;
    return JBP_OK;
}

int jbp_code_hgp(int m, int n, int col_w, int row_w, uint64_t seed, jbp_code **out, char *err,
                 size_t errlen) {
    if (!out || m <= 0 || n <= 0 || col_w <= 0 || row_w <= 0) {
        set_err(err, errlen, "bad HGP arguments");
        return JBP_ERR_ARG;
    }
    std::vector<std::vector<int>> H;
    if (!random_regular(m, n, col_w, row_w, seed, H)) {
        set_err(err, errlen, "no simple (%d,%d)-regular %dx%d matrix found", col_w, row_w, m, n);
        return JBP_ERR_ARG;
    }
    std::vector<std::vector<int>> Ht(n);  // rows of H^T: checks meeting each column
    for (int r = 0; r < m; ++r)
        for (int col : H[r]) Ht[col].push_back(r);
    const int nq = n * n + m * m;
    const int degx = row_w + col_w, degz = col_w + row_w;
    std::vector<int> xf, zf;
    xf.reserve(static_cast<size_t>(m) * n * degx);
    zf.reserve(static_cast<size_t>(n) * m * degz);
    // X check (a, j): qubits (i, j) for i in row a of H, then (a, b) for b in column j of H.
    for (int a = 0; a < m; ++a)
        for (int j = 0; j < n; ++j) {
            for (int i : H[a]) xf.push_back(i * n + j);
            for (int b : Ht[j]) xf.push_back(n * n + a * m + b);
        }
    // Z check (i, b): qubits (i, j) for j in row b of H, then (a, b) for a in column i of H.
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < m; ++b) {
            for (int j : H[b]) zf.push_back(i * n + j);
            for (int a : Ht[i]) zf.push_back(n * n + a * m + b);
        }
    int rc = jbp_code_from_checks(nq, m * n, degx, xf.data(), n * m, degz, zf.data(), out, err,
                                  errlen);
    if (rc == JBP_OK) {
        (*out)->hgp_m = m;
        (*out)->hgp_n = n;
        (*out)->hgp_H = std::move(H);
    }
    return rc;
}

int jbp_code_logicals(const jbp_code *c, int side, int64_t *n_rows, int64_t *n_entries,
                      int64_t *row_off, int *qubits) {
    if (!c || (side != 0 && side != 1) || !n_rows || !n_entries) return JBP_ERR_ARG;
    if (c->hgp_H.empty()) return JBP_ERR_UNSUPPORTED;
    const int m = c->hgp_m, n = c->hgp_n;
    const long long nn = static_cast<long long>(n) * n;
    std::vector<std::vector<int>> Ht(n);
    for (int r = 0; r < m; ++r)
        for (int col : c->hgp_H[r]) Ht[col].push_back(r);
    std::vector<std::vector<int>> kerH, kerHt;  // supports over [n] and [m]
    std::vector<int> freeH, freeHt;             // complements of row(H), col(H)
    gf2_kernel(c->hgp_H, n, kerH, freeH);
    gf2_kernel(Ht, m, kerHt, freeHt);
    // Qubit (i, j) of the n x n block is i*n + j; (a, b) of the m x m block
    // is n*n + a*m + b.  Side 0 (parities with the X part of a residual):
    // e_p (x) z and y (x) e_q; side 1: z (x) e_p and e_q (x) y.
    int64_t rows = 0, ent = 0;
    auto emit = [&](const std::vector<int> &sup, auto qubit) {
        if (row_off) row_off[rows] = ent;
        for (int t : sup) {
            if (qubits) qubits[ent] = static_cast<int>(qubit(t));
            ++ent;
        }
        ++rows;
    };
    for (int p : freeH)
        for (const auto &z : kerH)
            emit(z, [&](int t) { return side == 0 ? 1LL * p * n + t : 1LL * t * n + p; });
    for (int q : freeHt)
        for (const auto &y : kerHt)
            emit(y, [&](int t) { return side == 0 ? nn + 1LL * t * m + q : nn + 1LL * q * m + t; });
    if (row_off) row_off[rows] = ent;
    *n_rows = rows;
    *n_entries = ent;
    return JBP_OK;
}

void jbp_code_free(jbp_code *code) { delete code; }

int jbp_code_dims(const jbp_code *c, int64_t out[8]) {
    if (!c || !out) return JBP_ERR_ARG;
    out[0] = c->n;
    out[1] = mcount(*c, 0);
    out[2] = mcount(*c, 1);
    out[3] = static_cast<int64_t>(c->edge_var[0].size());
    out[4] = static_cast<int64_t>(c->edge_var[1].size());
    out[5] = c->P;
    out[6] = c->J;
    out[7] = c->L;
    return JBP_OK;
}

int jbp_code_graph(const jbp_code *c, jbp_graph_desc *g) {
    if (!c || !g) return JBP_ERR_ARG;
    std::memset(g, 0, sizeof(*g));
    g->nvars = c->n;
    g->mX = mcount(*c, 0);
    g->mZ = mcount(*c, 1);
    g->x_edges = static_cast<int>(c->edge_var[0].size());
    g->z_edges = static_cast<int>(c->edge_var[1].size());
    g->x_check_offsets = c->chk_off[0].data();
    g->x_check_edges = c->chk_edge[0].data();
    g->x_edge_var = c->edge_var[0].data();
    g->x_edge_check = c->edge_chk[0].data();
    g->x_edge_pos = c->edge_pos[0].data();
    g->x_var_offsets = c->var_off[0].data();
    g->x_var_edges = c->var_edge[0].data();
    g->z_check_offsets = c->chk_off[1].data();
    g->z_check_edges = c->chk_edge[1].data();
    g->z_edge_var = c->edge_var[1].data();
    g->z_edge_check = c->edge_chk[1].data();
    g->z_edge_pos = c->edge_pos[1].data();
    g->z_var_offsets = c->var_off[1].data();
    g->z_var_edges = c->var_edge[1].data();
    return JBP_OK;
}

int jbp_code_params(const jbp_code *c, int64_t *out) {
    if (!c || !out || c->abcd.empty()) return JBP_ERR_ARG;
    for (size_t i = 0; i < c->abcd.size(); ++i) out[i] = c->abcd[i];
    return JBP_OK;
}

int jbp_code_checks(const jbp_code *c, int side, int *flat) {
    if (!c || (side != 0 && side != 1)) return -1;
    if (flat) std::copy(c->checks[side].begin(), c->checks[side].end(), flat);
    return c->deg[side];
}

int jbp_code_diagnostics(const jbp_code *c, int with_girth, int64_t out[15]) {
    if (!c || !out) return JBP_ERR_ARG;
    out[0] = gf2_rank(*c, 0);
    out[1] = gf2_rank(*c, 1);
    for (int s = 0; s < 2; ++s) {
        int vmin = std::numeric_limits<int>::max(), vmax = 0, cmin = std::numeric_limits<int>::max(), cmax = 0;
        for (int v = 0; v < c->n; ++v) {
            int d = c->var_off[s][v + 1] - c->var_off[s][v];
            vmin = std::min(vmin, d);
            vmax = std::max(vmax, d);
        }
        for (int ch = 0; ch < mcount(*c, s); ++ch) {
            int d = c->chk_off[s][ch + 1] - c->chk_off[s][ch];
            cmin = std::min(cmin, d);
            cmax = std::max(cmax, d);
        }
        if (mcount(*c, s) == 0) cmin = 0;
        out[2 + 4 * s] = vmin;
        out[3 + 4 * s] = vmax;
        out[4 + 4 * s] = cmin;
        out[5 + 4 * s] = cmax;
    }
    out[10] = odd_pairs(*c);
    out[11] = with_girth ? girth(*c, 0) : 0;
    out[12] = with_girth ? girth(*c, 1) : 0;
    out[13] = 0;
    out[14] = 0;
    return JBP_OK;
}

}  // extern "C"
