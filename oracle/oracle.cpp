// ORACLE — test infrastructure, not product code.
//
// A CPU restatement of the reference's `--simulate --no-pp` path, used only
// by tests/, by the smoke check and by bench.py's CPU-baseline leg, always as
// the checker and never as the thing measured or shipped.  The product
// (joint_bp_plus_pp_amd/) neither links nor loads this file's library.
//
// Each function cites the reference lines (jointbp_ets.cpp) it restates.
// Arithmetic is kept in the reference's order (double precision, left-to-right
// sums, division-based normalisation, no FMA: built with -ffp-contract=off)
// so that per-trial estimates and iteration counts are identical.
//
// Pins (tests/test_oracle.py):
//   code construction  -> check lists equal to the reference's constructors,
//                         README known answers (n, m, rank, k, edges, girth)
//   seeds / sampler    -> closed form of splitmix64, the C++ standard's
//                         mt19937_64 known answer, golden error vectors
//   syndromes          -> brute force, golden vectors
//   BP decode          -> golden per-trial (iterations, est hash, success)
//                         written by the reference decoder; closed-form
//                         invariants on a single-check toy code
//   logical check      -> brute-force span membership on a tiny code,
//                         golden success flags
//   learned-weight BP  -> reduces to plain BP at unit weights (exact);
//                         gradient vs central finite differences
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <random>
#include <thread>
#include <vector>

namespace {

using Msg = std::array<double, 4>;

// ---------------------------------------------------------------- code ----

// jointbp_ets.cpp:62-66
long long mod_norm(long long x, long long m) {
    long long r = x % m;
    return r < 0 ? r + m : r;
}

// Modular inverse by the extended Euclidean algorithm; -1 when none exists
// (jointbp_ets.cpp:68-88 returns nullopt in that case).
long long mod_inverse(long long a, long long m) {
    a = mod_norm(a, m);
    if (m == 1) return 0;
    long long old_r = a, r = m, old_s = 1, s = 0;
    while (r != 0) {
        long long q = old_r / r;
        long long t = old_r - q * r; old_r = r; r = t;
        t = old_s - q * s; old_s = s; s = t;
    }
    if (old_r != 1) return -1;
    return mod_norm(old_s, m);
}

struct Code {
    int n = 0;
    int m[2] = {0, 0};                 // [0] = X checks, [1] = Z checks
    // per side: check-major edges (reference edge order, jointbp_ets.cpp:949-964)
    std::vector<int> chk_off[2];       // m+1
    std::vector<int> edge_var[2];      // E
    // per side: var -> edge ids in ascending check order (jointbp_ets.cpp:922-935)
    std::vector<int> var_off[2];       // n+1
    std::vector<int> var_edge[2];
    // GF(2) row basis per side (jointbp_ets.cpp:827-885)
    int words = 0;
    std::vector<std::vector<uint64_t>> basis_rows[2];
    std::vector<int> basis_pivot[2];
    std::vector<int> basis_order[2];
};

// Build from per-side check lists; -1 entries are absent blocks and are
// skipped, as build_var_adjacency and compute_syndrome skip them
// (jointbp_ets.cpp:928-931; the syndrome indexes err[v] only for the
// APM codes, where every entry is present).
void finish_code(Code &code, const std::vector<std::vector<int>> checks[2]) {
    for (int s = 0; s < 2; ++s) {
        code.m[s] = static_cast<int>(checks[s].size());
        code.chk_off[s].assign(code.m[s] + 1, 0);
        code.edge_var[s].clear();
        for (int c = 0; c < code.m[s]; ++c) {
            code.chk_off[s][c] = static_cast<int>(code.edge_var[s].size());
            for (int v : checks[s][c]) {
                if (v < 0 || v >= code.n) continue;
                code.edge_var[s].push_back(v);
            }
        }
        code.chk_off[s][code.m[s]] = static_cast<int>(code.edge_var[s].size());
        std::vector<int> deg(code.n, 0);
        for (int v : code.edge_var[s]) deg[v]++;
        code.var_off[s].assign(code.n + 1, 0);
        for (int v = 0; v < code.n; ++v) code.var_off[s][v + 1] = code.var_off[s][v] + deg[v];
        code.var_edge[s].assign(code.edge_var[s].size(), 0);
        std::vector<int> fill(code.var_off[s].begin(), code.var_off[s].end() - 1);
        for (int e = 0; e < static_cast<int>(code.edge_var[s].size()); ++e) {
            code.var_edge[s][fill[code.edge_var[s][e]]++] = e;  // ascending check order
        }
    }
    // build_row_basis (jointbp_ets.cpp:845-885): incremental elimination on
    // the highest set bit.
    code.words = (code.n + 63) / 64;
    for (int s = 0; s < 2; ++s) {
        code.basis_rows[s].clear();
        code.basis_pivot[s].clear();
        std::vector<int> pivot_to_row(code.n, -1);
        for (int c = 0; c < code.m[s]; ++c) {
            std::vector<uint64_t> row(code.words, 0);
            for (int e = code.chk_off[s][c]; e < code.chk_off[s][c + 1]; ++e) {
                int v = code.edge_var[s][e];
                row[v >> 6] |= 1ULL << (v & 63);
            }
            for (;;) {
                int pivot = -1;
                for (int w = code.words - 1; w >= 0; --w) {
                    if (row[w]) { pivot = w * 64 + 63 - __builtin_clzll(row[w]); break; }
                }
                if (pivot < 0) break;
                int ex = pivot_to_row[pivot];
                if (ex < 0) {
                    pivot_to_row[pivot] = static_cast<int>(code.basis_rows[s].size());
                    code.basis_rows[s].push_back(row);
                    code.basis_pivot[s].push_back(pivot);
                    break;
                }
                for (int w = 0; w < code.words; ++w) row[w] ^= code.basis_rows[s][ex][w];
            }
        }
        code.basis_order[s].resize(code.basis_rows[s].size());
        for (size_t i = 0; i < code.basis_order[s].size(); ++i) code.basis_order[s][i] = static_cast<int>(i);
        std::sort(code.basis_order[s].begin(), code.basis_order[s].end(),
                  [&](int a, int b) { return code.basis_pivot[s][a] > code.basis_pivot[s][b]; });
    }
}

// read_params + build_block_matrix_hx/hz + build_check_neighbors
// (jointbp_ets.cpp:98-180, 628-648).
bool code_from_params(const char *path, Code &code) {
    std::ifstream in(path);
    if (!in) return false;
    long long P = 0;
    int J = 0, L = 0;
    unsigned long long unused_seed = 0;
    if (!(in >> P >> J >> L)) return false;
    if (!(in >> unused_seed)) return false;
    if (L % 2 != 0) return false;
    int L2 = L / 2;
    std::vector<long long> a(L2), b(L2), c(L2), d(L2);
    for (auto *arr : {&a, &b, &c, &d})
        for (int i = 0; i < L2; ++i)
            if (!(in >> (*arr)[i])) return false;
    // blocks[side][r][cc] = (alpha, beta) or (-1, -1) when absent
    std::vector<std::array<long long, 2>> blk[2];
    for (int s = 0; s < 2; ++s) blk[s].assign(static_cast<size_t>(J) * L, {-1, -1});
    for (int r = 0; r < J; ++r) {
        for (int cc = 0; cc < L; ++cc) {
            bool left = cc < L2;
            int t = left ? cc : cc - L2;
            int k = ((t - r) % L2 + L2) % L2;
            // Hx: left half uses (a, b), right half (c, d), index (cc - r) mod L2.
            if (left) blk[0][r * L + cc] = {mod_norm(a[k], P), mod_norm(b[k], P)};
            else blk[0][r * L + cc] = {mod_norm(c[k], P), mod_norm(d[k], P)};
            // Hz: inverse maps with the halves swapped, index -(cc - r) mod L2.
            int idx = ((-k) % L2 + L2) % L2;
            long long num = left ? c[idx] : a[idx];
            long long off = left ? d[idx] : b[idx];
            long long inv = mod_inverse(mod_norm(num, P), P);
            if (inv >= 0) blk[1][r * L + cc] = {inv, mod_norm(-inv * mod_norm(off, P), P)};
        }
    }
    code.n = L * static_cast<int>(P);
    std::vector<std::vector<int>> checks[2];
    for (int s = 0; s < 2; ++s) {
        checks[s].assign(static_cast<size_t>(J) * P, std::vector<int>(L, -1));
        for (int r = 0; r < J; ++r)
            for (long long x = 0; x < P; ++x)
                for (int cc = 0; cc < L; ++cc) {
                    const auto &bb = blk[s][r * L + cc];
                    if (bb[0] < 0) continue;
                    checks[s][r * P + x][cc] = cc * static_cast<int>(P) +
                                               static_cast<int>(mod_norm(bb[0] * x + bb[1], P));
                }
    }
    finish_code(code, checks);
    return true;
}

// ------------------------------------------------------------- helpers ----

// jointbp_ets.cpp:650-657 (left-to-right sum, divide each entry)
inline void normalize(Msg &m) {
    double sum = m[0] + m[1] + m[2] + m[3];
    if (sum <= 0.0) { m = {0.25, 0.25, 0.25, 0.25}; return; }
    for (double &v : m) v /= sum;
}
inline int xbit(int s) { return (s == 1 || s == 3) ? 1 : 0; }   // :659-661
inline int zbit(int s) { return (s == 2 || s == 3) ? 1 : 0; }   // :663-665
inline Msg mul(const Msg &a, const Msg &b) {                     // :1851-1853
    return Msg{a[0] * b[0], a[1] * b[1], a[2] * b[2], a[3] * b[3]};
}
// jointbp_ets.cpp:1903-1911
inline Msg scaled_mix(const Msg &nw, const Msg &old, double damping) {
    if (damping <= 0.0) return nw;
    Msg out;
    for (int i = 0; i < 4; ++i) out[i] = (1.0 - damping) * nw[i] + damping * old[i];
    normalize(out);
    return out;
}

// jointbp_ets.cpp:4401-4411
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
inline uint64_t trial_seed(uint64_t base, uint64_t idx) {
    return splitmix64(base ^ (idx * 0x9e3779b97f4a7c15ull));
}

// jointbp_ets.cpp:5932-5946: one U[0,1) draw per qubit in qubit order from
// std::mt19937_64 through std::uniform_real_distribution<double>.
void sample(const Code &code, uint64_t base, uint64_t idx, double p, int *err) {
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::mt19937_64 rng(trial_seed(base, idx));
    for (int i = 0; i < code.n; ++i) {
        double u = dist(rng);
        if (u < (1.0 - p)) {
            err[i] = 0;
        } else {
            double v = (u - (1.0 - p)) / p;
            err[i] = (v < 1.0 / 3.0) ? 1 : (v < 2.0 / 3.0) ? 2 : 3;
        }
    }
}

// jointbp_ets.cpp:1834-1849
void syndrome(const Code &code, int side, const int *err, int *out) {
    for (int c = 0; c < code.m[side]; ++c) {
        int par = 0;
        for (int e = code.chk_off[side][c]; e < code.chk_off[side][c + 1]; ++e) {
            int s = err[code.edge_var[side][e]];
            par ^= side == 0 ? xbit(s) : zbit(s);
        }
        out[c] = par;
    }
}

// fill_diff_bits + in_row_space (jointbp_ets.cpp:887-920); side 0 tests
// membership in the row space of Hx, side 1 in that of Hz.
bool in_row_space(const Code &code, int basis_side, const int *truth, const int *est, bool use_xbit) {
    std::vector<uint64_t> vec(code.words, 0);
    for (int i = 0; i < code.n; ++i) {
        int t = use_xbit ? xbit(truth[i]) : zbit(truth[i]);
        int e = use_xbit ? xbit(est[i]) : zbit(est[i]);
        if (t ^ e) vec[i >> 6] |= 1ULL << (i & 63);
    }
    for (int idx : code.basis_order[basis_side]) {
        int pivot = code.basis_pivot[basis_side][idx];
        if (!(vec[pivot >> 6] & (1ULL << (pivot & 63)))) continue;
        const auto &row = code.basis_rows[basis_side][idx];
        for (int w = 0; w < code.words; ++w) vec[w] ^= row[w];
    }
    for (uint64_t w : vec) if (w) return false;
    return true;
}

// jointbp_ets.cpp:6199-6207: residual X part must lie in rowspace(Hz) and
// residual Z part in rowspace(Hx).
bool stab_equiv(const Code &code, const int *err, const int *est) {
    return in_row_space(code, 1, err, est, true) && in_row_space(code, 0, err, est, false);
}

// ------------------------------------------------------------- decoder ----

// Optional learned weights (no reference counterpart): multiply the check
// output LLR and the extrinsic variable LLR of block `blk` at iteration `t`.
// Plain BP never uses this path.
struct Workspace {
    std::vector<Msg> v2c[2], c2v[2];
    std::vector<int> est, shat;
};

// joint_bp_decode with pp_ctx == nullptr (jointbp_ets.cpp:1943-3079):
// init 1970-1982; X checks 2373-2410; Z checks 2412-2449; joint variable
// update + hard decision 2451-2519; syndrome test 2521-2525, 2733-2735;
// freeze-syn 1884-1901, 2526-2535; unconverged return 3078.
int decode(const Code &code, const int *sx, const int *sz, const Msg &prior, int max_iter,
           double damping, bool freeze_syn, Workspace &ws, int *est_out, int *syn_match) {
    const int *target[2] = {sx, sz};
    for (int s = 0; s < 2; ++s) {
        size_t E = code.edge_var[s].size();
        ws.v2c[s].assign(E, prior);
        ws.c2v[s].assign(E, Msg{0.25, 0.25, 0.25, 0.25});
    }
    ws.est.assign(code.n, 0);
    bool frozen[2] = {false, false};
    std::vector<double> pe, po, se, so, q0, q1;
    for (int iter = 0; iter < max_iter; ++iter) {
        for (int s = 0; s < 2; ++s) {
            if (frozen[s]) continue;
            for (int c = 0; c < code.m[s]; ++c) {
                int off = code.chk_off[s][c];
                int deg = code.chk_off[s][c + 1] - off;
                q0.assign(deg, 0.0); q1.assign(deg, 0.0);
                for (int i = 0; i < deg; ++i) {
                    const Msg &m = ws.v2c[s][off + i];
                    // X side pairs (I,Z) vs (X,Y); Z side pairs (I,X) vs (Z,Y)
                    q0[i] = s == 0 ? m[0] + m[2] : m[0] + m[1];
                    q1[i] = s == 0 ? m[1] + m[3] : m[2] + m[3];
                }
                pe.assign(deg + 1, 0.0); po.assign(deg + 1, 0.0);
                se.assign(deg + 1, 0.0); so.assign(deg + 1, 0.0);
                pe[0] = 1.0; po[0] = 0.0;
                for (int i = 0; i < deg; ++i) {
                    double ne = pe[i] * q0[i] + po[i] * q1[i];
                    double no = pe[i] * q1[i] + po[i] * q0[i];
                    pe[i + 1] = ne; po[i + 1] = no;
                }
                se[deg] = 1.0; so[deg] = 0.0;
                for (int i = deg - 1; i >= 0; --i) {
                    double ne = se[i + 1] * q0[i] + so[i + 1] * q1[i];
                    double no = se[i + 1] * q1[i] + so[i + 1] * q0[i];
                    se[i] = ne; so[i] = no;
                }
                for (int i = 0; i < deg; ++i) {
                    double p_even = pe[i] * se[i + 1] + po[i] * so[i + 1];
                    double p_odd = pe[i] * so[i + 1] + po[i] * se[i + 1];
                    double v0 = target[s][c] == 0 ? p_even : p_odd;
                    double v1 = target[s][c] == 0 ? p_odd : p_even;
                    Msg out = s == 0 ? Msg{v0, v1, v0, v1} : Msg{v0, v0, v1, v1};
                    normalize(out);
                    ws.c2v[s][off + i] = out;
                }
            }
        }
        const Msg ident{1.0, 1.0, 1.0, 1.0};
        Msg pref[2][64], suf[2][64];
        for (int v = 0; v < code.n; ++v) {
            int deg[2];
            for (int s = 0; s < 2; ++s) {
                int b = code.var_off[s][v];
                deg[s] = code.var_off[s][v + 1] - b;
                pref[s][0] = ident;
                for (int i = 0; i < deg[s]; ++i)
                    pref[s][i + 1] = mul(pref[s][i], ws.c2v[s][code.var_edge[s][b + i]]);
                suf[s][deg[s]] = ident;
                for (int i = deg[s]; i-- > 0;)
                    suf[s][i] = mul(suf[s][i + 1], ws.c2v[s][code.var_edge[s][b + i]]);
            }
            Msg total = mul(prior, mul(pref[0][deg[0]], pref[1][deg[1]]));
            normalize(total);
            int best = 0;
            for (int k = 1; k < 4; ++k)
                if (total[k] > total[best]) best = k;
            ws.est[v] = best;
            if (!frozen[0]) {
                int b = code.var_off[0][v];
                for (int i = 0; i < deg[0]; ++i) {
                    Msg out = mul(mul(prior, mul(pref[0][i], suf[0][i + 1])), pref[1][deg[1]]);
                    normalize(out);
                    Msg &slot = ws.v2c[0][code.var_edge[0][b + i]];
                    slot = scaled_mix(out, slot, damping);
                }
            }
            if (!frozen[1]) {
                int b = code.var_off[1][v];
                for (int i = 0; i < deg[1]; ++i) {
                    Msg out = mul(mul(prior, pref[0][deg[0]]), mul(pref[1][i], suf[1][i + 1]));
                    normalize(out);
                    Msg &slot = ws.v2c[1][code.var_edge[1][b + i]];
                    slot = scaled_mix(out, slot, damping);
                }
            }
        }
        bool ok[2];
        for (int s = 0; s < 2; ++s) {
            ws.shat.assign(code.m[s], 0);
            syndrome(code, s, ws.est.data(), ws.shat.data());
            ok[s] = std::equal(ws.shat.begin(), ws.shat.end(), target[s]);
        }
        if (freeze_syn) {
            for (int s = 0; s < 2; ++s) {
                if (!ok[s] || frozen[s]) continue;
                frozen[s] = true;
                // freeze_side_msgs (jointbp_ets.cpp:1884-1901)
                for (int e = 0; e < static_cast<int>(code.edge_var[s].size()); ++e) {
                    int bit = s == 0 ? xbit(ws.est[code.edge_var[s][e]]) : zbit(ws.est[code.edge_var[s][e]]);
                    Msg m = s == 0 ? (bit ? Msg{0.0, 0.5, 0.0, 0.5} : Msg{0.5, 0.0, 0.5, 0.0})
                                   : (bit ? Msg{0.0, 0.0, 0.5, 0.5} : Msg{0.5, 0.5, 0.0, 0.0});
                    ws.v2c[s][e] = m;
                    ws.c2v[s][e] = m;
                }
            }
        }
        if (ok[0] && ok[1]) {
            std::copy(ws.est.begin(), ws.est.end(), est_out);
            *syn_match = 1;
            return iter + 1;
        }
    }
    std::copy(ws.est.begin(), ws.est.end(), est_out);
    *syn_match = 0;
    return max_iter;
}

// jointbp_ets.cpp:5484-5485
Msg make_prior(double p) {
    Msg prior{1.0 - p, p / 3.0, p / 3.0, p / 3.0};
    normalize(prior);
    return prior;
}

}  // namespace

extern "C" {

void *orc_code_from_params(const char *path) {
    auto *code = new Code();
    if (!code_from_params(path, *code)) { delete code; return nullptr; }
    return code;
}

// Flat check lists: `deg` entries per check, -1 = absent.
void *orc_code_from_checks(int n, int mx, int xdeg, const int *xflat, int mz, int zdeg, const int *zflat) {
    auto *code = new Code();
    code->n = n;
    std::vector<std::vector<int>> checks[2];
    checks[0].assign(mx, std::vector<int>(xdeg));
    checks[1].assign(mz, std::vector<int>(zdeg));
    for (int c = 0; c < mx; ++c) for (int i = 0; i < xdeg; ++i) checks[0][c][i] = xflat[c * xdeg + i];
    for (int c = 0; c < mz; ++c) for (int i = 0; i < zdeg; ++i) checks[1][c][i] = zflat[c * zdeg + i];
    finish_code(*code, checks);
    return code;
}

void orc_code_free(void *h) { delete static_cast<Code *>(h); }

// out: n, mx, mz, rank_x, rank_z, edges_x, edges_z
void orc_code_info(const void *h, long long *out) {
    const Code &c = *static_cast<const Code *>(h);
    out[0] = c.n; out[1] = c.m[0]; out[2] = c.m[1];
    out[3] = static_cast<long long>(c.basis_rows[0].size());
    out[4] = static_cast<long long>(c.basis_rows[1].size());
    out[5] = static_cast<long long>(c.edge_var[0].size());
    out[6] = static_cast<long long>(c.edge_var[1].size());
}

// Check-major CSR of one side: offsets (m+1) and edge vars (E).
void orc_code_csr(const void *h, int side, int *off, int *vars) {
    const Code &c = *static_cast<const Code *>(h);
    std::copy(c.chk_off[side].begin(), c.chk_off[side].end(), off);
    std::copy(c.edge_var[side].begin(), c.edge_var[side].end(), vars);
}

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }
uint64_t orc_trial_seed(uint64_t base, uint64_t idx) { return trial_seed(base, idx); }

// Raw mt19937_64 outputs for a seed (the engine the sampler uses).
void orc_mt_outputs(uint64_t seed, long long count, uint64_t *out) {
    std::mt19937_64 rng(seed);
    for (long long i = 0; i < count; ++i) out[i] = rng();
}

void orc_sample(const void *h, uint64_t base, uint64_t idx, double p, int *err) {
    sample(*static_cast<const Code *>(h), base, idx, p, err);
}

void orc_syndrome(const void *h, const int *err, int *sx, int *sz) {
    const Code &c = *static_cast<const Code *>(h);
    syndrome(c, 0, err, sx);
    syndrome(c, 1, err, sz);
}

void orc_prior(double p, double *out4) {
    Msg m = make_prior(p);
    std::copy(m.begin(), m.end(), out4);
}

int orc_decode(const void *h, const int *sx, const int *sz, const double *prior4, int max_iter,
               double damping, int freeze_syn, int *est, int *syn_match) {
    Workspace ws;
    Msg prior{prior4[0], prior4[1], prior4[2], prior4[3]};
    return decode(*static_cast<const Code *>(h), sx, sz, prior, max_iter, damping, freeze_syn != 0,
                  ws, est, syn_match);
}

int orc_stab_equiv(const void *h, const int *err, const int *est) {
    return stab_equiv(*static_cast<const Code *>(h), err, est) ? 1 : 0;
}

// Membership of an arbitrary bit vector (one int per qubit) in the row
// space of side `basis_side` (0 = Hx, 1 = Hz).
int orc_in_row_space(const void *h, int basis_side, const int *bits) {
    const Code &c = *static_cast<const Code *>(h);
    std::vector<int> zero(c.n, 0), ones(c.n);
    for (int i = 0; i < c.n; ++i) ones[i] = bits[i] ? 1 : 0;  // xbit(1) = 1
    return in_row_space(c, basis_side, zero.data(), ones.data(), true) ? 1 : 0;
}

// The simulate loop with PP off (jointbp_ets.cpp:5931-6207) over trial
// indices first..first+count-1 on `threads` workers.  Per-trial outputs:
// iterations, syndrome match, success, exact.  `est_hash` (may be null)
// receives a 64-bit FNV-1a hash of the final estimate.  Returns seconds.
double orc_simulate(const void *h, uint64_t base, long long first, long long count, double p,
                    int max_iter, double damping, int threads, int *iters, unsigned char *syn_ok,
                    unsigned char *success, unsigned char *exact, uint64_t *est_hash) {
    const Code &code = *static_cast<const Code *>(h);
    Msg prior = make_prior(p);
    std::atomic<long long> next{0};
    auto worker = [&]() {
        Workspace ws;
        std::vector<int> err(code.n), est(code.n), sx(code.m[0]), sz(code.m[1]);
        for (;;) {
            long long t = next.fetch_add(1);
            if (t >= count) break;
            sample(code, base, static_cast<uint64_t>(first + t), p, err.data());
            syndrome(code, 0, err.data(), sx.data());
            syndrome(code, 1, err.data(), sz.data());
            int ok = 0;
            iters[t] = decode(code, sx.data(), sz.data(), prior, max_iter, damping, false, ws, est.data(), &ok);
            bool eq = ok && stab_equiv(code, err.data(), est.data());
            syn_ok[t] = static_cast<unsigned char>(ok);
            success[t] = (ok && eq) ? 1 : 0;
            exact[t] = (est == err) ? 1 : 0;
            if (est_hash) {
                uint64_t hsh = 1469598103934665603ull;
                for (int v : est) { hsh ^= static_cast<uint64_t>(v); hsh *= 1099511628211ull; }
                est_hash[t] = hsh;
            }
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

}  // extern "C"
