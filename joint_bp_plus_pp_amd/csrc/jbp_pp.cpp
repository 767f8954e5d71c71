// Host side of post-processing: trapping-set tables, the ETS / flip / OSD
// repairs and the trigger state machine (declarations and reference
// citations in jbp_pp.h).
//
// The linear algebra differs in form from the reference and agrees in
// result.  The reference solves each repair system by dense Gauss-Jordan
// elimination, and for OSD re-eliminates once per probe of two binary
// searches over the prefix length k (jointbp_ets.cpp:3882-3949, 4166-4239).
// Every quantity it reads from those eliminations is fixed by the column
// space alone: whether the syndrome difference lies in the span of the
// chosen columns, whether those columns are independent, and then the unique
// combination.  SpanBasis below answers all three while the columns are
// added one at a time, so OSD walks the |LLR| order once and stops at the
// first prefix that reaches the difference (or at the first dependent
// column, where the reference's uniqueness test fails).
#include "jbp_pp.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace jbp {
namespace pp {

namespace {

constexpr int kUsLimit = 20;          // pp_us_limit, osd_us_limit (jointbp_ets.cpp:2008-2009)
constexpr int kWarmupIters = 100;     // pp_warmup_iters (2003)
constexpr int kUsWindow = 100;        // us_osc_window (2004)
constexpr int kUsMinTurns = 2;        // us_osc_min_turns (2005)
constexpr int kUsMinRange = 10;       // us_osc_min_range (2006)
constexpr double kUsMinRangeFrac = 0.2;  // us_osc_min_range_frac (2007)
constexpr int kOsdWeightLimit = 30;   // osd_weight_limit (4257)

inline int part(int side, uint8_t s) { return side == 0 ? (s & 1) : ((s >> 1) & 1); }
inline void toggle(int side, uint8_t &s) { s ^= static_cast<uint8_t>(side == 0 ? 1 : 2); }

std::string trimmed(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

// Comma-separated fields, each trimmed; a trailing empty field is dropped.
std::vector<std::string> fields(const std::string &s) {
    std::vector<std::string> out;
    size_t start = 0;
    for (;;) {
        size_t comma = s.find(',', start);
        if (comma == std::string::npos) {
            if (start < s.size()) out.push_back(trimmed(s.substr(start)));
            return out;
        }
        out.push_back(trimmed(s.substr(start, comma - start)));
        start = comma + 1;
    }
}

// "<tag><block>:<index>" -> block * P + index.
bool node_id(const std::string &tok, char tag, long long P, long long &id) {
    if (tok.size() < 3 || std::toupper(static_cast<unsigned char>(tok[0])) != tag) return false;
    const size_t colon = tok.find(':');
    if (colon == std::string::npos) return false;
    const std::string a = tok.substr(1, colon - 1), b = tok.substr(colon + 1);
    char *end = nullptr;
    const long long blk = std::strtoll(a.c_str(), &end, 10);
    if (end == a.c_str()) return false;
    const long long idx = std::strtoll(b.c_str(), &end, 10);
    if (end == b.c_str()) return false;
    id = blk * P + idx;
    return true;
}

// Solution of rows * x = rhs over GF(2) for at most 32 unknowns (bit i of a
// row = unknown i), free unknowns set to 0; false when inconsistent.
bool solve32(std::vector<uint32_t> rows, std::vector<uint8_t> rhs, int nv, uint32_t &x) {
    const int m = static_cast<int>(rows.size());
    int pivot_row[32];
    for (int c = 0; c < 32; ++c) pivot_row[c] = -1;
    int r = 0;
    for (int c = 0; c < nv && r < m; ++c) {
        const uint32_t bit = 1u << c;
        int sel = r;
        while (sel < m && !(rows[sel] & bit)) ++sel;
        if (sel == m) continue;
        std::swap(rows[sel], rows[r]);
        std::swap(rhs[sel], rhs[r]);
        for (int i = 0; i < m; ++i)
            if (i != r && (rows[i] & bit)) {
                rows[i] ^= rows[r];
                rhs[i] ^= rhs[r];
            }
        pivot_row[c] = r++;
    }
    for (int i = r; i < m; ++i)
        if (rhs[i]) return false;  // every row at or below r is zero by now
    x = 0;
    for (int c = 0; c < nv; ++c)
        if (pivot_row[c] >= 0 && rhs[pivot_row[c]]) x |= 1u << c;
    return true;
}

// Span of columns over GF(2)^m, grown one column at a time, with a target
// vector kept reduced against it.  Each basis vector is stored with its
// lowest set bit as its pivot (distinct across the basis) and the set of
// added columns it is the sum of.
class SpanBasis {
public:
    SpanBasis(int m, int max_cols)
        : m_(m), w_((m + 63) / 64), cw_((std::max(max_cols, 1) + 63) / 64), owner_(m, -1),
          tv_(w_, 0), tc_(cw_, 0), tv_tmp_(w_), tc_tmp_(cw_) {}

    void set_target(const uint8_t *bits) {
        std::fill(tv_.begin(), tv_.end(), 0);
        std::fill(tc_.begin(), tc_.end(), 0);
        for (int i = 0; i < m_; ++i)
            if (bits[i]) tv_[i >> 6] |= 1ull << (i & 63);
        reduce(tv_, tc_, stuck_);
    }
    bool target_reached() const { return stuck_ < 0; }
    const std::vector<uint64_t> &target_combination() const { return tc_; }

    // Adds column j (rows = the set bits, duplicates count once); false when
    // it is a sum of earlier columns (the basis is unchanged then).
    bool add(const int *rows, int nrows, int j) {
        std::fill(tv_tmp_.begin(), tv_tmp_.end(), 0);
        std::fill(tc_tmp_.begin(), tc_tmp_.end(), 0);
        for (int i = 0; i < nrows; ++i) tv_tmp_[rows[i] >> 6] |= 1ull << (rows[i] & 63);
        tc_tmp_[j >> 6] |= 1ull << (j & 63);
        int lead = -1;
        reduce(tv_tmp_, tc_tmp_, lead);
        if (lead < 0) return false;
        owner_[lead] = nb_++;
        vec_.insert(vec_.end(), tv_tmp_.begin(), tv_tmp_.end());
        comb_.insert(comb_.end(), tc_tmp_.begin(), tc_tmp_.end());
        if (lead == stuck_) reduce(tv_, tc_, stuck_);
        return true;
    }

private:
    // Clears low bits with basis vectors until the lowest set bit has no
    // owner; lead = that bit, or -1 when v became zero.
    void reduce(std::vector<uint64_t> &v, std::vector<uint64_t> &c, int &lead) {
        int w = 0;
        for (;;) {
            while (w < w_ && v[w] == 0) ++w;
            if (w == w_) {
                lead = -1;
                return;
            }
            const int bit = (w << 6) + __builtin_ctzll(v[w]);
            const int b = owner_[bit];
            if (b < 0) {
                lead = bit;
                return;
            }
            const uint64_t *bv = &vec_[static_cast<size_t>(b) * w_];
            const uint64_t *bc = &comb_[static_cast<size_t>(b) * cw_];
            for (int k = w; k < w_; ++k) v[k] ^= bv[k];
            for (int k = 0; k < cw_; ++k) c[k] ^= bc[k];
        }
    }

    int m_, w_, cw_;
    int nb_ = 0;
    std::vector<int> owner_;
    std::vector<uint64_t> vec_, comb_;
    std::vector<uint64_t> tv_, tc_;
    int stuck_ = -1;
    std::vector<uint64_t> tv_tmp_, tc_tmp_;
};

// Rows (checks of this side) of variable v.
inline const int *col_rows(const Graph &g, int side, int v, int &count) {
    count = g.voff[side][v + 1] - g.voff[side][v];
    return g.vchk[side].data() + g.voff[side][v];
}

// Variable v's side bit toggled: the bits of its checks flip in the
// syndrome difference.
inline void diff_toggle(const Graph &g, int side, SynDiff &sd, int v) {
    int cnt = 0;
    const int *rows = col_rows(g, side, v, cnt);
    for (int k = 0; k < cnt; ++k) {
        uint8_t &b = sd.d[rows[k]];
        b ^= 1;
        sd.ones += b ? 1 : -1;
    }
}

// One table of ETS repair (jointbp_ets.cpp:3525-3642).
// (diff: est's syndrome difference, whose two set checks form `key`)
bool ets_one(const Graph &g, const EtsSide &es, int side, SynDiff &diff, uint64_t key,
             std::vector<uint8_t> &est) {
    const int nv = es.nv;
    if (es.count() == 0 || nv <= 0 || nv > 32) return false;
    auto it = es.by_pair.find(key);
    if (it == es.by_pair.end()) return false;
    std::vector<int> touched;
    std::vector<uint32_t> rows;
    std::vector<uint8_t> rhs;
    for (int idx : it->second) {
        const int *vars = &es.vars[static_cast<size_t>(idx) * nv];
        touched.clear();
        for (int i = 0; i < nv; ++i) {
            int cnt = 0;
            const int *cs = col_rows(g, side, vars[i], cnt);
            for (int k = 0; k < cnt; ++k)
                if (std::find(touched.begin(), touched.end(), cs[k]) == touched.end())
                    touched.push_back(cs[k]);
        }
        // (the reference also rejects an entry whose checks include an
        // unsatisfied one other than the pair; with exactly two unsatisfied
        // checks that cannot happen)
        rows.clear();
        rhs.clear();
        for (int c : touched) {
            uint32_t mask = 0;
            for (int e = g.coff[side][c]; e < g.coff[side][c + 1]; ++e) {
                const int v = g.cvar[side][e];
                for (int i = 0; i < nv; ++i)
                    if (vars[i] == v) {
                        mask |= 1u << i;
                        break;
                    }
            }
            if (!mask) continue;
            rows.push_back(mask);
            rhs.push_back(diff.d[c]);
        }
        uint32_t x = 0;
        if (!solve32(rows, rhs, nv, x)) continue;
        // the entry's flips must clear the whole difference
        for (int i = 0; i < nv; ++i)
            if (x & (1u << i)) diff_toggle(g, side, diff, vars[i]);
        if (diff.ones == 0) {
            for (int i = 0; i < nv; ++i)
                if (x & (1u << i)) toggle(side, est[vars[i]]);
            return true;
        }
        for (int i = 0; i < nv; ++i)
            if (x & (1u << i)) diff_toggle(g, side, diff, vars[i]);
    }
    return false;
}

// detect_us_oscillation (jointbp_ets.cpp:2028-2059)
bool us_oscillation(const std::deque<int> &h) {
    if (static_cast<int>(h.size()) < kUsWindow) return false;
    int lo = h.front(), hi = h.front();
    long long sum = 0;
    for (int v : h) {
        lo = std::min(lo, v);
        hi = std::max(hi, v);
        sum += v;
    }
    const double mean = sum / static_cast<double>(h.size());
    if (mean <= 0.0) return false;
    const int need = std::max(kUsMinRange, static_cast<int>(mean * kUsMinRangeFrac));
    if (hi - lo < need) return false;
    int turns = 0, dir = 0;
    for (size_t i = 1; i < h.size(); ++i) {
        const int d = h[i] - h[i - 1];
        const int s = d > 0 ? 1 : (d < 0 ? -1 : 0);
        if (!s) continue;
        if (dir && s != dir) ++turns;
        dir = s;
    }
    return turns >= kUsMinTurns;
}

std::vector<int> window_union(const std::deque<std::vector<int>> &win, int n) {
    std::vector<uint8_t> seen(n, 0);
    for (const auto &w : win)
        for (int v : w)
            if (v >= 0 && v < n) seen[v] = 1;
    std::vector<int> out;
    for (int v = 0; v < n; ++v)
        if (seen[v]) out.push_back(v);
    return out;
}

// The flip-set loops of both sides (jointbp_ets.cpp:2951-2978 and
// 6144-6166): the first set that repairs a side ends that side's loop.
bool flip_pass(const Graph &g, const Params &prm, const Trial &tr, std::vector<uint8_t> &est,
               SynDiff (&diff)[2]) {
    bool any = false;
    for (int side = 0; side < 2; ++side) {
        std::vector<int> uni;
        if (tr.hist_on && prm.hist_window > 0) uni = window_union(tr.win[side], g.n);
        // est only changes when a set applies, which ends this side's loop
        const std::vector<int> *tried = nullptr;
        for (const std::vector<int> *set : flip_sets(uni, tr.win[side], tr.last[side])) {
            if (set->empty()) continue;
            // the same set again (the union of a one-entry history is that
            // entry) fails again
            if (tried && *tried == *set) continue;
            tried = set;
            if (flip_fix_diff(g, side, *set, diff[side], est)) {
                any = true;
                break;
            }
        }
    }
    return any;
}

}  // namespace

Graph graph_from_csr(int n, const int *m, const int *const *chk_off, const int *const *edge_var) {
    Graph g;
    g.n = n;
    for (int s = 0; s < 2; ++s) {
        g.m[s] = m[s];
        const int E = chk_off[s][m[s]];
        g.coff[s].assign(chk_off[s], chk_off[s] + m[s] + 1);
        g.cvar[s].assign(edge_var[s], edge_var[s] + E);
        g.voff[s].assign(n + 1, 0);
        for (int e = 0; e < E; ++e) g.voff[s][g.cvar[s][e] + 1]++;
        for (int v = 0; v < n; ++v) g.voff[s][v + 1] += g.voff[s][v];
        g.vchk[s].assign(E, 0);
        std::vector<int> fill(g.voff[s].begin(), g.voff[s].end() - 1);
        for (int c = 0; c < m[s]; ++c)
            for (int e = g.coff[s][c]; e < g.coff[s][c + 1]; ++e) g.vchk[s][fill[g.cvar[s][e]]++] = c;
    }
    return g;
}

int load_ets_file(const std::string &path, long long P, int n, int m_side, int nv, char side,
                  EtsSide &out) {
    out = EtsSide();
    out.nv = nv;
    std::ifstream in(path);
    if (!in || nv <= 0 || nv > 32) return -1;
    std::string line;
    std::vector<int> vars(nv);
    while (std::getline(in, line)) {
        line = trimmed(line);
        if (line.empty()) continue;
        char line_side = 0;
        if (line.size() >= 2 && std::isspace(static_cast<unsigned char>(line[1]))) {
            const char c = static_cast<char>(std::toupper(static_cast<unsigned char>(line[0])));
            if (c == 'X' || c == 'Z') {
                line_side = c;
                line = trimmed(line.substr(1));
            }
        }
        if (line_side && line_side != side) continue;
        const size_t pv = line.find("vars="), pu = line.find("unsat=");
        if (pv == std::string::npos || pu == std::string::npos) continue;
        const auto vt = fields(trimmed(line.substr(pv + 5, pu - (pv + 5))));
        const auto ct = fields(trimmed(line.substr(pu + 6)));
        if (static_cast<int>(vt.size()) != nv || ct.size() != 2) continue;
        bool good = true;
        for (int i = 0; i < nv && good; ++i) {
            long long id = 0;
            good = node_id(vt[i], 'V', P, id) && id >= 0 && id < n;
            vars[i] = static_cast<int>(id);
        }
        int chk[2] = {0, 0};
        for (int i = 0; i < 2 && good; ++i) {
            long long id = 0;
            good = node_id(ct[i], 'C', P, id) && id >= 0 && id < m_side;
            chk[i] = static_cast<int>(id);
        }
        if (!good) continue;
        std::sort(vars.begin(), vars.end());
        if (chk[0] > chk[1]) std::swap(chk[0], chk[1]);
        const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(chk[0])) << 32) |
                             static_cast<uint32_t>(chk[1]);
        out.by_pair[key].push_back(out.count());
        out.vars.insert(out.vars.end(), vars.begin(), vars.end());
        out.unsat.push_back(chk[0]);
        out.unsat.push_back(chk[1]);
    }
    return out.count();
}

void syndrome(const Graph &g, int side, const uint8_t *est, uint8_t *out) {
    const auto &off = g.coff[side];
    const auto &var = g.cvar[side];
    for (int c = 0; c < g.m[side]; ++c) {
        int par = 0;
        for (int e = off[c]; e < off[c + 1]; ++e) par ^= part(side, est[var[e]]);
        out[c] = static_cast<uint8_t>(par);
    }
}

int unsatisfied(const Graph &g, int side, const uint8_t *est, const uint8_t *target) {
    const auto &off = g.coff[side];
    const auto &var = g.cvar[side];
    int count = 0;
    for (int c = 0; c < g.m[side]; ++c) {
        int par = 0;
        for (int e = off[c]; e < off[c + 1]; ++e) par ^= part(side, est[var[e]]);
        count += par != target[c];
    }
    return count;
}

void syn_diff_init(const Graph &g, int side, const uint8_t *est, const uint8_t *target, SynDiff &sd) {
    const int m = g.m[side];
    sd.d.resize(m);
    syndrome(g, side, est, sd.d.data());
    int ones = 0;
    for (int c = 0; c < m; ++c) {
        sd.d[c] ^= target[c] ? 1 : 0;
        ones += sd.d[c];
    }
    sd.ones = ones;
}

void syn_diff_from_bits(const Graph &g, int side, const uint32_t *bits, SynDiff &sd) {
    const int m = g.m[side];
    sd.d.resize(m);
    int ones = 0;
    for (int c = 0; c < m; ++c) {
        const uint8_t b = static_cast<uint8_t>((bits[c >> 5] >> (c & 31)) & 1u);
        sd.d[c] = b;
        ones += b;
    }
    sd.ones = ones;
}

EtsFix ets_fix(const Graph &g, const Tables &t, int side, const uint8_t *target,
               std::vector<uint8_t> &est) {
    if (t.groups.empty()) return EtsFix();
    SynDiff sd;
    syn_diff_init(g, side, est.data(), target, sd);
    return ets_fix_diff(g, t, side, sd, est);
}

EtsFix ets_fix_diff(const Graph &g, const Tables &t, int side, SynDiff &diff, std::vector<uint8_t> &est) {
    EtsFix r;
    if (t.groups.empty()) return r;
    // every table needs exactly two unsatisfied checks, and est only changes
    // when a table applies (which ends the search)
    if (diff.ones != 2) return r;
    int unsat[2], nu = 0;
    for (int c = 0; c < g.m[side] && nu < 2; ++c)
        if (diff.d[c]) unsat[nu++] = c;
    const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(unsat[0])) << 32) |
                         static_cast<uint32_t>(unsat[1]);
    for (const EtsGroup &grp : t.groups) {
        if (!grp.has[side]) continue;
        if (ets_one(g, grp.side[side], side, diff, key, est)) {
            r.applied = true;
            r.size = grp.size;
            r.label = grp.label;
            break;
        }
    }
    return r;
}

bool flip_fix(const Graph &g, int side, const std::vector<int> &vars_in, const uint8_t *target,
              std::vector<uint8_t> &est) {
    SynDiff sd;
    syn_diff_init(g, side, est.data(), target, sd);
    return flip_fix_diff(g, side, vars_in, sd, est);
}

bool flip_fix_diff(const Graph &g, int side, const std::vector<int> &vars_in, SynDiff &diff,
                   std::vector<uint8_t> &est) {
    std::vector<int> vars;
    for (int v : vars_in)
        if (v >= 0 && v < g.n) vars.push_back(v);
    std::sort(vars.begin(), vars.end());
    vars.erase(std::unique(vars.begin(), vars.end()), vars.end());
    if (vars.empty()) return false;
    const int m = g.m[side];
    const int nv = static_cast<int>(vars.size());
    // more columns than rows cannot all be independent
    if (nv > m) return false;
    // an unsatisfied check that no variable of the set touches makes the
    // system unsolvable (the reference's empty row with a nonzero right-hand
    // side, jointbp_ets.cpp:3822-3831); found without any elimination
    {
        std::vector<uint8_t> hit(m, 0);
        for (int v : vars) {
            int cnt = 0;
            const int *rows = col_rows(g, side, v, cnt);
            for (int k = 0; k < cnt; ++k) hit[rows[k]] = 1;
        }
        for (int c = 0; c < m; ++c)
            if (diff.d[c] && !hit[c]) return false;
    }
    SpanBasis basis(m, nv);
    basis.set_target(diff.d.data());
    for (int j = 0; j < nv; ++j) {
        int cnt = 0;
        const int *rows = col_rows(g, side, vars[j], cnt);
        if (!basis.add(rows, cnt, j)) return false;  // not unique
    }
    if (!basis.target_reached()) return false;  // not solvable
    const auto &comb = basis.target_combination();
    for (int j = 0; j < nv; ++j)
        if ((comb[j >> 6] >> (j & 63)) & 1) diff_toggle(g, side, diff, vars[j]);
    if (diff.ones != 0) {  // (the combination reaches the difference: never)
        for (int j = 0; j < nv; ++j)
            if ((comb[j >> 6] >> (j & 63)) & 1) diff_toggle(g, side, diff, vars[j]);
        return false;
    }
    for (int j = 0; j < nv; ++j)
        if ((comb[j >> 6] >> (j & 63)) & 1) toggle(side, est[vars[j]]);
    return true;
}

namespace {

// Clamped pair sums of a marginal: P(bit = 1), P(bit = 0) of the side's
// Pauli bit, as jointbp_ets.cpp:2483-2493 forms them.
inline void llr_pair(const double *t, int side, double &p1, double &p0) {
    const double eps = 1e-300;
    p1 = side == 0 ? t[1] + t[3] : t[2] + t[3];
    p0 = side == 0 ? t[0] + t[2] : t[0] + t[1];
    p1 = std::max(p1, eps);
    p0 = std::max(p0, eps);
}

}  // namespace

bool osd_fix(const Graph &g, int side, const LlrSide &llr, const uint8_t *target,
             std::vector<uint8_t> &est) {
    if (llr.empty() || g.m[side] == 0) return false;
    SynDiff sd;
    syn_diff_init(g, side, est.data(), target, sd);
    return osd_fix_diff(g, side, llr, sd, est);
}

bool osd_fix_diff(const Graph &g, int side, const LlrSide &llr, SynDiff &diff,
                  std::vector<uint8_t> &est) {
    if (llr.empty() || g.m[side] == 0) return false;
    const int m = g.m[side], n = g.n;
    if (diff.ones == 0) return false;
    // |LLR| order with ties by index (the reference's stable sort); only
    // the prefix the walk reaches is sorted, in growing chunks.
    //
    // From marginals, |LLR| = |log p1 - log p0| (the reference's value, with
    // its std::log) is computed only for the candidates of each chunk.  The
    // candidates come from the key k = max(p1, p0) / min(p1, p0): with p1, p0
    // in [1e-300, 1e300] and logs within 2 ulp, |LLR| and ln k differ by less
    // than E = 1e-12.  Take kt, the chunk's last key in key order, and as
    // candidates every remaining variable with k <= kt (1 + 1e-9).  At least
    // a chunk's worth of variables have k <= kt, so |LLR| <= ln kt + E; a
    // variable with k above the bound has |LLR| > ln kt + 1e-9 - E, strictly
    // more, and is not in the chunk.  Sorting the candidates by the exact
    // value therefore gives the chunk exactly.
    std::vector<int> order(n);
    for (int v = 0; v < n; ++v) order[v] = v;
    std::vector<double> val, key;
    const double *L = llr.abs_llr;
    bool lazy = false;
    auto exact = [&](int v) {
        double p1, p0;
        llr_pair(llr.tot4 + 4 * static_cast<size_t>(v), side, p1, p0);
        return std::fabs(std::log(p1) - std::log(p0));
    };
    if (!L) {
        val.assign(n, -1.0);
        key.resize(n);
        lazy = true;
        for (int v = 0; v < n && lazy; ++v) {
            double p1, p0;
            llr_pair(llr.tot4 + 4 * static_cast<size_t>(v), side, p1, p0);
            key[v] = p1 >= p0 ? p1 / p0 : p0 / p1;
            lazy = p1 <= 1e300 && p0 <= 1e300 && std::isfinite(key[v]);
        }
        if (!lazy)  // outside the bound's range: every value, as the reference
            for (int v = 0; v < n; ++v) val[v] = exact(v);
        L = val.data();
    }
    const auto before = [&](int a, int b) { return L[a] < L[b] || (L[a] == L[b] && a < b); };
    int sorted = 0;
    auto extend = [&](int next) {
        auto rb = order.begin() + sorted;
        if (lazy) {
            const int want = next - sorted;
            std::nth_element(rb, rb + (want - 1), order.end(), [&](int a, int b) { return key[a] < key[b]; });
            const double lim = key[*(rb + (want - 1))] * (1.0 + 1e-9);
            auto mid = std::partition(rb, order.end(), [&](int v) { return key[v] <= lim; });
            for (auto it = rb; it != mid; ++it)
                if (!(val[*it] >= 0.0)) val[*it] = exact(*it);
            std::partial_sort(rb, order.begin() + next, mid, before);
        } else {
            std::partial_sort(rb, order.begin() + next, order.end(), before);
        }
        sorted = next;
    };
    SpanBasis basis(m, std::min(n, m + 1));
    basis.set_target(diff.d.data());
    for (int j = 0; j < n; ++j) {
        if (j == sorted) extend(std::min(n, std::max(2 * sorted, sorted + 256)));
        int cnt = 0;
        const int *rows = col_rows(g, side, order[j], cnt);
        if (!basis.add(rows, cnt, j)) return false;
        if (!basis.target_reached()) continue;
        // k = j + 1 columns reach the difference and are independent
        const auto &comb = basis.target_combination();
        auto used = [&](int i) { return (comb[i >> 6] >> (i & 63)) & 1; };
        int weight = 0;
        for (int i = 0; i <= j; ++i) weight += used(i) ? 1 : 0;
        if (weight > kOsdWeightLimit) return false;
        for (int i = 0; i <= j; ++i)
            if (used(i)) diff_toggle(g, side, diff, order[i]);
        if (diff.ones != 0) {  // (the combination reaches the difference: never)
            for (int i = 0; i <= j; ++i)
                if (used(i)) diff_toggle(g, side, diff, order[i]);
            return false;
        }
        for (int i = 0; i <= j; ++i)
            if (used(i)) toggle(side, est[order[i]]);
        return true;
    }
    return false;
}

std::vector<const std::vector<int> *> flip_sets(const std::vector<int> &uni,
                                               const std::deque<std::vector<int>> &window,
                                               const std::vector<int> &last) {
    std::vector<const std::vector<int> *> sets;
    if (!uni.empty()) sets.push_back(&uni);
    if (!window.empty()) {
        for (size_t i = window.size(); i-- > 0;)
            if (!window[i].empty()) sets.push_back(&window[i]);
    } else if (!last.empty() && uni.empty()) {
        sets.push_back(&last);
    }
    return sets;
}

void abs_llr_from_marginals(int n, const double *tot4, double *llr_x, double *llr_z) {
    const double eps = 1e-300;
    for (int v = 0; v < n; ++v) {
        const double *t = tot4 + 4 * static_cast<size_t>(v);
        const double px1 = t[1] + t[3], px0 = t[0] + t[2];
        const double pz1 = t[2] + t[3], pz0 = t[0] + t[1];
        llr_x[v] = std::fabs(std::log(std::max(px1, eps)) - std::log(std::max(px0, eps)));
        llr_z[v] = std::fabs(std::log(std::max(pz1, eps)) - std::log(std::max(pz0, eps)));
    }
}

Step observe(const Graph &g, const Params &prm, Trial &tr, int t, const IterRec &rec,
             const uint8_t *est, const uint8_t *prev) {
    int us_signal = 0;
    bool skip = false;
    if (!rec.syn_all) {
        us_signal = std::max(rec.us[0], rec.us[1]);
        skip = us_signal > kUsLimit;
        tr.us_hist.push_back(us_signal);
        if (static_cast<int>(tr.us_hist.size()) > kUsWindow) tr.us_hist.pop_front();
    }
    const bool same = t >= 1 && rec.same;
    const bool osc = t >= 2 && rec.osc && !same;
    auto start_history = [&]() {
        if (!tr.hist_on && prm.hist_window > 0) {
            tr.hist_on = true;
            tr.win[0].clear();
            tr.win[1].clear();
        }
    };
    if (!rec.syn_all) {
        const bool allow = t + 1 >= kWarmupIters || t + 1 == prm.max_iter;
        tr.stall_same = same ? tr.stall_same + 1 : 0;
        if (!tr.stall && (tr.stall_same >= 1 || t + 1 == prm.max_iter) && allow) {
            tr.stall = true;
            start_history();
        }
        if (!tr.oscill && osc && allow) {
            tr.oscill = true;
            start_history();
        }
        if (!tr.us_oscill && allow && us_oscillation(tr.us_hist)) {
            tr.us_oscill = true;
            start_history();
        }
    }
    if (prev && est) {
        // qubits whose X part (bit 0) or Z part (bit 1) changed, ascending;
        // eight qubits per word compare, most words unchanged
        tr.last[0].clear();
        tr.last[1].clear();
        auto changed = [&](int v, unsigned d) {
            if (d & 1u) tr.last[0].push_back(v);
            if (d & 2u) tr.last[1].push_back(v);
        };
        int v = 0;
        for (; v + 8 <= g.n; v += 8) {
            uint64_t a, b;
            std::memcpy(&a, prev + v, 8);
            std::memcpy(&b, est + v, 8);
            const uint64_t x = (a ^ b) & 0x0303030303030303ull;
            if (x == 0) continue;
            for (int k = 0; k < 8; ++k) changed(v + k, static_cast<unsigned>(x >> (8 * k)) & 3u);
        }
        for (; v < g.n; ++v) changed(v, static_cast<unsigned>(prev[v] ^ est[v]) & 3u);
        for (int s = 0; s < 2; ++s) {
            if (tr.hist_on) {
                tr.win[s].push_back(tr.last[s]);
                if (static_cast<int>(tr.win[s].size()) > prm.hist_window) tr.win[s].pop_front();
            }
        }
    }
    if (rec.syn_all) {
        tr.done = tr.bp_ok = tr.ok = true;
        tr.iterations = t + 1;
        return Step::kConverged;
    }
    if ((tr.stall || tr.oscill || tr.us_oscill) && !skip) return Step::kAttempt;
    return Step::kContinue;
}

namespace {

bool all_clear(const SynDiff (&diff)[2]) { return diff[0].ones == 0 && diff[1].ones == 0; }

}  // namespace

void attempt_begin(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                   const std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz,
                   Attempt &a, const uint32_t *const *dbits) {
    a.ok = a.want_osd = a.ets_used = a.ets6 = a.ets12 = a.ets_ok = a.flip_ok = false;
    a.trial = est;
    if (dbits) {
        syn_diff_from_bits(g, 0, dbits[0], a.diff[0]);
        syn_diff_from_bits(g, 1, dbits[1], a.diff[1]);
    } else {
        syn_diff_init(g, 0, a.trial.data(), sx, a.diff[0]);
        syn_diff_init(g, 1, a.trial.data(), sz, a.diff[1]);
    }
    const EtsFix ex = ets_fix_diff(g, t, 0, a.diff[0], a.trial);
    const EtsFix ez = ets_fix_diff(g, t, 1, a.diff[1], a.trial);
    a.ets_used = ex.applied || ez.applied;
    a.ets6 = ex.size == 6 || ez.size == 6;
    a.ets12 = ex.size == 12 || ez.size == 12;
    if (a.ets_used) a.ets_ok = a.ok = all_clear(a.diff);
    if (!a.ok && flip_pass(g, prm, tr, a.trial, a.diff)) a.flip_ok = a.ok = all_clear(a.diff);
    if (!a.ok) a.want_osd = std::max(a.diff[0].ones, a.diff[1].ones) <= kUsLimit;
}

void attempt_osd(const Graph &g, Attempt &a, const LlrSide &llr_x, const LlrSide &llr_z) {
    if (!a.want_osd) return;
    const bool ox = osd_fix_diff(g, 0, llr_x, a.diff[0], a.trial);
    const bool oz = osd_fix_diff(g, 1, llr_z, a.diff[1], a.trial);
    if (ox || oz) a.ok = all_clear(a.diff);
}

bool attempt_end(Trial &tr, Attempt &a, int t, std::vector<uint8_t> &est) {
    if (!a.ok) return false;
    est.swap(a.trial);
    tr.done = tr.ok = tr.pp_success = true;
    tr.bp_ok = false;
    tr.iterations = t + 1;
    tr.pp_ets = a.ets_ok;
    tr.pp_flip = a.flip_ok;
    tr.used_any = a.ets_used;
    tr.used6 = a.ets6;
    tr.used12 = a.ets12;
    return true;
}

PostOut post_begin(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                   std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz, SynDiff (&diff)[2],
                   const uint32_t *const *dbits) {
    PostOut o;
    if (dbits) {
        syn_diff_from_bits(g, 0, dbits[0], diff[0]);
        syn_diff_from_bits(g, 1, dbits[1], diff[1]);
    } else {
        syn_diff_init(g, 0, est.data(), sx, diff[0]);
        syn_diff_init(g, 1, est.data(), sz, diff[1]);
    }
    const EtsFix ex = ets_fix_diff(g, t, 0, diff[0], est);
    const EtsFix ez = ets_fix_diff(g, t, 1, diff[1], est);
    o.ets_applied = ex.applied || ez.applied;
    o.used6 = ex.size == 6 || ez.size == 6;
    o.used12 = ex.size == 12 || ez.size == 12;
    if (o.ets_applied) o.ets_ok = o.ok = all_clear(diff);
    if (!o.ok && flip_pass(g, prm, tr, est, diff)) o.flip_ok = o.ok = all_clear(diff);
    if (!o.ok) o.want_osd = std::max(diff[0].ones, diff[1].ones) <= kUsLimit;
    return o;
}

void post_osd(const Graph &g, PostOut &o, std::vector<uint8_t> &est, SynDiff (&diff)[2],
              const LlrSide &llr_x, const LlrSide &llr_z) {
    if (!o.want_osd) return;
    const bool ox = osd_fix_diff(g, 0, llr_x, diff[0], est);
    const bool oz = osd_fix_diff(g, 1, llr_z, diff[1], est);
    if (ox || oz) o.osd_ok = o.ok = all_clear(diff);
}

PostOut post_decode(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                    std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz,
                    const double *llr_x, const double *llr_z) {
    SynDiff diff[2];
    PostOut o = post_begin(g, t, prm, tr, est, sx, sz, diff);
    post_osd(g, o, est, diff, LlrSide::values(llr_x), LlrSide::values(llr_z));
    return o;
}

}  // namespace pp
}  // namespace jbp
