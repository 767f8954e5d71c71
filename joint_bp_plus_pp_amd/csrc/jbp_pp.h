// Post-processing (PP) of the reference decoder, host side: trapping-set
// tables (ETS), the three repair methods (ETS lookup, flip-set solve, OSD)
// and the per-trial trigger state machine that decides when they run.
//
// The reference runs these inside its CPU decoder loop
// (jointbp_ets.cpp:2536-2735 triggers, 2736-3074 in-loop attempt) and again
// after the decode (5993-6197).  Here belief propagation itself runs on the
// device (jbp_pp_stage.hip); this module sees one iteration's outcome at a
// time and applies the repairs on host threads.  Estimates are one byte per
// qubit (0 I, 1 X, 2 Z, 3 Y); syndromes one byte per check.
#pragma once

#include <cstdint>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

namespace jbp {
namespace pp {

// Both Tanner graphs as check lists (check-major, the order the syndrome is
// defined in) and variable lists (checks of a variable in edge order).
struct Graph {
    int n = 0;
    int m[2] = {0, 0};
    std::vector<int> coff[2], cvar[2];  // check c: cvar[coff[c] .. coff[c+1])
    std::vector<int> voff[2], vchk[2];  // variable v: vchk[voff[v] .. voff[v+1])
};

// Graph from check-major CSR (offsets, variable per edge) of both sides.
Graph graph_from_csr(int n, const int *m, const int *const *chk_off, const int *const *edge_var);

// One side of one trapping-set table: entries of nv variables with the two
// checks they leave unsatisfied, indexed by that pair
// (jointbp_ets.cpp:3112-3123).
struct EtsSide {
    int nv = 0;
    std::vector<int> vars;   // entries x nv, ascending within an entry
    std::vector<int> unsat;  // entries x 2, ascending
    std::unordered_map<uint64_t, std::vector<int>> by_pair;
    int count() const { return nv ? static_cast<int>(vars.size() / nv) : 0; }
};

struct EtsGroup {
    int size = 0;
    std::string label;
    bool has[2] = {false, false};
    EtsSide side[2];
};

// Groups in the order they are tried (jointbp_ets.cpp:3644-3668).
struct Tables {
    std::vector<EtsGroup> groups;
};

// Reads one table file in the reference's text format (jointbp_ets.cpp:
// 3199-3258): lines "[X|Z] ... vars=Vc:i,... unsat=Cr:i,Cr:i", variable
// Vc:i = c*P + i, check Cr:i = r*P + i; lines of the other side, with a
// different variable count or malformed are skipped.  Returns -1 when the
// file cannot be read (or nv is outside 1..32), else the entries loaded.
int load_ets_file(const std::string &path, long long P, int n, int m_side, int nv, char side,
                  EtsSide &out);

struct EtsFix {
    bool applied = false;
    int size = 0;
    std::string label;
};

// Syndrome of est on one side (X checks read the X bit, Z checks the Z bit).
void syndrome(const Graph &g, int side, const uint8_t *est, uint8_t *out);
int unsatisfied(const Graph &g, int side, const uint8_t *est, const uint8_t *target);

// The syndrome difference of an estimate on one side, d[c] = syndrome[c] XOR
// target[c], with its count of ones (the unsatisfied checks).  The repairs
// keep it in step with the estimate: toggling a variable flips the bits of
// its checks, so a candidate is verified without a syndrome pass.
struct SynDiff {
    std::vector<uint8_t> d;
    int ones = 0;
};
void syn_diff_init(const Graph &g, int side, const uint8_t *est, const uint8_t *target, SynDiff &sd);
// The same from bit words already holding syndrome(est) XOR target (bit c of
// word c / 32), as a device computes them beside the estimate.
void syn_diff_from_bits(const Graph &g, int side, const uint32_t *bits, SynDiff &sd);

// ETS repair of one side (jointbp_ets.cpp:3525-3668): applies to est the
// first table entry, over the groups in order, whose variables explain the
// two unsatisfied checks.
EtsFix ets_fix(const Graph &g, const Tables &t, int side, const uint8_t *target,
               std::vector<uint8_t> &est);
// ets_fix on the estimate's syndrome difference (updated when it applies).
EtsFix ets_fix_diff(const Graph &g, const Tables &t, int side, SynDiff &diff, std::vector<uint8_t> &est);

// Flip-set repair (jointbp_ets.cpp:3726-3880): flips the variables of `vars`
// that make est match the target syndrome, when that set of flips exists
// and is unique.
bool flip_fix(const Graph &g, int side, const std::vector<int> &vars, const uint8_t *target,
              std::vector<uint8_t> &est);

// flip_fix on the estimate's syndrome difference (updated when it applies).
bool flip_fix_diff(const Graph &g, int side, const std::vector<int> &vars, SynDiff &diff,
                   std::vector<uint8_t> &est);

// One side's |LLR|, as values or as the normalised marginals [n][4] they
// come from (jointbp_ets.cpp:2483-2493).  From marginals OSD computes the
// exact values (std::log) only around the prefix of the order it walks.
struct LlrSide {
    const double *abs_llr = nullptr;
    const double *tot4 = nullptr;
    static LlrSide values(const double *v) { LlrSide l; l.abs_llr = v; return l; }
    static LlrSide marginals(const double *t) { LlrSide l; l.tot4 = t; return l; }
    bool empty() const { return !abs_llr && !tot4; }
};

// Ordered-statistics repair (jointbp_ets.cpp:4078-4386): with the variables
// ordered by increasing |LLR|, the shortest prefix whose columns reach the
// syndrome difference; applied when those columns are independent and the
// correction has weight <= 30.
bool osd_fix(const Graph &g, int side, const LlrSide &llr, const uint8_t *target,
             std::vector<uint8_t> &est);
bool osd_fix_diff(const Graph &g, int side, const LlrSide &llr, SynDiff &diff,
                  std::vector<uint8_t> &est);

// The reference's sets of flip candidates (jointbp_ets.cpp:3670-3689):
// the union of the recent flip history, each recent iteration's flips
// newest first, or the last flips alone when there is no history.
std::vector<const std::vector<int> *> flip_sets(const std::vector<int> &uni,
                                               const std::deque<std::vector<int>> &window,
                                               const std::vector<int> &last);

// |LLR| of the X and Z parts of a normalised marginal (jointbp_ets.cpp:
// 2483-2493).
void abs_llr_from_marginals(int n, const double *tot4, double *llr_x, double *llr_z);

// One BP iteration's outcome as seen by the trigger logic.
struct IterRec {
    int us[2] = {0, 0};  // unsatisfied checks per side
    bool syn_all = false;
    bool same = false;   // estimate equals the previous iteration's
    bool osc = false;    // equals the one before that and differs from the previous
};

// Decode-loop state of one trial with PP on (jointbp_ets.cpp:1986-2010,
// 2536-2629) and its outcome.
struct Trial {
    int stall_same = 0;
    bool stall = false, oscill = false, us_oscill = false, hist_on = false;
    std::deque<int> us_hist;
    std::deque<std::vector<int>> win[2];
    std::vector<int> last[2];
    // outcome
    bool done = false;
    bool bp_ok = false;
    bool ok = false;
    int iterations = 0;
    bool pp_success = false, pp_ets = false, pp_flip = false;
    bool used_any = false, used6 = false, used12 = false;
};

enum class Step { kContinue, kConverged, kAttempt };

struct Params {
    int max_iter = 0;
    int hist_window = 5;  // --flip-hist
};

// Bookkeeping of iteration t (0-based): triggers, flip history; est and
// prev (prev = nullptr when t == 0) are this and the previous iteration's
// estimates.  Returns whether the trial converged, needs a PP attempt now,
// or continues.
Step observe(const Graph &g, const Params &prm, Trial &tr, int t, const IterRec &rec,
             const uint8_t *est, const uint8_t *prev);

// The in-loop attempt at iteration t (jointbp_ets.cpp:2761-3073) in two
// parts: attempt_begin tries ETS and the flip sets on a copy of est and
// says whether OSD is to run (it needs this iteration's |LLR|, which the
// caller fetches only then); attempt_osd runs it; attempt_end commits a
// successful repair (est receives it, the trial is done).
struct Attempt {
    std::vector<uint8_t> trial;
    SynDiff diff[2];  // of trial, X and Z
    bool ok = false, want_osd = false;
    bool ets_used = false, ets6 = false, ets12 = false, ets_ok = false, flip_ok = false;
};
// (dbits: optional syndrome differences of est, X and Z, as bit words; else
// they are computed from sx, sz)
void attempt_begin(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                   const std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz,
                   Attempt &a, const uint32_t *const *dbits = nullptr);
void attempt_osd(const Graph &g, Attempt &a, const LlrSide &llr_x, const LlrSide &llr_z);
bool attempt_end(Trial &tr, Attempt &a, int t, std::vector<uint8_t> &est);

// The post-decode pass over a trial that ended unconverged
// (jointbp_ets.cpp:5993-6197), in place on est.  Fills the post-decode
// flags below; llr may be nullptr (no marginals kept).
struct PostOut {
    bool ok = false;
    bool ets_applied = false, ets_ok = false, flip_ok = false, osd_ok = false;
    bool used6 = false, used12 = false;
    bool want_osd = false;  // (post_begin) OSD is to run: it needs |LLR|
};
PostOut post_decode(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                    std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz,
                    const double *llr_x, const double *llr_z);
// The same pass in two parts, so that |LLR| are fetched only when OSD runs:
// post_begin does ETS and the flip sets and sets want_osd; post_osd finishes.
// diff: the caller's scratch, carried from one to the other.
PostOut post_begin(const Graph &g, const Tables &t, const Params &prm, const Trial &tr,
                   std::vector<uint8_t> &est, const uint8_t *sx, const uint8_t *sz, SynDiff (&diff)[2],
                   const uint32_t *const *dbits = nullptr);
void post_osd(const Graph &g, PostOut &o, std::vector<uint8_t> &est, SynDiff (&diff)[2],
              const LlrSide &llr_x, const LlrSide &llr_z);

}  // namespace pp
}  // namespace jbp

// The C ABI's opaque table set (include/jbp.h).
struct jbp_pp_tables {
    jbp::pp::Tables t;
};

