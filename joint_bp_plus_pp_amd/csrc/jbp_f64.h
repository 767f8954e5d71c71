// FP64 joint BP in the reference's exact operation order
// (jointbp_ets.cpp:1943-2535): shared by the parity decoder and the
// post-processing stepper.  Translation units including this header are
// compiled with -ffp-contract=off (no FMA), and normalisation divides by the
// left-to-right sum as normalize_msg does (jointbp_ets.cpp:650-657).
#pragma once

#include "jbp_common.h"

namespace jbp {
namespace f64 {

struct M4 {
    double v[4];
};

__device__ __forceinline__ M4 mul4(const M4 &a, const M4 &b) {
    return M4{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}};
}

__device__ __forceinline__ void normalize4(M4 &m) {
    double sum = m.v[0] + m.v[1] + m.v[2] + m.v[3];
    if (sum <= 0.0) {
        m.v[0] = m.v[1] = m.v[2] = m.v[3] = 0.25;
        return;
    }
    m.v[0] /= sum;
    m.v[1] /= sum;
    m.v[2] /= sum;
    m.v[3] /= sum;
}

// scaled_mix (jointbp_ets.cpp:1903-1911)
__device__ __forceinline__ M4 mix4(const M4 &nw, const M4 &old, double d) {
    if (d <= 0.0) return nw;
    M4 out;
#pragma unroll
    for (int i = 0; i < 4; ++i) out.v[i] = (1.0 - d) * nw.v[i] + d * old.v[i];
    normalize4(out);
    return out;
}

// Variable-to-check messages are read by the check update only through the
// pair sums q0 = P(even), q1 = P(odd) of the side's Pauli bit (X side:
// m0 + m2, m1 + m3; Z side: m0 + m1, m2 + m3).  Without damping the variable
// update stores those two sums (PAIR: 2 doubles per row) instead of the
// four components (4 per row, needed to mix old and new messages when
// damping); the sums are formed in the same order either way.
__device__ __forceinline__ void pair_sums(const double *m, int side, double &q0, double &q1) {
    q0 = side == 0 ? m[0] + m[2] : m[0] + m[1];
    q1 = side == 0 ? m[1] + m[3] : m[2] + m[3];
}

template <bool PAIR>
__device__ __forceinline__ void store_v2c(double *v2c, size_t row, int side, const M4 &m) {
    if (PAIR) {
        double2 o;  // one 16-byte store: a wavefront writes whole cache lines
        pair_sums(m.v, side, o.x, o.y);
        *reinterpret_cast<double2 *>(v2c + 2 * row) = o;
    } else {
        double *o = v2c + 4 * row;
        o[0] = m.v[0]; o[1] = m.v[1]; o[2] = m.v[2]; o[3] = m.v[3];
    }
}

// Check update of one check (jointbp_ets.cpp:2373-2449).  Every position's
// incoming message is loaded before any is used (positions past the
// check's degree load row 0 and are ignored): loads under the degree test
// would each wait for the previous one.  Row i of a side's messages starts
// at a block-uniform address, so each load is that base plus the check's
// 32-bit offset.
// (QC: every check has degree MAXC, no offsets read)
template <int MAXC, bool PAIR, bool QC = false>
__device__ void check_update(const DevSide &sd, int side, int c, int syn, const double *v2c,
                             double *c2v) {
    const int deg = QC ? MAXC : sd.chk_off[c + 1] - sd.chk_off[c];
    const size_t m = static_cast<size_t>(sd.m);
    constexpr int W = PAIR ? 2 : 4;  // doubles per v2c row
    double q0[MAXC], q1[MAXC], pe[MAXC + 1], po[MAXC + 1];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        const double *row = v2c + W * (i < deg ? i : 0) * m;
        if (PAIR) {
            const double2 mv = *reinterpret_cast<const double2 *>(row + 2 * c);
            q0[i] = mv.x;
            q1[i] = mv.y;
        } else {
            pair_sums(row + 4 * c, side, q0[i], q1[i]);
        }
    }
    pe[0] = 1.0;
    po[0] = 0.0;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        if (i < deg) {
            pe[i + 1] = pe[i] * q0[i] + po[i] * q1[i];
            po[i + 1] = pe[i] * q1[i] + po[i] * q0[i];
        }
    }
    double se = 1.0, so = 0.0;
#pragma unroll
    for (int i = MAXC - 1; i >= 0; --i) {
        if (i < deg) {
            double p_even = pe[i] * se + po[i] * so;
            double p_odd = pe[i] * so + po[i] * se;
            double v0 = syn == 0 ? p_even : p_odd;
            double v1 = syn == 0 ? p_odd : p_even;
            double sum = side == 0 ? ((v0 + v1) + v0) + v1 : ((v0 + v0) + v1) + v1;
            double2 o;
            if (sum <= 0.0) {
                o.x = 0.25;
                o.y = 0.25;
            } else {
                o.x = v0 / sum;
                o.y = v1 / sum;
            }
            *reinterpret_cast<double2 *>(c2v + 2 * i * m + 2 * c) = o;
            double ne = se * q0[i] + so * q1[i];
            double no = se * q1[i] + so * q0[i];
            se = ne;
            so = no;
        }
    }
}

// The check update split in two for a software-pipelined pass over a
// thread's checks (quasi-cyclic codes, pair sums): check_load fetches a
// check's twelve incoming pairs, check_finish computes and stores its
// outgoing messages exactly as check_update does.  RES >= 0: position RES's
// messages are resident in LDS (res[c], the column block the variable pass
// left there) and are read and written there instead of in HBM.
template <int MAXC, int RES>
__device__ __forceinline__ void check_load(const DevSide &sd, int c, const double *v2c, const double2 *res,
                                           double (&q0)[MAXC], double (&q1)[MAXC]) {
    const size_t m = static_cast<size_t>(sd.m);
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        const double2 mv = i == RES ? res[c] : *reinterpret_cast<const double2 *>(v2c + 2 * i * m + 2 * c);
        q0[i] = mv.x;
        q1[i] = mv.y;
    }
}

template <int MAXC, int RES>
__device__ __forceinline__ void check_finish(const DevSide &sd, int side, int c, int syn, const double (&q0)[MAXC],
                                             const double (&q1)[MAXC], double *c2v, double2 *res) {
    const size_t m = static_cast<size_t>(sd.m);
    double pe[MAXC + 1], po[MAXC + 1];
    pe[0] = 1.0;
    po[0] = 0.0;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        pe[i + 1] = pe[i] * q0[i] + po[i] * q1[i];
        po[i + 1] = pe[i] * q1[i] + po[i] * q0[i];
    }
    double se = 1.0, so = 0.0;
#pragma unroll
    for (int i = MAXC - 1; i >= 0; --i) {
        double p_even = pe[i] * se + po[i] * so;
        double p_odd = pe[i] * so + po[i] * se;
        double v0 = syn == 0 ? p_even : p_odd;
        double v1 = syn == 0 ? p_odd : p_even;
        double sum = side == 0 ? ((v0 + v1) + v0) + v1 : ((v0 + v0) + v1) + v1;
        double2 o;
        if (sum <= 0.0) {
            o.x = 0.25;
            o.y = 0.25;
        } else {
            o.x = v0 / sum;
            o.y = v1 / sum;
        }
        if (i == RES) res[c] = o;
        else *reinterpret_cast<double2 *>(c2v + 2 * i * m + 2 * c) = o;
        double ne = se * q0[i] + so * q1[i];
        double no = se * q1[i] + so * q0[i];
        se = ne;
        so = no;
    }
}

// A thread's checks c0, c0 + step, ... of one side: the next check's loads
// are issued before the current one is computed (two register sets, no copy
// between them, so the loads stay in flight during the arithmetic).
template <int MAXC, int RES>
__device__ void check_pass_pipelined(const DevSide &sd, int side, int c0, int step, const uint32_t *ts,
                                     const double *v2c, double *c2v, double2 *res) {
    const int m = sd.m;
    if (c0 >= m) return;
    double a0[MAXC], a1[MAXC], b0[MAXC], b1[MAXC];
    int c = c0;
    check_load<MAXC, RES>(sd, c, v2c, res, a0, a1);
    for (;;) {
        const int c1 = c + step;
        if (c1 < m) check_load<MAXC, RES>(sd, c1, v2c, res, b0, b1);
        check_finish<MAXC, RES>(sd, side, c, (ts[c >> 5] >> (c & 31)) & 1, a0, a1, c2v, res);
        if (c1 >= m) break;
        const int c2 = c1 + step;
        if (c2 < m) check_load<MAXC, RES>(sd, c2, v2c, res, a0, a1);
        check_finish<MAXC, RES>(sd, side, c1, (ts[c1 >> 5] >> (c1 & 31)) & 1, b0, b1, c2v, res);
        if (c2 >= m) break;
        c = c2;
    }
}

// Both sides' check passes, with the column block `res` (-1: none, 0 or
// MAXC - 1) resident in LDS (lds: 2*m pairs, side 0 then side 1).
template <int MAXC>
__device__ void check_passes_qc(const DevGraph &g, int res, const uint32_t *tsx, const uint32_t *tsz,
                                double *const *v2c, double *const *c2v, double2 *lds) {
    const int tid = threadIdx.x, bs = blockDim.x, m = g.side[0].m;
    if (res < 0) {
        check_pass_pipelined<MAXC, -1>(g.side[0], 0, tid, bs, tsx, v2c[0], c2v[0], lds);
        check_pass_pipelined<MAXC, -1>(g.side[1], 1, tid, bs, tsz, v2c[1], c2v[1], lds + m);
    } else if (res == 0) {
        check_pass_pipelined<MAXC, 0>(g.side[0], 0, tid, bs, tsx, v2c[0], c2v[0], lds);
        check_pass_pipelined<MAXC, 0>(g.side[1], 1, tid, bs, tsz, v2c[1], c2v[1], lds + m);
    } else {
        check_pass_pipelined<MAXC, MAXC - 1>(g.side[0], 0, tid, bs, tsx, v2c[0], c2v[0], lds);
        check_pass_pipelined<MAXC, MAXC - 1>(g.side[1], 1, tid, bs, tsz, v2c[1], c2v[1], lds + m);
    }
}

__device__ __forceinline__ const double2 *c2v_rows(const double *c2v) {
    return reinterpret_cast<const double2 *>(c2v);
}

// (r: the edge's row)
__device__ __forceinline__ M4 c2v_msg(const double *c2v, int r, int side) {
    double a = c2v[2 * static_cast<size_t>(r)];
    double b = c2v[2 * static_cast<size_t>(r) + 1];
    return side == 0 ? M4{{a, b, a, b}} : M4{{a, a, b, b}};
}

__device__ __forceinline__ M4 c2v_pair(double a, double b, int side) {
    return side == 0 ? M4{{a, b, a, b}} : M4{{a, a, b, b}};
}

// q = a / P for 0 <= a < 2^22 through the float reciprocal, corrected to the
// exact quotient (the estimate is off by at most one).
__device__ __forceinline__ int div_small(int a, int P, float inv_p) {
    int q = static_cast<int>(static_cast<float>(a) * inv_p);
    const int r = a - q * P;
    q += r >= P ? 1 : (r < 0 ? -1 : 0);
    return q;
}

// Joint variable update + hard decision (jointbp_ets.cpp:2451-2519).  The
// suffix products are formed on the fly while walking the edges backwards;
// each outgoing message uses the same operands as the reference's
// pref[i] (*) suf[i+1].
//
// Messages: v2c[s] (4 doubles per edge), c2v[s] (2 per edge).  Returns the
// hard decision; tot (optional) receives the normalised marginal.  Up to
// degree 4 the incoming messages are all loaded first (positions past the
// degree load the side's row 0 and are ignored) and kept in registers.
//
// QC (a quasi-cyclic code with MAXV row blocks; qc = its block maps, as in
// DevGraph::qc_tab): variable v = cc*P + y meets, in row block r, check
// r*P + x with x = ainv*(y - b) mod P, at position cc of that check, so its
// rows are computed instead of read from the offset and row tables.
template <int MAXV, bool PAIR, bool QC = false>
__device__ int var_update(const DevGraph &g, int v, const M4 &prior, double damping,
                          double *const *v2c, const double *const *c2v, double *tot,
                          const int *qc = nullptr) {
    constexpr bool kBatch = MAXV <= 4 || QC;
    const M4 ident{{1.0, 1.0, 1.0, 1.0}};
    constexpr int NB = kBatch ? MAXV : 1;
    int b[2], deg[2];
    int row[2][NB];
    double ca[2][NB], cb[2][NB];
    if (QC) {
        const int P = g.qc_P, JL = MAXV * g.qc_L;
        const float inv_p = 1.0f / static_cast<float>(P);
        const int cc = div_small(v, P, inv_p), y = v - cc * P;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            b[s] = 0;
            deg[s] = MAXV;
            const int *t = qc + 3 * s * JL;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                const int blk = r * g.qc_L + cc;
                const int u = t[2 * JL + blk] * (y - t[JL + blk] + P);  // ainv * (y - b + P)
                const int x = u - div_small(u, P, inv_p) * P;
                row[s][r] = cc * g.side[s].m + r * P + x;
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            b[s] = g.side[s].var_off[v];
            deg[s] = g.side[s].var_off[v + 1] - b[s];
        }
    }
    if (kBatch) {
        if (!QC) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int i = 0; i < NB; ++i) row[s][i] = g.side[s].vrow[i < deg[s] ? b[s] + i : 0];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const double2 m = *reinterpret_cast<const double2 *>(c2v[s] + 2 * static_cast<size_t>(row[s][i]));
                ca[s][i] = m.x;
                cb[s][i] = m.y;
            }
    }
    // row and incoming message of position i on side s
    auto erow = [&](int s, int i) -> int { return kBatch ? row[s][kBatch ? i : 0] : g.side[s].vrow[b[s] + i]; };
    auto msg = [&](int s, int i) -> M4 {
        if (kBatch) return c2v_pair(ca[s][kBatch ? i : 0], cb[s][kBatch ? i : 0], s);
        return c2v_msg(c2v[s], erow(s, i), s);
    };
    M4 pref[2][MAXV + 1];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        pref[s][0] = ident;
#pragma unroll
        for (int i = 0; i < MAXV; ++i)
            if (i < deg[s]) pref[s][i + 1] = mul4(pref[s][i], msg(s, i));
    }
    M4 allx = ident, allz = ident;
#pragma unroll
    for (int i = 0; i <= MAXV; ++i) {
        if (i == deg[0]) allx = pref[0][i];
        if (i == deg[1]) allz = pref[1][i];
    }
    M4 total = mul4(prior, mul4(allx, allz));
    normalize4(total);
    int best = 0;
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (total.v[k] > total.v[best]) best = k;
    if (tot) {
        double *t = tot + 4 * static_cast<size_t>(v);
        t[0] = total.v[0]; t[1] = total.v[1]; t[2] = total.v[2]; t[3] = total.v[3];
    }

    // X edges: ((prior (*) x_excl) (*) z_all)
    M4 suf = ident;
#pragma unroll
    for (int i = MAXV - 1; i >= 0; --i) {
        if (i < deg[0]) {
            const int e = erow(0, i);
            M4 out = mul4(mul4(prior, mul4(pref[0][i], suf)), allz);
            normalize4(out);
            if (!PAIR && damping > 0.0) {
                const double *slot = v2c[0] + 4 * static_cast<size_t>(e);
                M4 old{{slot[0], slot[1], slot[2], slot[3]}};
                out = mix4(out, old, damping);
            }
            store_v2c<PAIR>(v2c[0], static_cast<size_t>(e), 0, out);
            suf = mul4(suf, msg(0, i));
        }
    }
    // Z edges: ((prior (*) x_all) (*) z_excl)
    suf = ident;
    const M4 px = mul4(prior, allx);
#pragma unroll
    for (int i = MAXV - 1; i >= 0; --i) {
        if (i < deg[1]) {
            const int e = erow(1, i);
            M4 out = mul4(px, mul4(pref[1][i], suf));
            normalize4(out);
            if (!PAIR && damping > 0.0) {
                const double *slot = v2c[1] + 4 * static_cast<size_t>(e);
                M4 old{{slot[0], slot[1], slot[2], slot[3]}};
                out = mix4(out, old, damping);
            }
            store_v2c<PAIR>(v2c[1], static_cast<size_t>(e), 1, out);
            suf = mul4(suf, msg(1, i));
        }
    }
    return best;
}

// var_update's arithmetic for one variable of degree MAXV on both sides with
// pair sums, on messages already loaded (same operands, same order):
// ca/cb[s][i] is the incoming pair of position i of side s; o[s][i] receives
// the outgoing pair sums and total the normalised marginal.
template <int MAXV>
__device__ __forceinline__ int var_core_pair(const M4 &prior, const double (&ca)[2][MAXV], const double (&cb)[2][MAXV],
                                             double2 (&o)[2][MAXV], M4 &total) {
    const M4 ident{{1.0, 1.0, 1.0, 1.0}};
    M4 pref[2][MAXV + 1];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        pref[s][0] = ident;
#pragma unroll
        for (int i = 0; i < MAXV; ++i) pref[s][i + 1] = mul4(pref[s][i], c2v_pair(ca[s][i], cb[s][i], s));
    }
    const M4 allx = pref[0][MAXV], allz = pref[1][MAXV];
    total = mul4(prior, mul4(allx, allz));
    normalize4(total);
    int best = 0;
    double bv = total.v[0];
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (total.v[k] > bv) {
            best = k;
            bv = total.v[k];
        }
    M4 suf = ident;
#pragma unroll
    for (int i = MAXV - 1; i >= 0; --i) {
        M4 out = mul4(mul4(prior, mul4(pref[0][i], suf)), allz);
        normalize4(out);
        pair_sums(out.v, 0, o[0][i].x, o[0][i].y);
        suf = mul4(suf, c2v_pair(ca[0][i], cb[0][i], 0));
    }
    suf = ident;
    const M4 px = mul4(prior, allx);
#pragma unroll
    for (int i = MAXV - 1; i >= 0; --i) {
        M4 out = mul4(px, mul4(pref[1][i], suf));
        normalize4(out);
        pair_sums(out.v, 1, o[1][i].x, o[1][i].y);
        suf = mul4(suf, c2v_pair(ca[1][i], cb[1][i], 1));
    }
    return best;
}

// Copies column block cc's variable-to-check rows from LDS to v2c, each
// thread its own rows (tid + j * blockDim.x); cc < 0: nothing.
__device__ __forceinline__ void var_pass_flush(const DevGraph &g, int cc, double *const *v2c, const double2 *lds) {
    if (cc < 0) return;
    const int m = g.side[0].m, rows = 2 * m;
    const size_t base = static_cast<size_t>(cc) * m;
    double2 *out0 = reinterpret_cast<double2 *>(v2c[0]) + base;
    double2 *out1 = reinterpret_cast<double2 *>(v2c[1]) + base;
    for (int k = threadIdx.x; k < rows; k += blockDim.x) {
        const double2 mv = lds[k];
        if (k < m) out0[k] = mv;
        else out1[k - m] = mv;
    }
}

// Bytes of LDS var_pass_qc_lds needs: both sides' rows of one column block.
inline size_t var_pass_lds_bytes(int m) { return 2 * static_cast<size_t>(m) * sizeof(double2); }

// Variable pass of a quasi-cyclic code (MAXV row blocks of size P, m = MAXV*P
// checks per side) with pair sums, staged through LDS.  The variables of
// column block cc meet every check at position cc, so all their messages
// lie in rows cc*m + [0, m) of each side, the row of check r*P + x being
// cc*m + r*P + x.  The workgroup copies those rows of c2v into LDS with
// coalesced loads, updates the block's P variables against LDS (each
// outgoing pair overwrites the incoming pair of the same edge, which no
// other variable touches), and copies the rows out to v2c.  lds holds 2*m
// pairs (side 0 then side 1).  visit(v, best, total) receives each
// variable's hard decision and normalised marginal.  The caller synchronises
// the workgroup before (c2v complete) and after (v2c complete).  Workgroup
// sizes are multiples of 64.
//
// res: on entry the column block already in LDS (its check messages, as the
// check pass left them) or -1; on exit the block left there (its
// variable-to-check messages, not copied out: the next check pass uses them
// in LDS, or var_pass_flush writes them back).  The blocks are walked from
// the resident end (ascending from 0, descending from L - 1), so each pass
// starts with the block the previous one ended on.
template <int MAXV, class Visit>
__device__ void var_pass_qc_lds(const DevGraph &g, const M4 &prior, double *const *v2c, const double *const *c2v,
                                const int *qc, double2 *lds, int &res, Visit visit) {
    const int tid = threadIdx.x, bs = blockDim.x;
    const int P = g.qc_P, L = g.qc_L, JL = MAXV * L, m = g.side[0].m, rows = 2 * m;
    const float inv_p = 1.0f / static_cast<float>(P);
    // LDS row k of column block cc from c2v: 16-byte loads written to LDS by
    // the load itself (lane l of a wavefront writes at the wavefront's LDS
    // base + 16*l, so thread tid owns LDS rows tid + j*bs, as in the copy out)
    auto fetch = [&](int cc) {
        const size_t base = static_cast<size_t>(cc) * m;
        for (int k = tid; k < rows; k += bs) {
            const double2 *src = k < m ? c2v_rows(c2v[0]) + base + k : c2v_rows(c2v[1]) + base + (k - m);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                             (__attribute__((address_space(3))) void *)(lds + (k & ~63)), 16, 0, 0);
        }
    };
    const bool down = L > 1 && res == L - 1;
    const int first = down ? L - 1 : 0, dir = down ? -1 : 1;
    if (res != first) fetch(first);
    for (int k = 0; k < L; ++k) {
        const int cc = first + dir * k;
        __builtin_amdgcn_s_waitcnt(0);  // this thread's LDS loads have landed
        __syncthreads();
        for (int y = tid; y < P; y += bs) {
            int idx[2][MAXV];
            double ca[2][MAXV], cb[2][MAXV];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int *t = qc + 3 * s * JL;
#pragma unroll
                for (int r = 0; r < MAXV; ++r) {
                    const int blk = r * L + cc;
                    const int u = t[2 * JL + blk] * (y - t[JL + blk] + P);  // ainv * (y - b + P)
                    idx[s][r] = s * m + r * P + (u - div_small(u, P, inv_p) * P);
                }
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int r = 0; r < MAXV; ++r) {
                    const double2 mv = lds[idx[s][r]];
                    ca[s][r] = mv.x;
                    cb[s][r] = mv.y;
                }
            double2 o[2][MAXV];
            M4 total;
            const int best = var_core_pair<MAXV>(prior, ca, cb, o, total);
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int r = 0; r < MAXV; ++r) lds[idx[s][r]] = o[s][r];
            visit(cc * P + y, best, total);
        }
        __syncthreads();
        if (k + 1 == L) break;  // the last block stays resident
        // copy out (each thread its own LDS rows), then the next block's rows
        // in over the same LDS rows once the reads have returned
        var_pass_flush(g, cc, v2c, lds);
        __builtin_amdgcn_s_waitcnt(0);
        fetch(cc + dir);
    }
    res = first + dir * (L - 1);
}

// Parity of check c of one side over est (X checks read the X bit, Z checks
// the Z bit), its variables' states loaded together.
template <int MAXC>
__device__ __forceinline__ int check_parity(const DevSide &sd, int side, int c, const uint8_t *est) {
    const int b = sd.chk_off[c], deg = sd.chk_off[c + 1] - b;
    int st[MAXC];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) st[i] = est[sd.edge_var[i < deg ? b + i : 0]];
    int par = 0;
#pragma unroll
    for (int i = 0; i < MAXC; ++i)
        if (i < deg) par ^= side == 0 ? (st[i] & 1) : (st[i] >> 1);
    return par;
}

}  // namespace f64
}  // namespace jbp
