// Learned (neural-weighted) joint BP: training kernels.
//
// No counterpart in the reference program.  The decoder extends the
// reference's joint BP (jointbp_ets.cpp:2373-2519) in the log domain with a
// weight per (side, permutation block, iteration) on every variable-to-check
// message and on every check-to-variable message, and a 2-16-1 MLP that
// damps the check-to-variable messages from (t / T, unsatisfied-check
// fraction of the estimate two iterations back):
//   mu_e      = wv[g][b][t] * (S_g(v) - lambda_e + F_g(S_other(v)))
//   lambda_e  = (1 - d_t) * wc[g][b][t] * CN_e(mu) + d_t * lambda_e(t-1)
// Loss: mean binary cross-entropy of the per-qubit X / Z marginals against
// the sampled error, averaged over iterations 1..T.
//
// One workgroup per trial (persistent over the batch).  The forward pass
// stores every iteration's check messages (the tape, T+1 x E floats per
// trial: 5.5 MB at T = 25, 23 GB for a 4096-trial batch, resident in HBM);
// the backward pass walks the tape in reverse, recomputing each iteration's
// variable and check pass from it, and accumulates the weight gradients in
// LDS before one global atomic per block and iteration.
//
// Check-node adjoint in O(deg): with t_k = tanh(mu_k / 2) and
// h_e = dL/dlambda_e * sign * cosh^2(lambda_e / 2),
//   dL/dmu_j = sech^2(mu_j / 2) * dQ/dt_j,   Q = sum_e h_e prod_{k != e} t_k,
// and all dQ/dt_j come from one prefix and one suffix recursion.
//
// KIND = 1 is the normalised min-sum form of the same model (the
// production learned decoder): CN_e = (-1)^syn prod_{k != e} sign(mu_k) *
// min_{k != e} |mu_k| and F(L) = max(ln a + L, ln b) - max(ln c + L, ln d),
// hard decisions from the signs of the X / Z marginals.  Its check adjoint
// routes each edge's gradient to the input that attains its minimum: the
// overall arg-minimum i1 for every edge but i1 itself, the second i2 for i1.
#include "jbp_common.h"
#include "jbp_learned.h"

namespace jbp {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDeg = 16;

template <typename R> __device__ __forceinline__ R dexp(R x);
template <> __device__ __forceinline__ float dexp(float x) { return expf(x); }
template <> __device__ __forceinline__ double dexp(double x) { return exp(x); }
template <typename R> __device__ __forceinline__ R dlog(R x);
template <> __device__ __forceinline__ float dlog(float x) { return logf(x); }
template <> __device__ __forceinline__ double dlog(double x) { return log(x); }
template <typename R> __device__ __forceinline__ R dtanh(R x);
template <> __device__ __forceinline__ float dtanh(float x) { return tanhf(x); }
template <> __device__ __forceinline__ double dtanh(double x) { return tanh(x); }
template <typename R> __device__ __forceinline__ R dcosh(R x);
template <> __device__ __forceinline__ float dcosh(float x) { return coshf(x); }
template <> __device__ __forceinline__ double dcosh(double x) { return cosh(x); }

template <typename R>
__device__ __forceinline__ R sigm(R x) {
    return x >= R(0) ? R(1) / (R(1) + dexp(-x)) : dexp(x) / (R(1) + dexp(x));
}
template <typename R>
__device__ __forceinline__ R softplus(R x) {
    return x > R(0) ? x + dlog(R(1) + dexp(-x)) : dlog(R(1) + dexp(x));
}
// F(L) = ln((k0 e^L + k1) / (k2 e^L + k3)) and F'(L)
template <typename R>
__device__ __forceinline__ R fterm(R L, const R *k) {
    if (L >= R(0)) return dlog(k[0] + k[1] * dexp(-L)) - dlog(k[2] + k[3] * dexp(-L));
    return dlog(k[0] * dexp(L) + k[1]) - dlog(k[2] * dexp(L) + k[3]);
}
template <typename R>
__device__ __forceinline__ R fterm_d(R L, const R *k) {
    return sigm(L + dlog(k[0] / k[1])) - sigm(L + dlog(k[2] / k[3]));
}

template <typename R>
struct Model {
    const R *wc, *wv, *mlp;  // [2][nb][T], [2][nb][T], [65]
    const int *blk0, *blk1;  // block of each edge, per side
    int nb, T, use_mlp;
    int lpb;                 // layered kind: blocks per layer
};

// Channel constants of one trial: (a, b, c, d) of each side's joint term and
// their logarithms (the min-sum form uses the logarithms).
template <typename R>
struct Chan {
    R fx[4], fz[4], lfx[4], lfz[4], lp[4];
};

// KIND 0: F(L) = ln((a e^L + b) / (c e^L + d)); KIND 1: its max-log form.
template <int KIND, typename R>
__device__ __forceinline__ R fterm_k(R L, const R *k, const R *lk) {
    if (KIND == 1) {
        const R u = lk[0] + L, v = lk[2] + L;
        return (u > lk[1] ? u : lk[1]) - (v > lk[3] ? v : lk[3]);
    }
    return fterm(L, k);
}
template <int KIND, typename R>
__device__ __forceinline__ R fterm_kd(R L, const R *k, const R *lk) {
    if (KIND == 1) return (lk[0] + L > lk[1] ? R(1) : R(0)) - (lk[2] + L > lk[3] ? R(1) : R(0));
    return fterm_d(L, k);
}

template <typename R>
struct Tape {
    R *act;             // [batch][T+1][Et]
    R *scratch;         // per slot
    size_t slot_stride; // elements
    R *tape_d;          // [batch][T]
    R *tape_u;          // [batch][T]
    const double *prior;// [batch][4]
    R *loss;            // [1]
    R *gwc, *gwv, *gmlp;
    int *conv;          // [batch]
};

template <typename R>
__device__ R mlp_forward(const R *mlp, R x0, R x1, R *h) {
    R z = mlp[kMlpHidden * 4];
    for (int i = 0; i < kMlpHidden; ++i) {
        h[i] = dtanh(mlp[2 * i] * x0 + mlp[2 * i + 1] * x1 + mlp[2 * kMlpHidden + i]);
        z += mlp[3 * kMlpHidden + i] * h[i];
    }
    return sigm(z);
}

// Block-wide sum through one LDS cell (the caller zeroes it and syncs).
template <typename R>
__device__ __forceinline__ void block_add(R *cell, R v) {
    atomicAdd(cell, v);
}

// Forward check node of one check: raw and clamped outputs (natural log).
template <typename R>
__device__ void check_node(const R *mu, int deg, int syn, R *raw, R *out) {
    const R lmax = R(64.0 * 0.69314718055994530942);
    R e[kMaxDeg], pe[kMaxDeg + 1], po[kMaxDeg + 1];
    uint32_t neg = 0, negall = static_cast<uint32_t>(syn & 1);
    pe[0] = R(1);
    po[0] = R(0);
#pragma unroll
    for (int i = 0; i < kMaxDeg; ++i) {
        if (i < deg) {
            R m = mu[i];
            uint32_t ng = m < R(0) ? 1u : 0u;
            neg |= ng << i;
            negall ^= ng;
            e[i] = dexp(-fabs(m));
            pe[i + 1] = pe[i] + po[i] * e[i];
            po[i + 1] = pe[i] * e[i] + po[i];
        }
    }
    R se = R(1), so = R(0);
#pragma unroll
    for (int i = kMaxDeg - 1; i >= 0; --i) {
        if (i < deg) {
            R PE = pe[i] * se + po[i] * so;
            R PO = pe[i] * so + po[i] * se;
            R mag = dlog(PE) - dlog(PO);
            R sg = ((negall ^ (neg >> i)) & 1u) ? R(-1) : R(1);
            raw[i] = sg * mag;
            out[i] = sg * (mag < lmax ? mag : lmax);
            R ne = se + so * e[i], no = se * e[i] + so;
            se = ne;
            so = no;
        }
    }
}

// Min-sum check node of one check: raw (unclamped) and clamped outputs.
template <typename R>
__device__ void check_node_ms(const R *mu, int deg, int syn, R *raw, R *out) {
    const R lmax = R(64.0 * 0.69314718055994530942);
    R m1 = R(HUGE_VAL), m2 = R(HUGE_VAL);
    int i1 = -1;
    uint32_t neg = 0, negall = static_cast<uint32_t>(syn & 1);
    for (int i = 0; i < deg; ++i) {
        const R m = mu[i];
        const R a = fabs(m);
        const uint32_t ng = m < R(0) ? 1u : 0u;
        neg |= ng << i;
        negall ^= ng;
        if (a < m1) {
            m2 = m1;
            m1 = a;
            i1 = i;
        } else if (a < m2) {
            m2 = a;
        }
    }
    for (int i = 0; i < deg; ++i) {
        const R mag = i == i1 ? m2 : m1;
        const R sg = ((negall ^ (neg >> i)) & 1u) ? R(-1) : R(1);
        raw[i] = sg * mag;
        out[i] = sg * (mag < lmax ? mag : lmax);
    }
}

template <typename R>
struct Slot {
    R *mu, *a, *raw, *out, *gc, *gp, *S;  // Et, Et, Et, Et, Et, Et, 2n
    int *est;                              // n
};

template <typename R>
__device__ Slot<R> slot_of(const DevGraph &g, const Tape<R> &tb) {
    const size_t Et = static_cast<size_t>(g.side[0].E) + g.side[1].E;
    R *base = tb.scratch + static_cast<size_t>(blockIdx.x) * tb.slot_stride;
    Slot<R> s;
    s.mu = base;
    s.a = base + Et;
    s.raw = base + 2 * Et;
    s.out = base + 3 * Et;
    s.gc = base + 4 * Et;
    s.gp = base + 5 * Et;
    s.S = base + 6 * Et;
    s.est = reinterpret_cast<int *>(s.S + 2 * static_cast<size_t>(g.n));
    return s;
}

template <typename R>
__device__ void priors(const double *p4, Chan<R> &ch) {
    const int ox[4] = {0, 2, 1, 3};
    for (int i = 0; i < 4; ++i) {
        ch.fx[i] = R(p4[ox[i]]);
        ch.fz[i] = R(p4[i]);
        ch.lfx[i] = R(log(p4[ox[i]]));
        ch.lfz[i] = R(log(p4[i]));
        ch.lp[i] = R(log(p4[i]));
    }
}

__device__ __forceinline__ int syn_bit(const DevGraph &g, const TrialPlanes &tp, int b, int s, int c) {
    const uint32_t *pl = (s == 0 ? tp.syn_x : tp.syn_z) + static_cast<size_t>(b) * g.wm[s];
    return (pl[c >> 5] >> (c & 31)) & 1;
}

__device__ __forceinline__ int err_state(const DevGraph &g, const TrialPlanes &tp, int b, int v) {
    size_t w = static_cast<size_t>(b) * g.wn + (v >> 5);
    return static_cast<int>(((tp.err_x[w] >> (v & 31)) & 1u) | (((tp.err_z[w] >> (v & 31)) & 1u) << 1));
}

// Variable pass of iteration t from lambda(t-1) = prev: S sums, then a and mu.
template <int KIND, typename R>
__device__ void var_pass(const DevGraph &g, const Model<R> &M, int t, const R *prev,
                         const Chan<R> &ch, Slot<R> &sl) {
    const int n = g.n, Ex = g.side[0].E;
    for (int v = threadIdx.x; v < n; v += kBlock) {
        R sx = R(0), sz = R(0);
        for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += prev[g.side[0].var_edge[k]];
        for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += prev[Ex + g.side[1].var_edge[k]];
        sl.S[v] = sx;
        sl.S[n + v] = sz;
        const R f0 = fterm_k<KIND>(sz, ch.fx, ch.lfx), f1 = fterm_k<KIND>(sx, ch.fz, ch.lfz);
        for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) {
            int e = g.side[0].var_edge[k];
            R a = sx - prev[e] + f0;
            sl.a[e] = a;
            sl.mu[e] = M.wv[(M.blk0[e]) * M.T + (t - 1)] * a;
        }
        for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) {
            int e = g.side[1].var_edge[k];
            R a = sz - prev[Ex + e] + f1;
            sl.a[Ex + e] = a;
            sl.mu[Ex + e] = M.wv[(M.nb + M.blk1[e]) * M.T + (t - 1)] * a;
        }
    }
}

// Check pass: raw and clamped check outputs of every edge from mu.
template <int KIND, typename R>
__device__ void check_pass(const DevGraph &g, const TrialPlanes &tp, int b, Slot<R> &sl) {
    const int mx = g.side[0].m, mz = g.side[1].m, Ex = g.side[0].E;
    for (int c = threadIdx.x; c < mx + mz; c += kBlock) {
        const int s = c < mx ? 0 : 1;
        const int cc = c < mx ? c : c - mx;
        const int base = s == 0 ? 0 : Ex;
        const int off = g.side[s].chk_off[cc], deg = g.side[s].chk_off[cc + 1] - off;
        if (KIND == 1)
            check_node_ms(sl.mu + base + off, deg, syn_bit(g, tp, b, s, cc), sl.raw + base + off,
                          sl.out + base + off);
        else
            check_node(sl.mu + base + off, deg, syn_bit(g, tp, b, s, cc), sl.raw + base + off,
                       sl.out + base + off);
    }
}

template <typename R, int KIND>
__global__ void __launch_bounds__(kBlock) learned_fwd_kernel(DevGraph g, Model<R> M, TrialPlanes tp,
                                                            Tape<R> tb, int batch, int *counter) {
    __shared__ int s_trial;
    __shared__ R s_d;
    __shared__ R s_loss;
    __shared__ int s_cnt;
    __shared__ R s_ubar[kMaxLearnedT + 1];
    const int T = M.T, n = g.n;
    const int Ex = g.side[0].E;
    const size_t Et = static_cast<size_t>(Ex) + g.side[1].E;
    const int mx = g.side[0].m, mz = g.side[1].m;
    const R mtot = R(mx + mz);
    const R scale = R(1) / (R(batch) * R(T) * R(2 * n));
    Slot<R> sl = slot_of(g, tb);
    for (;;) {
        if (threadIdx.x == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = s_trial;
        __syncthreads();
        if (b >= batch) return;
        Chan<R> ch;
        priors(tb.prior + 4 * static_cast<size_t>(b), ch);
        R *lam = tb.act + static_cast<size_t>(b) * (T + 1) * Et;
        for (size_t e = threadIdx.x; e < Et; e += kBlock) lam[e] = R(0);
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_loss = R(0);
        }
        __syncthreads();
        int local = 0;
        for (int w = threadIdx.x; w < g.wm[0]; w += kBlock) local += __popc(tp.syn_x[static_cast<size_t>(b) * g.wm[0] + w]);
        for (int w = threadIdx.x; w < g.wm[1]; w += kBlock) local += __popc(tp.syn_z[static_cast<size_t>(b) * g.wm[1] + w]);
        atomicAdd(&s_cnt, local);
        __syncthreads();
        if (threadIdx.x == 0) s_ubar[0] = R(s_cnt) / mtot;
        int conv = 0;
        R loss_local = R(0);
        for (int t = 1; t <= T; ++t) {
            const R *prev = lam + static_cast<size_t>(t - 1) * Et;
            R *cur = lam + static_cast<size_t>(t) * Et;
            var_pass<KIND>(g, M, t, prev, ch, sl);
            if (threadIdx.x == 0) {
                R h[kMlpHidden];
                R x0 = R(t) / R(T), x1 = s_ubar[t >= 2 ? t - 2 : 0];
                R d = M.use_mlp ? mlp_forward(M.mlp, x0, x1, h) : R(0);
                s_d = d;
                tb.tape_d[static_cast<size_t>(b) * T + t - 1] = d;
                tb.tape_u[static_cast<size_t>(b) * T + t - 1] = x1;
                s_cnt = 0;
            }
            __syncthreads();
            check_pass<KIND>(g, tp, b, sl);
            __syncthreads();
            const R d = s_d;
            for (size_t e = threadIdx.x; e < Et; e += kBlock) {
                const int blk = e < static_cast<size_t>(Ex) ? M.blk0[e] : M.nb + M.blk1[e - Ex];
                cur[e] = (R(1) - d) * M.wc[blk * T + t - 1] * sl.out[e] + d * prev[e];
            }
            __syncthreads();
            for (int v = threadIdx.x; v < n; v += kBlock) {
                R sx = R(0), sz = R(0);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += cur[g.side[0].var_edge[k]];
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += cur[Ex + g.side[1].var_edge[k]];
                const R lx = sx + fterm_k<KIND>(sz, ch.fx, ch.lfx);
                const R lz = sz + fterm_k<KIND>(sx, ch.fz, ch.lfz);
                const int st = err_state(g, tp, b, v);
                loss_local += softplus(R(2 * (st & 1) - 1) * lx) + softplus(R(2 * (st >> 1) - 1) * lz);
                if (KIND == 1) {  // per-bit decisions from the marginals' signs
                    sl.est[v] = (lx < R(0) ? 1 : 0) | (lz < R(0) ? 2 : 0);
                    continue;
                }
                R l0 = ch.lp[0] + sx + sz, l1 = ch.lp[1] + sz, l2 = ch.lp[2] + sx, l3 = ch.lp[3];
                int best = 0;
                R bv = l0;
                if (l1 > bv) { best = 1; bv = l1; }
                if (l2 > bv) { best = 2; bv = l2; }
                if (l3 > bv) { best = 3; }
                sl.est[v] = best;
            }
            __syncthreads();
            int mis = 0;
            for (int c = threadIdx.x; c < mx + mz; c += kBlock) {
                const int s = c < mx ? 0 : 1;
                const int cc = c < mx ? c : c - mx;
                int par = 0;
                for (int e = g.side[s].chk_off[cc]; e < g.side[s].chk_off[cc + 1]; ++e) {
                    int st = sl.est[g.side[s].edge_var[e]];
                    par ^= s == 0 ? (st & 1) : (st >> 1);
                }
                mis += par != syn_bit(g, tp, b, s, cc);
            }
            atomicAdd(&s_cnt, mis);
            __syncthreads();
            if (threadIdx.x == 0) s_ubar[t] = R(s_cnt) / mtot;
            if (s_cnt == 0 && conv == 0) conv = t;
            __syncthreads();
        }
        block_add(&s_loss, loss_local);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(tb.loss, s_loss * scale);
            tb.conv[b] = conv;
        }
        __syncthreads();
    }
}

template <typename R, int KIND>
__global__ void __launch_bounds__(kBlock) learned_bwd_kernel(DevGraph g, Model<R> M, TrialPlanes tp,
                                                            Tape<R> tb, int batch, int *counter) {
    __shared__ int s_trial;
    __shared__ R s_gd;
    __shared__ R s_gw[4 * kMaxLearnedBlocks];  // wc side0 | wc side1 | wv side0 | wv side1
    const int T = M.T, n = g.n, nb = M.nb;
    const int Ex = g.side[0].E;
    const size_t Et = static_cast<size_t>(Ex) + g.side[1].E;
    const int mx = g.side[0].m, mz = g.side[1].m;
    const R scale = R(1) / (R(batch) * R(T) * R(2 * n));
    const R lmax = R(64.0 * 0.69314718055994530942);
    Slot<R> sl = slot_of(g, tb);
    for (;;) {
        if (threadIdx.x == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = s_trial;
        __syncthreads();
        if (b >= batch) return;
        Chan<R> ch;
        priors(tb.prior + 4 * static_cast<size_t>(b), ch);
        const R *lam = tb.act + static_cast<size_t>(b) * (T + 1) * Et;
        R *gc = sl.gc, *gp = sl.gp;
        for (size_t e = threadIdx.x; e < Et; e += kBlock) gc[e] = R(0);
        for (int i = threadIdx.x; i < 4 * nb; i += kBlock) s_gw[i] = R(0);
        __syncthreads();
        for (int t = T; t >= 1; --t) {
            const R *prev = lam + static_cast<size_t>(t - 1) * Et;
            const R *cur = lam + static_cast<size_t>(t) * Et;
            // (1) loss term at iteration t
            for (int v = threadIdx.x; v < n; v += kBlock) {
                R sx = R(0), sz = R(0);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += cur[g.side[0].var_edge[k]];
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += cur[Ex + g.side[1].var_edge[k]];
                const R lx = sx + fterm_k<KIND>(sz, ch.fx, ch.lfx);
                const R lz = sz + fterm_k<KIND>(sx, ch.fz, ch.lfz);
                const int st = err_state(g, tp, b, v);
                const R kx = R(2 * (st & 1) - 1), kz = R(2 * (st >> 1) - 1);
                const R glx = scale * kx * sigm(kx * lx), glz = scale * kz * sigm(kz * lz);
                const R gsx = glx + glz * fterm_kd<KIND>(sx, ch.fz, ch.lfz);
                const R gsz = glz + glx * fterm_kd<KIND>(sz, ch.fx, ch.lfx);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) gc[g.side[0].var_edge[k]] += gsx;
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) gc[Ex + g.side[1].var_edge[k]] += gsz;
            }
            // (2) recompute the variable pass, (3) the check pass
            var_pass<KIND>(g, M, t, prev, ch, sl);
            if (threadIdx.x == 0) s_gd = R(0);
            __syncthreads();
            check_pass<KIND>(g, tp, b, sl);
            __syncthreads();
            // (4) damping and check weights
            const R d = tb.tape_d[static_cast<size_t>(b) * T + t - 1];
            R gd_local = R(0);
            for (size_t e = threadIdx.x; e < Et; e += kBlock) {
                const bool zs = e >= static_cast<size_t>(Ex);
                const int blk = zs ? M.blk1[e - Ex] : M.blk0[e];
                const R w = M.wc[((zs ? nb : 0) + blk) * T + t - 1];
                const R ge = gc[e];
                gd_local += ge * (prev[e] - w * sl.out[e]);
                gp[e] = d * ge;
                const R gtil = (R(1) - d) * ge;
                atomicAdd(&s_gw[(zs ? nb : 0) + blk], gtil * sl.out[e]);
                gc[e] = fabs(sl.raw[e]) > lmax ? R(0) : w * gtil;
            }
            atomicAdd(&s_gd, gd_local);
            __syncthreads();
            if (threadIdx.x == 0 && M.use_mlp) {
                R h[kMlpHidden];
                const R x0 = R(t) / R(T), x1 = tb.tape_u[static_cast<size_t>(b) * T + t - 1];
                mlp_forward(M.mlp, x0, x1, h);
                const R gz = s_gd * d * (R(1) - d);
                atomicAdd(&tb.gmlp[4 * kMlpHidden], gz);
                for (int i = 0; i < kMlpHidden; ++i) {
                    atomicAdd(&tb.gmlp[3 * kMlpHidden + i], gz * h[i]);
                    const R ga = gz * M.mlp[3 * kMlpHidden + i] * (R(1) - h[i] * h[i]);
                    atomicAdd(&tb.gmlp[2 * i], ga * x0);
                    atomicAdd(&tb.gmlp[2 * i + 1], ga * x1);
                    atomicAdd(&tb.gmlp[2 * kMlpHidden + i], ga);
                }
            }
            // (5) check-node adjoint: gc (= dL/dlambda-hat) -> dL/dmu into out
            for (int c = threadIdx.x; KIND == 1 && c < mx + mz; c += kBlock) {
                const int s = c < mx ? 0 : 1;
                const int cc = c < mx ? c : c - mx;
                const int base = (s == 0 ? 0 : Ex) + g.side[s].chk_off[cc];
                const int deg = g.side[s].chk_off[cc + 1] - g.side[s].chk_off[cc];
                // arg-minima of |mu| (lowest index among equals)
                int i1 = 0;
                for (int i = 1; i < deg; ++i)
                    if (fabs(sl.mu[base + i]) < fabs(sl.mu[base + i1])) i1 = i;
                int i2 = i1 == 0 ? 1 : 0;
                for (int i = 0; i < deg; ++i)
                    if (i != i1 && fabs(sl.mu[base + i]) < fabs(sl.mu[base + i2])) i2 = i;
                R to1 = R(0), to2 = R(0);
                for (int i = 0; i < deg; ++i) {
                    const R ge = gc[base + i];
                    const R se = sl.raw[base + i] < R(0) ? -ge : ge;
                    if (i == i1) to2 += se;
                    else to1 += se;
                }
                for (int i = 0; i < deg; ++i) sl.out[base + i] = R(0);
                sl.out[base + i1] = sl.mu[base + i1] < R(0) ? -to1 : to1;
                if (deg > 1) sl.out[base + i2] = sl.mu[base + i2] < R(0) ? -to2 : to2;
            }
            for (int c = threadIdx.x; KIND == 0 && c < mx + mz; c += kBlock) {
                const int s = c < mx ? 0 : 1;
                const int cc = c < mx ? c : c - mx;
                const int base = (s == 0 ? 0 : Ex) + g.side[s].chk_off[cc];
                const int deg = g.side[s].chk_off[cc + 1] - g.side[s].chk_off[cc];
                const R sgn = syn_bit(g, tp, b, s, cc) ? R(-1) : R(1);
                R tk[kMaxDeg], hk[kMaxDeg], A[kMaxDeg + 1], Bq[kMaxDeg + 1];
                A[0] = R(1);
                Bq[0] = R(0);
#pragma unroll
                for (int i = 0; i < kMaxDeg; ++i) {
                    if (i < deg) {
                        tk[i] = dtanh(sl.mu[base + i] * R(0.5));
                        const R ch = dcosh(sl.raw[base + i] * R(0.5));
                        const R ge = gc[base + i];
                        hk[i] = ge == R(0) ? R(0) : ge * sgn * ch * ch;
                        A[i + 1] = A[i] * tk[i];
                        Bq[i + 1] = Bq[i] * tk[i] + hk[i] * A[i];
                    }
                }
                R SA = R(1), SB = R(0);
#pragma unroll
                for (int i = kMaxDeg - 1; i >= 0; --i) {
                    if (i < deg) {
                        const R dq = Bq[i] * SA + A[i] * SB;
                        const R cj = dcosh(sl.mu[base + i] * R(0.5));
                        sl.out[base + i] = dq / (cj * cj);
                        const R nsb = SB * tk[i] + hk[i] * SA;
                        SA = SA * tk[i];
                        SB = nsb;
                    }
                }
            }
            __syncthreads();
            // (6) variable-pass adjoint into gp (dL/dlambda(t-1))
            for (int v = threadIdx.x; v < n; v += kBlock) {
                R gsum[2];
                for (int s = 0; s < 2; ++s) {
                    const int base = s == 0 ? 0 : Ex;
                    const int *blk = s == 0 ? M.blk0 : M.blk1;
                    R acc = R(0);
                    for (int k = g.side[s].var_off[v]; k < g.side[s].var_off[v + 1]; ++k) {
                        const int e = g.side[s].var_edge[k];
                        const int wi = (s * nb + blk[e]);
                        const R gmu = sl.out[base + e];
                        atomicAdd(&s_gw[2 * nb + wi], gmu * sl.a[base + e]);
                        const R ga = M.wv[wi * T + t - 1] * gmu;
                        acc += ga;
                        gp[base + e] -= ga;
                    }
                    gsum[s] = acc;
                }
                const R sx = sl.S[v], sz = sl.S[n + v];
                const R fdx = fterm_kd<KIND>(sz, ch.fx, ch.lfx);  // d a_x / d S_z
                const R fdz = fterm_kd<KIND>(sx, ch.fz, ch.lfz);  // d a_z / d S_x
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k)
                    gp[g.side[0].var_edge[k]] += gsum[0] + gsum[1] * fdz;
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k)
                    gp[Ex + g.side[1].var_edge[k]] += gsum[1] + gsum[0] * fdx;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < 4 * nb; i += kBlock) {
                const int which = i / (2 * nb);  // 0 wc, 1 wv
                const int wi = i % (2 * nb);
                R *dst = which == 0 ? tb.gwc : tb.gwv;
                if (s_gw[i] != R(0)) atomicAdd(&dst[wi * T + t - 1], s_gw[i]);
                s_gw[i] = R(0);
            }
            R *tmp = gc;
            gc = gp;
            gp = tmp;
            __syncthreads();
        }
    }
}

// ---- layered min-sum (kind 2) ----
// One workgroup per trial as above.  Layer of an edge: its weight block /
// M.lpb (block row r of the permutation-block code).  Forward: tape row t
// starts as a copy of row t-1 and each layer's update overwrites its edges
// in place, so row t ends as iteration t's messages.  The estimate a layer
// starts from is tested as the device decoder does: iteration t converges
// when all its layers start from the same estimate and every check accepts
// it.

template <typename R>
__device__ __forceinline__ int edge_layer(const Model<R> &M, int s, int e) {
    return (s == 0 ? M.blk0[e] : M.blk1[e]) / M.lpb;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) learned_fwd_lms_kernel(DevGraph g, Model<R> M, TrialPlanes tp,
                                                                Tape<R> tb, int batch, int *counter) {
    __shared__ int s_trial;
    __shared__ R s_loss;
    __shared__ int s_unstable, s_bad;
    const int T = M.T, n = g.n;
    const int Ex = g.side[0].E;
    const size_t Et = static_cast<size_t>(Ex) + g.side[1].E;
    const int mx = g.side[0].m, mz = g.side[1].m;
    const int nlay = M.nb / M.lpb;
    const R scale = R(1) / (R(batch) * R(T) * R(2 * n));
    Slot<R> sl = slot_of(g, tb);
    for (;;) {
        if (threadIdx.x == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = s_trial;
        __syncthreads();
        if (b >= batch) return;
        Chan<R> ch;
        priors(tb.prior + 4 * static_cast<size_t>(b), ch);
        R *lam = tb.act + static_cast<size_t>(b) * (T + 1) * Et;
        for (size_t e = threadIdx.x; e < Et; e += kBlock) lam[e] = R(0);
        if (threadIdx.x == 0) s_loss = R(0);
        __syncthreads();
        int conv = 0;
        R loss_local = R(0);
        for (int t = 1; t <= T; ++t) {
            R *cur = lam + static_cast<size_t>(t) * Et;
            const R *prev = cur - Et;
            for (size_t e = threadIdx.x; e < Et; e += kBlock) cur[e] = prev[e];
            if (threadIdx.x == 0) {
                s_unstable = 0;
                s_bad = 0;
            }
            __syncthreads();
            for (int r = 0; r < nlay; ++r) {
                // sums and the estimate this layer starts from
                int unstable = 0;
                for (int v = threadIdx.x; v < n; v += kBlock) {
                    R sx = R(0), sz = R(0);
                    for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += cur[g.side[0].var_edge[k]];
                    for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += cur[Ex + g.side[1].var_edge[k]];
                    sl.S[v] = sx;
                    sl.S[n + v] = sz;
                    const R lx = sx + fterm_k<1>(sz, ch.fx, ch.lfx), lz = sz + fterm_k<1>(sx, ch.fz, ch.lfz);
                    const int est = (lx < R(0) ? 1 : 0) | (lz < R(0) ? 2 : 0);
                    if (r == 0) sl.est[v] = est;
                    else unstable |= est != sl.est[v];
                }
                if (unstable) atomicOr(&s_unstable, 1);
                __syncthreads();
                // this layer's checks against the starting estimate (when the
                // layers' estimates differ the iteration cannot converge)
                int bad = 0;
                for (int c = threadIdx.x; c < mx + mz; c += kBlock) {
                    const int s = c < mx ? 0 : 1;
                    const int cc = c < mx ? c : c - mx;
                    const int e0 = g.side[s].chk_off[cc], e1 = g.side[s].chk_off[cc + 1];
                    if (e1 == e0 || edge_layer(M, s, e0) != r) continue;
                    int par = 0;
                    for (int e = e0; e < e1; ++e) {
                        const int st = sl.est[g.side[s].edge_var[e]];
                        par ^= s == 0 ? (st & 1) : (st >> 1);
                    }
                    bad |= par != syn_bit(g, tp, b, s, cc);
                }
                if (bad) atomicOr(&s_bad, 1);
                // mu of this layer's edges
                for (int v = threadIdx.x; v < n; v += kBlock) {
                    const R sx = sl.S[v], sz = sl.S[n + v];
                    const R f0 = fterm_k<1>(sz, ch.fx, ch.lfx), f1 = fterm_k<1>(sx, ch.fz, ch.lfz);
                    for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) {
                        const int e = g.side[0].var_edge[k];
                        if (edge_layer(M, 0, e) != r) continue;
                        const R a = sx - cur[e] + f0;
                        sl.a[e] = a;
                        sl.mu[e] = M.wv[M.blk0[e] * T + (t - 1)] * a;
                    }
                    for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) {
                        const int e = g.side[1].var_edge[k];
                        if (edge_layer(M, 1, e) != r) continue;
                        const R a = sz - cur[Ex + e] + f1;
                        sl.a[Ex + e] = a;
                        sl.mu[Ex + e] = M.wv[(M.nb + M.blk1[e]) * T + (t - 1)] * a;
                    }
                }
                __syncthreads();
                // this layer's check update, weighted into the state
                for (int c = threadIdx.x; c < mx + mz; c += kBlock) {
                    const int s = c < mx ? 0 : 1;
                    const int cc = c < mx ? c : c - mx;
                    const int off = g.side[s].chk_off[cc], deg = g.side[s].chk_off[cc + 1] - off;
                    if (deg == 0 || edge_layer(M, s, off) != r) continue;
                    const int base = s == 0 ? 0 : Ex;
                    check_node_ms(sl.mu + base + off, deg, syn_bit(g, tp, b, s, cc), sl.raw + base + off,
                                  sl.out + base + off);
                    const int *blk = s == 0 ? M.blk0 : M.blk1;
                    for (int i = 0; i < deg; ++i)
                        cur[base + off + i] = M.wc[(s * M.nb + blk[off + i]) * T + (t - 1)] * sl.out[base + off + i];
                }
                __syncthreads();
            }
            if (s_unstable == 0 && s_bad == 0 && conv == 0) conv = t;
            // loss on the marginals at the end of the iteration
            for (int v = threadIdx.x; v < n; v += kBlock) {
                R sx = R(0), sz = R(0);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += cur[g.side[0].var_edge[k]];
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += cur[Ex + g.side[1].var_edge[k]];
                const R lx = sx + fterm_k<1>(sz, ch.fx, ch.lfx);
                const R lz = sz + fterm_k<1>(sx, ch.fz, ch.lfz);
                const int st = err_state(g, tp, b, v);
                loss_local += softplus(R(2 * (st & 1) - 1) * lx) + softplus(R(2 * (st >> 1) - 1) * lz);
            }
            __syncthreads();
        }
        block_add(&s_loss, loss_local);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(tb.loss, s_loss * scale);
            tb.conv[b] = conv;
        }
        __syncthreads();
    }
}

template <typename R>
__global__ void __launch_bounds__(kBlock) learned_bwd_lms_kernel(DevGraph g, Model<R> M, TrialPlanes tp,
                                                                Tape<R> tb, int batch, int *counter) {
    __shared__ int s_trial;
    __shared__ R s_gw[4 * kMaxLearnedBlocks];  // wc side0 | wc side1 | wv side0 | wv side1
    const int T = M.T, n = g.n, nb = M.nb;
    const int Ex = g.side[0].E;
    const size_t Et = static_cast<size_t>(Ex) + g.side[1].E;
    const int mx = g.side[0].m, mz = g.side[1].m;
    const int nlay = nb / M.lpb;
    const R scale = R(1) / (R(batch) * R(T) * R(2 * n));
    const R lmax = R(64.0 * 0.69314718055994530942);
    Slot<R> sl = slot_of(g, tb);
    for (;;) {
        if (threadIdx.x == 0) s_trial = atomicAdd(counter, 1);
        __syncthreads();
        const int b = s_trial;
        __syncthreads();
        if (b >= batch) return;
        Chan<R> ch;
        priors(tb.prior + 4 * static_cast<size_t>(b), ch);
        const R *lam = tb.act + static_cast<size_t>(b) * (T + 1) * Et;
        R *gc = sl.gc;  // dL/d(messages) of the state after the step being undone
        for (size_t e = threadIdx.x; e < Et; e += kBlock) gc[e] = R(0);
        for (int i = threadIdx.x; i < 4 * nb; i += kBlock) s_gw[i] = R(0);
        __syncthreads();
        for (int t = T; t >= 1; --t) {
            const R *prev = lam + static_cast<size_t>(t - 1) * Et;
            const R *cur = lam + static_cast<size_t>(t) * Et;
            // loss term at the end of iteration t
            for (int v = threadIdx.x; v < n; v += kBlock) {
                R sx = R(0), sz = R(0);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) sx += cur[g.side[0].var_edge[k]];
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) sz += cur[Ex + g.side[1].var_edge[k]];
                const R lx = sx + fterm_k<1>(sz, ch.fx, ch.lfx);
                const R lz = sz + fterm_k<1>(sx, ch.fz, ch.lfz);
                const int st = err_state(g, tp, b, v);
                const R kx = R(2 * (st & 1) - 1), kz = R(2 * (st >> 1) - 1);
                const R glx = scale * kx * sigm(kx * lx), glz = scale * kz * sigm(kz * lz);
                const R gsx = glx + glz * fterm_kd<1>(sx, ch.fz, ch.lfz);
                const R gsz = glz + glx * fterm_kd<1>(sz, ch.fx, ch.lfx);
                for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) gc[g.side[0].var_edge[k]] += gsx;
                for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) gc[Ex + g.side[1].var_edge[k]] += gsz;
            }
            __syncthreads();
            for (int r = nlay - 1; r >= 0; --r) {
                // the state before step (t, r): iteration t's messages on the
                // layers before r, iteration t-1's on the rest; its sums, and
                // a and mu of layer r's edges
                for (int v = threadIdx.x; v < n; v += kBlock) {
                    R sx = R(0), sz = R(0);
                    for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) {
                        const int e = g.side[0].var_edge[k];
                        sx += edge_layer(M, 0, e) < r ? cur[e] : prev[e];
                    }
                    for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) {
                        const int e = g.side[1].var_edge[k];
                        sz += edge_layer(M, 1, e) < r ? cur[Ex + e] : prev[Ex + e];
                    }
                    sl.S[v] = sx;
                    sl.S[n + v] = sz;
                    const R f0 = fterm_k<1>(sz, ch.fx, ch.lfx), f1 = fterm_k<1>(sx, ch.fz, ch.lfz);
                    for (int k = g.side[0].var_off[v]; k < g.side[0].var_off[v + 1]; ++k) {
                        const int e = g.side[0].var_edge[k];
                        if (edge_layer(M, 0, e) != r) continue;
                        const R a = sx - prev[e] + f0;
                        sl.a[e] = a;
                        sl.mu[e] = M.wv[M.blk0[e] * T + (t - 1)] * a;
                    }
                    for (int k = g.side[1].var_off[v]; k < g.side[1].var_off[v + 1]; ++k) {
                        const int e = g.side[1].var_edge[k];
                        if (edge_layer(M, 1, e) != r) continue;
                        const R a = sz - prev[Ex + e] + f1;
                        sl.a[Ex + e] = a;
                        sl.mu[Ex + e] = M.wv[(nb + M.blk1[e]) * T + (t - 1)] * a;
                    }
                }
                __syncthreads();
                // the layer's check update again, its weights' gradient, and
                // the check adjoint (gradient to the input at each edge's
                // minimum over the other edges) into sl.out
                for (int c = threadIdx.x; c < mx + mz; c += kBlock) {
                    const int s = c < mx ? 0 : 1;
                    const int cc = c < mx ? c : c - mx;
                    const int off = g.side[s].chk_off[cc], deg = g.side[s].chk_off[cc + 1] - off;
                    if (deg == 0 || edge_layer(M, s, off) != r) continue;
                    const int base = (s == 0 ? 0 : Ex) + off;
                    const int *blk = s == 0 ? M.blk0 : M.blk1;
                    check_node_ms(sl.mu + base, deg, syn_bit(g, tp, b, s, cc), sl.raw + base, sl.out + base);
                    for (int i = 0; i < deg; ++i) {
                        const int wi = s * nb + blk[off + i];
                        const R ge = gc[base + i];
                        atomicAdd(&s_gw[wi], ge * sl.out[base + i]);
                        gc[base + i] = fabs(sl.raw[base + i]) > lmax ? R(0) : M.wc[wi * T + t - 1] * ge;
                    }
                    int i1 = 0;
                    for (int i = 1; i < deg; ++i)
                        if (fabs(sl.mu[base + i]) < fabs(sl.mu[base + i1])) i1 = i;
                    int i2 = i1 == 0 ? 1 : 0;
                    for (int i = 0; i < deg; ++i)
                        if (i != i1 && fabs(sl.mu[base + i]) < fabs(sl.mu[base + i2])) i2 = i;
                    R to1 = R(0), to2 = R(0);
                    for (int i = 0; i < deg; ++i) {
                        const R ge = gc[base + i];
                        const R se = sl.raw[base + i] < R(0) ? -ge : ge;
                        if (i == i1) to2 += se;
                        else to1 += se;
                    }
                    for (int i = 0; i < deg; ++i) sl.out[base + i] = R(0);
                    sl.out[base + i1] = sl.mu[base + i1] < R(0) ? -to1 : to1;
                    if (deg > 1) sl.out[base + i2] = sl.mu[base + i2] < R(0) ? -to2 : to2;
                }
                __syncthreads();
                // variable side: gradient of the state before the step
                for (int v = threadIdx.x; v < n; v += kBlock) {
                    R A[2];
                    for (int s = 0; s < 2; ++s) {
                        const int base = s == 0 ? 0 : Ex;
                        const int *blk = s == 0 ? M.blk0 : M.blk1;
                        R acc = R(0);
                        for (int k = g.side[s].var_off[v]; k < g.side[s].var_off[v + 1]; ++k) {
                            const int e = g.side[s].var_edge[k];
                            if (edge_layer(M, s, e) != r) continue;
                            const int wi = s * nb + blk[e];
                            const R gmu = sl.out[base + e];
                            atomicAdd(&s_gw[2 * nb + wi], gmu * sl.a[base + e]);
                            acc += M.wv[wi * T + t - 1] * gmu;
                        }
                        A[s] = acc;
                    }
                    const R sx = sl.S[v], sz = sl.S[n + v];
                    // d S_x: a_x directly, a_z through F_z(S_x); d S_z likewise
                    const R dsx = A[0] + A[1] * fterm_kd<1>(sx, ch.fz, ch.lfz);
                    const R dsz = A[1] + A[0] * fterm_kd<1>(sz, ch.fx, ch.lfx);
                    for (int s = 0; s < 2; ++s) {
                        const int base = s == 0 ? 0 : Ex;
                        const int *blk = s == 0 ? M.blk0 : M.blk1;
                        const R ds = s == 0 ? dsx : dsz;
                        for (int k = g.side[s].var_off[v]; k < g.side[s].var_off[v + 1]; ++k) {
                            const int e = g.side[s].var_edge[k];
                            if (edge_layer(M, s, e) == r) {
                                const R ga = M.wv[(s * nb + blk[e]) * T + t - 1] * sl.out[base + e];
                                gc[base + e] = ds - ga;
                            } else {
                                gc[base + e] += ds;
                            }
                        }
                    }
                }
                __syncthreads();
            }
            for (int i = threadIdx.x; i < 4 * nb; i += kBlock) {
                const int which = i / (2 * nb);  // 0 wc, 1 wv
                const int wi = i % (2 * nb);
                R *dst = which == 0 ? tb.gwc : tb.gwv;
                if (s_gw[i] != R(0)) atomicAdd(&dst[wi * T + t - 1], s_gw[i]);
                s_gw[i] = R(0);
            }
            __syncthreads();
        }
    }
}

// Adam on the double master copy; refreshes the float copy the FP32
// kernels read.
__global__ void adam_kernel(double *p, const double *grad, double *m, double *v, float *pf,
                            int count, double lr, double b1, double b2, double eps, int step) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double gi = grad[i];
    m[i] = b1 * m[i] + (1.0 - b1) * gi;
    v[i] = b2 * v[i] + (1.0 - b2) * gi * gi;
    double mh = m[i] / (1.0 - pow(b1, step));
    double vh = v[i] / (1.0 - pow(b2, step));
    p[i] -= lr * mh / (sqrt(vh) + eps);
    pf[i] = static_cast<float>(p[i]);
}

template <typename R>
__global__ void cast_kernel(const R *src, double *dst, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = static_cast<double>(src[i]);
}

template <typename R>
Model<R> model_of(const LearnedDevParams &lp, const R *params) {
    Model<R> M;
    M.nb = lp.nb;
    M.T = lp.T;
    M.use_mlp = lp.use_mlp;
    M.wc = params;
    M.wv = params + 2 * lp.nb * lp.T;
    M.mlp = params + 4 * lp.nb * lp.T;
    M.blk0 = lp.blk0;
    M.blk1 = lp.blk1;
    M.lpb = lp.lpb > 0 ? lp.lpb : 1;
    return M;
}

}  // namespace

size_t learned_slot_elems(const DevGraph &g) {
    const size_t Et = static_cast<size_t>(g.side[0].E) + g.side[1].E;
    return 6 * Et + 2 * static_cast<size_t>(g.n) + static_cast<size_t>(g.n);  // est as ints
}

template <typename R>
void launch_learned_train(const DevGraph &g, const LearnedDevParams &lp, const R *params,
                          TrialPlanes tp, int batch, R *act, R *scratch, int slots, R *tape_d,
                          R *tape_u, const double *prior, R *loss, R *grad, int *conv,
                          int *counters, bool backward, hipStream_t st) {
    const bool ms = lp.kind == 1;
    Model<R> M = model_of<R>(lp, params);
    Tape<R> tb;
    tb.act = act;
    tb.scratch = scratch;
    tb.slot_stride = learned_slot_elems(g);
    tb.tape_d = tape_d;
    tb.tape_u = tape_u;
    tb.prior = prior;
    tb.loss = loss;
    tb.gwc = grad;
    tb.gwv = grad + 2 * lp.nb * lp.T;
    tb.gmlp = grad + 4 * lp.nb * lp.T;
    tb.conv = conv;
    const int grid = slots < batch ? slots : batch;
    launch_fill_i32(counters, 0, 2, st);
    if (lp.kind == kLearnedLMS) {
        hipLaunchKernelGGL((learned_fwd_lms_kernel<R>), dim3(grid), dim3(kBlock), 0, st, g, M, tp, tb,
                           batch, counters);
        if (backward)
            hipLaunchKernelGGL((learned_bwd_lms_kernel<R>), dim3(grid), dim3(kBlock), 0, st, g, M, tp,
                               tb, batch, counters + 1);
        return;
    }
    if (ms)
        hipLaunchKernelGGL((learned_fwd_kernel<R, 1>), dim3(grid), dim3(kBlock), 0, st, g, M, tp, tb,
                           batch, counters);
    else
        hipLaunchKernelGGL((learned_fwd_kernel<R, 0>), dim3(grid), dim3(kBlock), 0, st, g, M, tp, tb,
                           batch, counters);
    if (backward && ms)
        hipLaunchKernelGGL((learned_bwd_kernel<R, 1>), dim3(grid), dim3(kBlock), 0, st, g, M, tp, tb,
                           batch, counters + 1);
    else if (backward)
        hipLaunchKernelGGL((learned_bwd_kernel<R, 0>), dim3(grid), dim3(kBlock), 0, st, g, M, tp, tb,
                           batch, counters + 1);
}

template void launch_learned_train<float>(const DevGraph &, const LearnedDevParams &, const float *,
                                          TrialPlanes, int, float *, float *, int, float *, float *,
                                          const double *, float *, float *, int *, int *, bool,
                                          hipStream_t);
template void launch_learned_train<double>(const DevGraph &, const LearnedDevParams &,
                                           const double *, TrialPlanes, int, double *, double *,
                                           int, double *, double *, const double *, double *,
                                           double *, int *, int *, bool, hipStream_t);

void launch_adam(double *p, const double *grad, double *m, double *v, float *pf, int count,
                 double lr, int step, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3((count + 255) / 256), dim3(256), 0, st, p, grad, m, v, pf,
                       count, lr, 0.9, 0.999, 1e-8, step);
}

void launch_to_double(const float *src, double *dst, int count, hipStream_t st) {
    hipLaunchKernelGGL((cast_kernel<float>), dim3((count + 255) / 256), dim3(256), 0, st, src, dst,
                       count);
}

}  // namespace jbp
