// ORACLE — test infrastructure, not product code (see oracle.cpp's header).
//
// FP64 reference for the learned (neural-weighted) joint BP.  This decoder
// has no counterpart in the reference program; it extends the reference's
// joint BP (jointbp_ets.cpp:2373-2519) with
//   * a weight per (side, permutation block, iteration) on every
//     variable-to-check message:   mu = wv[g][blk][t] * (S_g - lambda + F_g(S_other))
//   * a weight per (side, block, iteration) on every check-to-variable
//     message:                     lambda~ = wc[g][blk][t] * CN(mu)
//   * damping of the check-to-variable messages set by a 2-16-1 MLP from
//     (t / T, unsatisfied-check fraction):
//                                  lambda_t = (1 - d_t) lambda~ + d_t lambda_{t-1}
// in natural-log likelihood ratios.  With all weights 1 and damping 0 it is
// the reference's joint BP in the log domain (pinned by
// tests/test_learned_oracle.py against the plain oracle's hard decisions).
//
// Loss: mean over iterations 1..T, trials and the 2n per-qubit X/Z bits of
// the binary cross-entropy of the marginal bit probabilities against the
// sampled error.  The gradient is the hand-derived adjoint below; it is
// pinned against central finite differences of orl_loss.
//
// Check-node derivative: written as the O(deg^3) product over all other
// inputs, independent of the O(deg) recursion the device adjoint uses.
//
// Second kind (kind = 1, normalised min-sum): the same model with the
// max-product (min-sum) form of both node updates,
//   check:  CN_e(mu) = (-1)^syn prod_{k != e} sign(mu_k) * min_{k != e} |mu_k|
//   joint variable term:  F(L) = max(ln a + L, ln b) - max(ln c + L, ln d)
//           in place of ln((a e^L + b) / (c e^L + d)),
// hard decisions from the signs of the per-qubit X / Z marginals, and the
// BCE loss on those marginals.  With unit weights and no damping it is
// max-product BP, exact on a cycle-free graph (pinned against brute-force
// maximisation over all error patterns on small tree codes).  The loss is
// piecewise smooth in the parameters; orl_loss2 also returns a hash of every
// discrete choice the forward pass made (arg-minima, signs, clamps, max
// branches, estimates), so a finite-difference test can confirm that both
// of its points lie on the same smooth piece.
//
// Third kind (kind = 2, layered min-sum): the min-sum model above on a
// layered schedule.  The checks of each side are split into layers (the
// permutation-block code: layer r = block row r, both sides' layer r
// together); an iteration updates the layers in order, each from the
// messages as the earlier layers of the same iteration left them:
//   S_g(v) = sum of v's current check messages on side g
//   mu_e   = wv[g][blk][t] * (S_g(v) - lambda_e + F_g(S_other(v)))   (e in layer r)
//   lambda_e <- wc[g][blk][t] * CN_e(mu)                            (checks of layer r)
// The loss is taken on the marginals at the end of each iteration.  The
// estimate tested for convergence at iteration t is the one each layer's
// update starts from: the decode stops at the first t whose layers all start
// from the same estimate and whose rows all accept it (the rule the device
// decoder can test inside the iteration).  The adjoint walks the layer steps
// backwards; the state before step (t, r) is iteration t's messages on
// layers < r and iteration t-1's on the rest, so the per-iteration tape of
// the other kinds serves it too.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kHidden = 16;
constexpr int kMlpParams = kHidden * 2 + kHidden + kHidden + 1;  // 65
constexpr double kLlrMax = 64.0 * 0.69314718055994530942;     // 64 bits, in nats

struct Graph {
    int n = 0;
    int m[2] = {0, 0};
    int nb = 0;  // weight blocks per side
    std::vector<int> chk_off[2], edge_var[2], var_off[2], var_edge[2], blk[2];
    int nlay = 0;                 // layers (kind 2)
    std::vector<int> edge_lay[2]; // layer of each edge's check (kind 2)
};

inline double sigmoid(double x) {
    return x >= 0 ? 1.0 / (1.0 + std::exp(-x)) : std::exp(x) / (1.0 + std::exp(x));
}
inline double softplus(double x) { return x > 0 ? x + std::log1p(std::exp(-x)) : std::log1p(std::exp(x)); }

// F(L) = ln((a e^L + b) / (c e^L + d)) and its derivative.
inline double fterm(double L, const double k[4]) {
    if (L >= 0) return std::log(k[0] + k[1] * std::exp(-L)) - std::log(k[2] + k[3] * std::exp(-L));
    return std::log(k[0] * std::exp(L) + k[1]) - std::log(k[2] * std::exp(L) + k[3]);
}
inline double fterm_d(double L, const double k[4]) {
    return sigmoid(L + std::log(k[0] / k[1])) - sigmoid(L + std::log(k[2] / k[3]));
}

struct Params {
    const double *wc;   // [2][nb][T]
    const double *wv;   // [2][nb][T]
    const double *mlp;  // W1[16][2] | b1[16] | W2[16] | b2
    bool use_mlp;
    int kind;           // 0 sum-product, 1 min-sum, 2 layered min-sum
};

// Hash of the discrete choices of a min-sum forward pass (FNV-1a).
struct Sig {
    uint64_t h = 1469598103934665603ull;
    void add(uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    }
};

// Max-log joint-variable term F(L) = max(ln k0 + L, ln k1) - max(ln k2 + L, ln k3)
// and its derivative (0, +-1) away from the branch points.
inline double fterm_ms(double L, const double k[4], Sig *sig = nullptr) {
    const double a = std::log(k[0]) + L, b = std::log(k[1]);
    const double c = std::log(k[2]) + L, d = std::log(k[3]);
    if (sig) sig->add((a > b ? 1u : 0u) | (c > d ? 2u : 0u));
    return std::max(a, b) - std::max(c, d);
}
inline double fterm_ms_d(double L, const double k[4]) {
    return (std::log(k[0]) + L > std::log(k[1]) ? 1.0 : 0.0) -
           (std::log(k[2]) + L > std::log(k[3]) ? 1.0 : 0.0);
}
inline double fterm_k(int kind, double L, const double k[4], Sig *sig = nullptr) {
    return kind >= 1 ? fterm_ms(L, k, sig) : fterm(L, k);
}
inline double fterm_kd(int kind, double L, const double k[4]) {
    return kind >= 1 ? fterm_ms_d(L, k) : fterm_d(L, k);
}

// Min-sum check update, written from its definition: for each edge the
// minimum and the sign product over the other edges.
void check_node_ms(const double *mu, int deg, int syn, double *raw, double *out, Sig *sig) {
    uint64_t negmask = 0;
    for (int i = 0; i < deg; ++i) negmask |= static_cast<uint64_t>(mu[i] < 0 ? 1 : 0) << i;
    for (int i = 0; i < deg; ++i) {
        double mag = HUGE_VAL;
        int neg = syn & 1;
        for (int j = 0; j < deg; ++j) {
            if (j == i) continue;
            mag = std::min(mag, std::fabs(mu[j]));
            neg ^= mu[j] < 0 ? 1 : 0;
        }
        raw[i] = neg ? -mag : mag;
        out[i] = neg ? -std::min(mag, kLlrMax) : std::min(mag, kLlrMax);
    }
    if (sig) {
        int i1 = 0;
        for (int j = 1; j < deg; ++j)
            if (std::fabs(mu[j]) < std::fabs(mu[i1])) i1 = j;
        int i2 = i1 == 0 ? 1 : 0;
        for (int j = 0; j < deg; ++j)
            if (j != i1 && std::fabs(mu[j]) < std::fabs(mu[i2])) i2 = j;
        sig->add(negmask);
        sig->add(static_cast<uint64_t>(i1) | static_cast<uint64_t>(i2) << 8 |
                 static_cast<uint64_t>(std::fabs(mu[i1]) > kLlrMax) << 16 |
                 static_cast<uint64_t>(std::fabs(mu[i2]) > kLlrMax) << 17);
    }
}

double mlp_forward(const double *mlp, double x0, double x1, double *h) {
    double z = mlp[kHidden * 2 + kHidden + kHidden];
    for (int i = 0; i < kHidden; ++i) {
        h[i] = std::tanh(mlp[2 * i] * x0 + mlp[2 * i + 1] * x1 + mlp[2 * kHidden + i]);
        z += mlp[3 * kHidden + i] * h[i];
    }
    return sigmoid(z);
}

struct TrialTape {
    std::vector<double> marg;                  // last iteration's X then Z marginal LLRs (2n)
    std::vector<std::vector<double>> lam[2];   // [t][e], t = 0..T
    std::vector<double> d, x0, x1;             // per iteration 1..T (index t-1)
    std::vector<int> est;                      // last estimate
    int converged_at = 0;
};

// Check update in the pair form used by the device kernels (natural log):
// raw lambda_e = sign * ln(PE/PO); returns the clamped value.
void check_node(const double *mu, int deg, int syn, double *raw, double *out) {
    std::vector<double> e(deg), pe(deg + 1), po(deg + 1);
    unsigned negall = syn & 1u;
    std::vector<unsigned> neg(deg);
    for (int i = 0; i < deg; ++i) {
        neg[i] = mu[i] < 0 ? 1u : 0u;
        negall ^= neg[i];
        e[i] = std::exp(-std::fabs(mu[i]));
    }
    pe[0] = 1.0;
    po[0] = 0.0;
    for (int i = 0; i < deg; ++i) {
        pe[i + 1] = pe[i] + po[i] * e[i];
        po[i + 1] = pe[i] * e[i] + po[i];
    }
    double se = 1.0, so = 0.0;
    for (int i = deg - 1; i >= 0; --i) {
        double PE = pe[i] * se + po[i] * so;
        double PO = pe[i] * so + po[i] * se;
        double mag = std::log(PE) - std::log(PO);
        double sg = (negall ^ neg[i]) ? -1.0 : 1.0;
        raw[i] = sg * mag;
        out[i] = sg * std::min(mag, kLlrMax);
        double ne = se + so * e[i], no = se * e[i] + so;
        se = ne;
        so = no;
    }
}

// One trial forward; fills the tape; returns the trial's loss sum (not yet
// averaged).  err: states 0..3; prior: 4-ary normalized.
double trial_forward(const Graph &g, const Params &P, int T, const int *err, const double *prior,
                     TrialTape &tape, Sig *sig = nullptr) {
    const int kind = P.kind;
    const double fx[4] = {prior[0], prior[2], prior[1], prior[3]};
    const double fz[4] = {prior[0], prior[1], prior[2], prior[3]};
    const double lp[4] = {std::log(prior[0]), std::log(prior[1]), std::log(prior[2]), std::log(prior[3])};
    std::vector<int> syn[2];
    int synw = 0;
    for (int s = 0; s < 2; ++s) {
        syn[s].assign(g.m[s], 0);
        for (int c = 0; c < g.m[s]; ++c) {
            int par = 0;
            for (int e = g.chk_off[s][c]; e < g.chk_off[s][c + 1]; ++e) {
                int st = err[g.edge_var[s][e]];
                par ^= s == 0 ? (st & 1) : (st >> 1);
            }
            syn[s][c] = par;
            synw += par;
        }
    }
    const double mtot = g.m[0] + g.m[1];
    for (int s = 0; s < 2; ++s) tape.lam[s].assign(T + 1, std::vector<double>(g.edge_var[s].size(), 0.0));
    tape.d.assign(T, 0.0);
    tape.x0.assign(T, 0.0);
    tape.x1.assign(T, 0.0);
    std::vector<double> ubar(T + 1, 0.0);  // unsat fraction of est_t; est_0 = all I
    ubar[0] = synw / mtot;
    std::vector<double> S[2], mu[2], raw, out;
    std::vector<int> est(g.n, 0);
    double loss = 0.0;
    tape.converged_at = 0;
    tape.marg.assign(2 * static_cast<size_t>(g.n), 0.0);
    for (int t = 1; t <= T; ++t) {
        const std::vector<double> *lp_prev[2] = {&tape.lam[0][t - 1], &tape.lam[1][t - 1]};
        for (int s = 0; s < 2; ++s) {
            S[s].assign(g.n, 0.0);
            for (int v = 0; v < g.n; ++v)
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) S[s][v] += (*lp_prev[s])[g.var_edge[s][k]];
        }
        for (int s = 0; s < 2; ++s) {
            mu[s].assign(g.edge_var[s].size(), 0.0);
            for (int v = 0; v < g.n; ++v) {
                double f = s == 0 ? fterm_k(kind, S[1][v], fx, sig) : fterm_k(kind, S[0][v], fz, sig);
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                    int e = g.var_edge[s][k];
                    double a = S[s][v] - (*lp_prev[s])[e] + f;
                    mu[s][e] = P.wv[(s * g.nb + g.blk[s][e]) * T + (t - 1)] * a;
                }
            }
        }
        const double uin = ubar[t >= 2 ? t - 2 : 0];
        double h[kHidden];
        const double x0 = static_cast<double>(t) / T;
        double d = P.use_mlp ? mlp_forward(P.mlp, x0, uin, h) : 0.0;
        tape.d[t - 1] = d;
        tape.x0[t - 1] = x0;
        tape.x1[t - 1] = uin;
        for (int s = 0; s < 2; ++s) {
            for (int c = 0; c < g.m[s]; ++c) {
                int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
                raw.assign(deg, 0.0);
                out.assign(deg, 0.0);
                if (kind == 1) check_node_ms(&mu[s][off], deg, syn[s][c], raw.data(), out.data(), sig);
                else check_node(&mu[s][off], deg, syn[s][c], raw.data(), out.data());
                for (int i = 0; i < deg; ++i) {
                    int e = off + i;
                    double w = P.wc[(s * g.nb + g.blk[s][e]) * T + (t - 1)];
                    tape.lam[s][t][e] = (1.0 - d) * w * out[i] + d * tape.lam[s][t - 1][e];
                }
            }
        }
        // marginals, loss, estimate
        for (int s = 0; s < 2; ++s) {
            S[s].assign(g.n, 0.0);
            for (int v = 0; v < g.n; ++v)
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) S[s][v] += tape.lam[s][t][g.var_edge[s][k]];
        }
        for (int v = 0; v < g.n; ++v) {
            double lx = S[0][v] + fterm_k(kind, S[1][v], fx, sig);
            double lz = S[1][v] + fterm_k(kind, S[0][v], fz, sig);
            int yx = err[v] & 1, yz = err[v] >> 1;
            loss += softplus((2 * yx - 1) * lx) + softplus((2 * yz - 1) * lz);
            tape.marg[v] = lx;
            tape.marg[g.n + v] = lz;
            if (kind == 1) {  // per-bit decisions from the marginals' signs
                est[v] = (lx < 0 ? 1 : 0) | (lz < 0 ? 2 : 0);
                if (sig) sig->add(static_cast<uint64_t>(est[v]));
                continue;
            }
            double l0 = lp[0] + S[0][v] + S[1][v], l1 = lp[1] + S[1][v], l2 = lp[2] + S[0][v], l3 = lp[3];
            int best = 0;
            double bv = l0;
            if (l1 > bv) { best = 1; bv = l1; }
            if (l2 > bv) { best = 2; bv = l2; }
            if (l3 > bv) { best = 3; }
            est[v] = best;
        }
        int unsat = 0;
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < g.m[s]; ++c) {
                int par = 0;
                for (int e = g.chk_off[s][c]; e < g.chk_off[s][c + 1]; ++e) {
                    int st = est[g.edge_var[s][e]];
                    par ^= s == 0 ? (st & 1) : (st >> 1);
                }
                unsat += par != syn[s][c];
            }
        ubar[t] = unsat / mtot;
        if (unsat == 0 && tape.converged_at == 0) tape.converged_at = t;
    }
    tape.est = est;
    return loss;
}

// Adjoint of trial_forward.  scale multiplies the trial's loss sum.
void trial_backward(const Graph &g, const Params &P, int T, const int *err, const double *prior,
                    const TrialTape &tape, double scale, double *gwc, double *gwv, double *gmlp) {
    const int kind = P.kind;
    const double fx[4] = {prior[0], prior[2], prior[1], prior[3]};
    const double fz[4] = {prior[0], prior[1], prior[2], prior[3]};
    std::vector<int> syn[2];
    for (int s = 0; s < 2; ++s) {
        syn[s].assign(g.m[s], 0);
        for (int c = 0; c < g.m[s]; ++c) {
            int par = 0;
            for (int e = g.chk_off[s][c]; e < g.chk_off[s][c + 1]; ++e) {
                int st = err[g.edge_var[s][e]];
                par ^= s == 0 ? (st & 1) : (st >> 1);
            }
            syn[s][c] = par;
        }
    }
    std::vector<double> G[2], Gp[2], S[2], Sp[2], mu[2], a[2], raw[2], out[2], Gmu[2];
    for (int s = 0; s < 2; ++s) G[s].assign(g.edge_var[s].size(), 0.0);
    for (int t = T; t >= 1; --t) {
        const auto &lam_t = tape.lam;
        // loss term at t
        for (int s = 0; s < 2; ++s) {
            S[s].assign(g.n, 0.0);
            for (int v = 0; v < g.n; ++v)
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) S[s][v] += lam_t[s][t][g.var_edge[s][k]];
        }
        for (int v = 0; v < g.n; ++v) {
            double lx = S[0][v] + fterm_k(kind, S[1][v], fx);
            double lz = S[1][v] + fterm_k(kind, S[0][v], fz);
            double kx = 2 * (err[v] & 1) - 1, kz = 2 * (err[v] >> 1) - 1;
            double glx = scale * kx * sigmoid(kx * lx);
            double glz = scale * kz * sigmoid(kz * lz);
            double gsx = glx + glz * fterm_kd(kind, S[0][v], fz);
            double gsz = glz + glx * fterm_kd(kind, S[1][v], fx);
            for (int k = g.var_off[0][v]; k < g.var_off[0][v + 1]; ++k) G[0][g.var_edge[0][k]] += gsx;
            for (int k = g.var_off[1][v]; k < g.var_off[1][v + 1]; ++k) G[1][g.var_edge[1][k]] += gsz;
        }
        // recompute the var pass and check pass of iteration t
        for (int s = 0; s < 2; ++s) {
            Sp[s].assign(g.n, 0.0);
            for (int v = 0; v < g.n; ++v)
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) Sp[s][v] += lam_t[s][t - 1][g.var_edge[s][k]];
        }
        for (int s = 0; s < 2; ++s) {
            mu[s].assign(g.edge_var[s].size(), 0.0);
            a[s].assign(g.edge_var[s].size(), 0.0);
            for (int v = 0; v < g.n; ++v) {
                double f = s == 0 ? fterm_k(kind, Sp[1][v], fx) : fterm_k(kind, Sp[0][v], fz);
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                    int e = g.var_edge[s][k];
                    a[s][e] = Sp[s][v] - lam_t[s][t - 1][e] + f;
                    mu[s][e] = P.wv[(s * g.nb + g.blk[s][e]) * T + (t - 1)] * a[s][e];
                }
            }
            raw[s].assign(g.edge_var[s].size(), 0.0);
            out[s].assign(g.edge_var[s].size(), 0.0);
            for (int c = 0; c < g.m[s]; ++c) {
                int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
                if (kind == 1) check_node_ms(&mu[s][off], deg, syn[s][c], &raw[s][off], &out[s][off], nullptr);
                else check_node(&mu[s][off], deg, syn[s][c], &raw[s][off], &out[s][off]);
            }
        }
        const double d = tape.d[t - 1];
        // damping and check weights
        double gd = 0.0;
        for (int s = 0; s < 2; ++s) {
            Gp[s].assign(g.edge_var[s].size(), 0.0);
            for (size_t e = 0; e < g.edge_var[s].size(); ++e) {
                size_t wi = (s * g.nb + g.blk[s][e]) * T + (t - 1);
                double w = P.wc[wi];
                gd += G[s][e] * (lam_t[s][t - 1][e] - w * out[s][e]);
                Gp[s][e] = d * G[s][e];
                double gtil = (1.0 - d) * G[s][e];
                gwc[wi] += gtil * out[s][e];
                G[s][e] = std::fabs(raw[s][e]) > kLlrMax ? 0.0 : w * gtil;  // now d loss / d lambda-hat
            }
        }
        if (P.use_mlp) {
            double h[kHidden];
            mlp_forward(P.mlp, tape.x0[t - 1], tape.x1[t - 1], h);
            double gz = gd * d * (1.0 - d);
            gmlp[2 * kHidden + kHidden + kHidden] += gz;
            for (int i = 0; i < kHidden; ++i) {
                gmlp[3 * kHidden + i] += gz * h[i];
                double ga = gz * P.mlp[3 * kHidden + i] * (1.0 - h[i] * h[i]);
                gmlp[2 * i] += ga * tape.x0[t - 1];
                gmlp[2 * i + 1] += ga * tape.x1[t - 1];
                gmlp[2 * kHidden + i] += ga;
            }
        }
        // check-node adjoint: d lambda_e / d mu_j = sign_c * P_ej sech^2(mu_j/2) / sech^2(raw_e/2)
        // (min-sum: d raw_e / d mu_j = sign(raw_e) sign(mu_j) for the j that
        // attains the minimum over the other edges, 0 for the rest)
        for (int s = 0; s < 2; ++s) {
            Gmu[s].assign(g.edge_var[s].size(), 0.0);
            for (int c = 0; kind == 1 && c < g.m[s]; ++c) {
                int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
                for (int e = 0; e < deg; ++e) {
                    double ge = G[s][off + e];
                    if (ge == 0.0) continue;
                    int js = -1;
                    for (int j = 0; j < deg; ++j)
                        if (j != e && (js < 0 || std::fabs(mu[s][off + j]) < std::fabs(mu[s][off + js]))) js = j;
                    double se = raw[s][off + e] < 0 ? -1.0 : 1.0;
                    double sj = mu[s][off + js] < 0 ? -1.0 : 1.0;
                    Gmu[s][off + js] += ge * se * sj;
                }
            }
            for (int c = 0; kind == 0 && c < g.m[s]; ++c) {
                int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
                double sgn = syn[s][c] ? -1.0 : 1.0;
                for (int e = 0; e < deg; ++e) {
                    double ge = G[s][off + e];
                    if (ge == 0.0) continue;
                    double ch = std::cosh(raw[s][off + e] / 2.0);
                    for (int j = 0; j < deg; ++j) {
                        if (j == e) continue;
                        double prod = 1.0;
                        for (int k = 0; k < deg; ++k)
                            if (k != e && k != j) prod *= std::tanh(mu[s][off + k] / 2.0);
                        double cj = std::cosh(mu[s][off + j] / 2.0);
                        Gmu[s][off + j] += ge * sgn * prod * (ch * ch) / (cj * cj);
                    }
                }
            }
        }
        // variable-pass adjoint
        for (int s = 0; s < 2; ++s) {
            const int o = 1 - s;
            for (int v = 0; v < g.n; ++v) {
                double ga_sum = 0.0;
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                    int e = g.var_edge[s][k];
                    size_t wi = (s * g.nb + g.blk[s][e]) * T + (t - 1);
                    gwv[wi] += Gmu[s][e] * a[s][e];
                    double ga = P.wv[wi] * Gmu[s][e];
                    ga_sum += ga;
                    Gp[s][e] -= ga;
                }
                for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) Gp[s][g.var_edge[s][k]] += ga_sum;
                // F_s(S_o) term: d a / d S_o = F_s'(S_o)
                double fd = s == 0 ? fterm_kd(kind, Sp[1][v], fx) : fterm_kd(kind, Sp[0][v], fz);
                for (int k = g.var_off[o][v]; k < g.var_off[o][v + 1]; ++k) Gp[o][g.var_edge[o][k]] += ga_sum * fd;
            }
        }
        G[0].swap(Gp[0]);
        G[1].swap(Gp[1]);
    }
}

// ---- kind 2: layered min-sum ----

std::vector<int> trial_syndromes(const Graph &g, const int *err, int s) {
    std::vector<int> syn(g.m[s], 0);
    for (int c = 0; c < g.m[s]; ++c) {
        int par = 0;
        for (int e = g.chk_off[s][c]; e < g.chk_off[s][c + 1]; ++e) {
            int st = err[g.edge_var[s][e]];
            par ^= s == 0 ? (st & 1) : (st >> 1);
        }
        syn[c] = par;
    }
    return syn;
}

// Per-variable sums of a message state.
void var_sums(const Graph &g, const std::vector<double> *lam, std::vector<double> *S) {
    for (int s = 0; s < 2; ++s) {
        S[s].assign(g.n, 0.0);
        for (int v = 0; v < g.n; ++v)
            for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) S[s][v] += lam[s][g.var_edge[s][k]];
    }
}

// Layer step (t, r) from the state `cur`: a and mu of layer r's edges (other
// edges untouched), raw and clamped check outputs of layer r's checks.
void layer_step(const Graph &g, const Params &P, int T, int t, int r, const std::vector<int> *syn,
                const double *fx, const double *fz, const std::vector<double> *cur,
                std::vector<double> *S, std::vector<double> *a, std::vector<double> *mu,
                std::vector<double> *raw, std::vector<double> *out, Sig *sig) {
    var_sums(g, cur, S);
    for (int s = 0; s < 2; ++s) {
        for (int v = 0; v < g.n; ++v) {
            bool any = false;
            for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1] && !any; ++k)
                any = g.edge_lay[s][g.var_edge[s][k]] == r;
            if (!any) continue;
            const double f = s == 0 ? fterm_ms(S[1][v], fx, sig) : fterm_ms(S[0][v], fz, sig);
            for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                const int e = g.var_edge[s][k];
                if (g.edge_lay[s][e] != r) continue;
                a[s][e] = S[s][v] - cur[s][e] + f;
                mu[s][e] = P.wv[(s * g.nb + g.blk[s][e]) * T + (t - 1)] * a[s][e];
            }
        }
        for (int c = 0; c < g.m[s]; ++c) {
            const int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
            if (deg == 0 || g.edge_lay[s][off] != r) continue;
            check_node_ms(&mu[s][off], deg, syn[s][c], &raw[s][off], &out[s][off], sig);
        }
    }
}

double trial_forward_layered(const Graph &g, const Params &P, int T, const int *err,
                             const double *prior, TrialTape &tape, Sig *sig) {
    const double fx[4] = {prior[0], prior[2], prior[1], prior[3]};
    const double fz[4] = {prior[0], prior[1], prior[2], prior[3]};
    std::vector<int> syn[2] = {trial_syndromes(g, err, 0), trial_syndromes(g, err, 1)};
    for (int s = 0; s < 2; ++s) tape.lam[s].assign(T + 1, std::vector<double>(g.edge_var[s].size(), 0.0));
    tape.d.assign(T, 0.0);
    tape.x0.assign(T, 0.0);
    tape.x1.assign(T, 0.0);
    tape.converged_at = 0;
    tape.marg.assign(2 * static_cast<size_t>(g.n), 0.0);
    std::vector<double> cur[2], S[2], a[2], mu[2], raw[2], out[2];
    for (int s = 0; s < 2; ++s) {
        cur[s].assign(g.edge_var[s].size(), 0.0);
        a[s].assign(g.edge_var[s].size(), 0.0);
        mu[s].assign(g.edge_var[s].size(), 0.0);
        raw[s].assign(g.edge_var[s].size(), 0.0);
        out[s].assign(g.edge_var[s].size(), 0.0);
    }
    std::vector<int> est(g.n, 0), est0(g.n, 0);
    double loss = 0.0;
    for (int t = 1; t <= T; ++t) {
        bool stable = true, rows_ok = true;
        for (int r = 0; r < g.nlay; ++r) {
            // the estimate this layer starts from, and its rows' parities
            var_sums(g, cur, S);
            for (int v = 0; v < g.n; ++v) {
                const double lx = S[0][v] + fterm_ms(S[1][v], fx), lz = S[1][v] + fterm_ms(S[0][v], fz);
                const int e = (lx < 0 ? 1 : 0) | (lz < 0 ? 2 : 0);
                if (r == 0) est0[v] = e;
                else if (e != est0[v]) stable = false;
                est[v] = e;
            }
            for (int s = 0; s < 2; ++s)
                for (int c = 0; c < g.m[s]; ++c) {
                    const int off = g.chk_off[s][c];
                    if (g.chk_off[s][c + 1] == off || g.edge_lay[s][off] != r) continue;
                    int par = 0;
                    for (int e = off; e < g.chk_off[s][c + 1]; ++e) {
                        const int st = est[g.edge_var[s][e]];
                        par ^= s == 0 ? (st & 1) : (st >> 1);
                    }
                    if (par != syn[s][c]) rows_ok = false;
                }
            layer_step(g, P, T, t, r, syn, fx, fz, cur, S, a, mu, raw, out, sig);
            for (int s = 0; s < 2; ++s)
                for (size_t e = 0; e < cur[s].size(); ++e)
                    if (g.edge_lay[s][e] == r)
                        cur[s][e] = P.wc[(s * g.nb + g.blk[s][e]) * T + (t - 1)] * out[s][e];
        }
        if (stable && rows_ok && tape.converged_at == 0) tape.converged_at = t;
        for (int s = 0; s < 2; ++s) tape.lam[s][t] = cur[s];
        var_sums(g, cur, S);
        for (int v = 0; v < g.n; ++v) {
            const double lx = S[0][v] + fterm_ms(S[1][v], fx, sig);
            const double lz = S[1][v] + fterm_ms(S[0][v], fz, sig);
            const int yx = err[v] & 1, yz = err[v] >> 1;
            loss += softplus((2 * yx - 1) * lx) + softplus((2 * yz - 1) * lz);
            tape.marg[v] = lx;
            tape.marg[g.n + v] = lz;
            est[v] = (lx < 0 ? 1 : 0) | (lz < 0 ? 2 : 0);
            if (sig) sig->add(static_cast<uint64_t>(est[v]));
        }
    }
    tape.est = est;
    return loss;
}

void trial_backward_layered(const Graph &g, const Params &P, int T, const int *err,
                            const double *prior, const TrialTape &tape, double scale, double *gwc,
                            double *gwv) {
    const double fx[4] = {prior[0], prior[2], prior[1], prior[3]};
    const double fz[4] = {prior[0], prior[1], prior[2], prior[3]};
    std::vector<int> syn[2] = {trial_syndromes(g, err, 0), trial_syndromes(g, err, 1)};
    std::vector<double> G[2], Gn[2], pre[2], S[2], a[2], mu[2], raw[2], out[2], Gmu[2], Ga[2];
    for (int s = 0; s < 2; ++s) {
        const size_t E = g.edge_var[s].size();
        G[s].assign(E, 0.0);
        a[s].assign(E, 0.0);
        mu[s].assign(E, 0.0);
        raw[s].assign(E, 0.0);
        out[s].assign(E, 0.0);
    }
    for (int t = T; t >= 1; --t) {
        // loss term on the marginals at the end of iteration t
        const std::vector<double> endv[2] = {tape.lam[0][t], tape.lam[1][t]};
        var_sums(g, endv, S);
        for (int v = 0; v < g.n; ++v) {
            const double lx = S[0][v] + fterm_ms(S[1][v], fx), lz = S[1][v] + fterm_ms(S[0][v], fz);
            const double kx = 2 * (err[v] & 1) - 1, kz = 2 * (err[v] >> 1) - 1;
            const double glx = scale * kx * sigmoid(kx * lx), glz = scale * kz * sigmoid(kz * lz);
            const double gsx = glx + glz * fterm_ms_d(S[0][v], fz);
            const double gsz = glz + glx * fterm_ms_d(S[1][v], fx);
            for (int k = g.var_off[0][v]; k < g.var_off[0][v + 1]; ++k) G[0][g.var_edge[0][k]] += gsx;
            for (int k = g.var_off[1][v]; k < g.var_off[1][v + 1]; ++k) G[1][g.var_edge[1][k]] += gsz;
        }
        for (int r = g.nlay - 1; r >= 0; --r) {
            // the state before step (t, r)
            for (int s = 0; s < 2; ++s) {
                pre[s].resize(g.edge_var[s].size());
                for (size_t e = 0; e < pre[s].size(); ++e)
                    pre[s][e] = g.edge_lay[s][e] < r ? tape.lam[s][t][e] : tape.lam[s][t - 1][e];
            }
            layer_step(g, P, T, t, r, syn, fx, fz, pre, S, a, mu, raw, out, nullptr);
            // lambda_e = wc * clamp(raw_e) on layer r
            for (int s = 0; s < 2; ++s) {
                Gmu[s].assign(g.edge_var[s].size(), 0.0);
                std::vector<double> Gout(g.edge_var[s].size(), 0.0);
                for (size_t e = 0; e < G[s].size(); ++e) {
                    if (g.edge_lay[s][e] != r) continue;
                    const size_t wi = (s * g.nb + g.blk[s][e]) * T + (t - 1);
                    gwc[wi] += G[s][e] * out[s][e];
                    Gout[e] = std::fabs(raw[s][e]) > kLlrMax ? 0.0 : P.wc[wi] * G[s][e];
                }
                // min-sum check adjoint: each edge's gradient to the input
                // attaining the minimum over the other edges
                for (int c = 0; c < g.m[s]; ++c) {
                    const int off = g.chk_off[s][c], deg = g.chk_off[s][c + 1] - off;
                    if (deg == 0 || g.edge_lay[s][off] != r) continue;
                    for (int e = 0; e < deg; ++e) {
                        const double ge = Gout[off + e];
                        if (ge == 0.0) continue;
                        int js = -1;
                        for (int j = 0; j < deg; ++j)
                            if (j != e && (js < 0 || std::fabs(mu[s][off + j]) < std::fabs(mu[s][off + js]))) js = j;
                        const double se = raw[s][off + e] < 0 ? -1.0 : 1.0;
                        const double sj = mu[s][off + js] < 0 ? -1.0 : 1.0;
                        Gmu[s][off + js] += ge * se * sj;
                    }
                }
            }
            // variable side: a_e = S_g(v) - lambda_e + F_g(S_o(v)), mu = wv a
            std::vector<double> A[2];
            for (int s = 0; s < 2; ++s) {
                Ga[s].assign(g.edge_var[s].size(), 0.0);
                A[s].assign(g.n, 0.0);
                for (int v = 0; v < g.n; ++v)
                    for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                        const int e = g.var_edge[s][k];
                        if (g.edge_lay[s][e] != r) continue;
                        const size_t wi = (s * g.nb + g.blk[s][e]) * T + (t - 1);
                        gwv[wi] += Gmu[s][e] * a[s][e];
                        Ga[s][e] = P.wv[wi] * Gmu[s][e];
                        A[s][v] += Ga[s][e];
                    }
            }
            for (int s = 0; s < 2; ++s) {
                const int o = 1 - s;
                const double *kf = s == 0 ? fz : fx;  // side o's term F_o(S_s)
                Gn[s].assign(g.edge_var[s].size(), 0.0);
                for (int v = 0; v < g.n; ++v) {
                    const double dS = A[s][v] + A[o][v] * fterm_ms_d(S[s][v], kf);
                    for (int k = g.var_off[s][v]; k < g.var_off[s][v + 1]; ++k) {
                        const int e = g.var_edge[s][k];
                        const bool in = g.edge_lay[s][e] == r;
                        Gn[s][e] = (in ? 0.0 : G[s][e]) + dS - (in ? Ga[s][e] : 0.0);
                    }
                }
            }
            G[0].swap(Gn[0]);
            G[1].swap(Gn[1]);
        }
    }
}

Graph make_graph(int n, int mx, int mz, const int *xoff, const int *xvar, const int *zoff,
                 const int *zvar, const int *xblk, const int *zblk, int nb) {
    Graph g;
    g.n = n;
    g.m[0] = mx;
    g.m[1] = mz;
    g.nb = nb;
    const int *off[2] = {xoff, zoff}, *var[2] = {xvar, zvar}, *blk[2] = {xblk, zblk};
    for (int s = 0; s < 2; ++s) {
        int E = off[s][g.m[s]];
        g.chk_off[s].assign(off[s], off[s] + g.m[s] + 1);
        g.edge_var[s].assign(var[s], var[s] + E);
        g.blk[s].assign(blk[s], blk[s] + E);
        std::vector<int> cnt(n + 1, 0);
        for (int v : g.edge_var[s]) cnt[v + 1]++;
        for (int v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
        g.var_off[s] = cnt;
        g.var_edge[s].assign(E, 0);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < E; ++e) g.var_edge[s][fill[g.edge_var[s][e]]++] = e;
    }
    return g;
}

void prior_of(double p, double *out) {
    double m[4] = {1.0 - p, p / 3.0, p / 3.0, p / 3.0};
    double s = m[0] + m[1] + m[2] + m[3];
    for (int i = 0; i < 4; ++i) out[i] = m[i] / s;
}

}  // namespace

extern "C" {

int orl_mlp_params() { return kMlpParams; }

// Mean loss over `batch` trials (errors batch x n, one p per trial) and,
// when grad_* are non-null, its gradient.  Per trial, conv_out (may be
// null) receives the first iteration whose estimate satisfies both
// syndromes (0 = none) and est_out (may be null) the final estimate.
// kind 0: sum-product, 1: min-sum, 2: layered min-sum (header; orl_loss3,
// with each side's check layers xlay / zlay, nlay layers).  signature (may be null): hash
// of the min-sum forward pass's discrete choices over the whole batch.
// marg_out (may be null): per trial the last iteration's per-qubit X then Z
// marginal log-likelihood ratios (batch x 2n).
double orl_loss3(int kind, int n, int mx, int mz, const int *xoff, const int *xvar,
                 const int *zoff, const int *zvar, const int *xblk, const int *zblk, int nb,
                 const int *xlay, const int *zlay, int nlay, int T, const double *wc,
                 const double *wv, const double *mlp, int use_mlp, int batch, const int *err,
                 const double *p, double *grad_wc, double *grad_wv, double *grad_mlp,
                 int *conv_out, int *est_out, uint64_t *signature, double *marg_out) {
    Graph g = make_graph(n, mx, mz, xoff, xvar, zoff, zvar, xblk, zblk, nb);
    Params P{wc, wv, mlp, use_mlp != 0, kind == 1 ? 1 : (kind == 2 ? 2 : 0)};
    if (P.kind == 2) {
        // layers per check -> per edge; the layered kind has no damping
        g.nlay = nlay;
        const int *lay[2] = {xlay, zlay};
        for (int s = 0; s < 2; ++s) {
            g.edge_lay[s].assign(g.edge_var[s].size(), 0);
            for (int c = 0; c < g.m[s]; ++c)
                for (int e = g.chk_off[s][c]; e < g.chk_off[s][c + 1]; ++e) g.edge_lay[s][e] = lay[s][c];
        }
        P.use_mlp = false;
    }
    Sig sig;
    const double scale = 1.0 / (static_cast<double>(batch) * T * 2.0 * n);
    if (grad_wc) std::memset(grad_wc, 0, sizeof(double) * 2 * nb * T);
    if (grad_wv) std::memset(grad_wv, 0, sizeof(double) * 2 * nb * T);
    if (grad_mlp) std::memset(grad_mlp, 0, sizeof(double) * kMlpParams);
    std::vector<double> dummy_wc(2 * nb * T), dummy_wv(2 * nb * T), dummy_mlp(kMlpParams);
    double total = 0.0;
    for (int b = 0; b < batch; ++b) {
        double prior[4];
        prior_of(p[b], prior);
        TrialTape tape;
        const int *eb = err + static_cast<size_t>(b) * n;
        if (P.kind == 2) total += trial_forward_layered(g, P, T, eb, prior, tape, signature ? &sig : nullptr);
        else total += trial_forward(g, P, T, eb, prior, tape, signature ? &sig : nullptr);
        if (conv_out) conv_out[b] = tape.converged_at;
        if (est_out) std::copy(tape.est.begin(), tape.est.end(), est_out + static_cast<size_t>(b) * n);
        if (marg_out) std::copy(tape.marg.begin(), tape.marg.end(), marg_out + static_cast<size_t>(b) * 2 * n);
        if (grad_wc || grad_wv || grad_mlp) {
            double *gc = grad_wc ? grad_wc : dummy_wc.data(), *gv = grad_wv ? grad_wv : dummy_wv.data();
            if (P.kind == 2) trial_backward_layered(g, P, T, eb, prior, tape, scale, gc, gv);
            else trial_backward(g, P, T, eb, prior, tape, scale, gc, gv, grad_mlp ? grad_mlp : dummy_mlp.data());
        }
    }
    if (signature) *signature = sig.h;
    return total * scale;
}

double orl_loss2(int kind, int n, int mx, int mz, const int *xoff, const int *xvar,
                 const int *zoff, const int *zvar, const int *xblk, const int *zblk, int nb, int T,
                 const double *wc, const double *wv, const double *mlp, int use_mlp, int batch,
                 const int *err, const double *p, double *grad_wc, double *grad_wv,
                 double *grad_mlp, int *conv_out, int *est_out, uint64_t *signature,
                 double *marg_out) {
    if (kind == 2) return 0.0;  // the layered kind needs its layers: orl_loss3
    return orl_loss3(kind, n, mx, mz, xoff, xvar, zoff, zvar, xblk, zblk, nb, nullptr, nullptr, 0, T,
                     wc, wv, mlp, use_mlp, batch, err, p, grad_wc, grad_wv, grad_mlp, conv_out,
                     est_out, signature, marg_out);
}

double orl_loss(int n, int mx, int mz, const int *xoff, const int *xvar, const int *zoff,
                const int *zvar, const int *xblk, const int *zblk, int nb, int T, const double *wc,
                const double *wv, const double *mlp, int use_mlp, int batch, const int *err,
                const double *p, double *grad_wc, double *grad_wv, double *grad_mlp, int *conv_out,
                int *est_out) {
    return orl_loss2(0, n, mx, mz, xoff, xvar, zoff, zvar, xblk, zblk, nb, T, wc, wv, mlp, use_mlp,
                     batch, err, p, grad_wc, grad_wv, grad_mlp, conv_out, est_out, nullptr, nullptr);
}

}  // extern "C"
