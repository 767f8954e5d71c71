// Learned-BP model description shared by the training kernels
// (jbp_learned.hip), the inference decoder (jbp_decode_f32.hip) and the host
// orchestration (jbp_api.hip).
//
// Parameter vector layout (one array of 4*nb*T + 65 values):
//   wc[2][nb][T]   check-to-variable weights per side, block, iteration
//   wv[2][nb][T]   variable-to-check weights
//   mlp[65]        damping MLP: W1[16][2] | b1[16] | W2[16] | b2
// For a permutation-block code the block of the edge at position cc of
// check r*P + x is r*L + cc (nb = J*L).
#pragma once

#include "jbp_common.h"

namespace jbp {

constexpr int kMlpHidden = 16;
constexpr int kMlpParams = 4 * kMlpHidden + 1;  // 65
constexpr int kMaxLearnedT = 64;
constexpr int kMaxLearnedBlocks = 64;

// Model kinds: sum-product (the reference's joint BP with weights),
// normalised min-sum (max-product updates) and layered min-sum (the same
// updates on a layered schedule: block row r of both sides is layer r, the
// layers of an iteration updated in turn; no damping).
constexpr int kLearnedSP = 0;
constexpr int kLearnedMS = 1;
constexpr int kLearnedLMS = 2;

struct LearnedDevParams {
    int nb = 0, T = 0, use_mlp = 1, kind = kLearnedSP;
    int lpb = 1;                // weight blocks per layer (the layer of block b is b / lpb)
    const int *blk0 = nullptr;  // block of each X edge (device)
    const int *blk1 = nullptr;  // block of each Z edge (device)
};

inline int learned_param_count(int nb, int T) { return 4 * nb * T + kMlpParams; }

size_t learned_slot_elems(const DevGraph &g);

template <typename R>
void launch_learned_train(const DevGraph &g, const LearnedDevParams &lp, const R *params,
                          TrialPlanes tp, int batch, R *act, R *scratch, int slots, R *tape_d,
                          R *tape_u, const double *prior, R *loss, R *grad, int *conv,
                          int *counters, bool backward, hipStream_t st);

// FP32 training of the flooding min-sum kind on the permutation-block code
// with P = 768, J = 3, L = 12 (jbp_learned_qc.hip): the same loss, gradient
// and convergence iterations as launch_learned_train<float>, with the
// messages on chip and the tape as [trial][T+1][72 edges][768] floats in
// act (the same size).
// chk: the forward pass's check records for the adjoint (the two smallest
// input magnitudes, where they are, and the signs of every check row of
// every iteration), learned_qc_chk_bytes(batch, T) bytes.
bool learned_qc_train_supported(const DevGraph &g, const LearnedDevParams &lp);
size_t learned_qc_chk_bytes(int batch, int T);
void launch_learned_train_qc(const DevGraph &g, const LearnedDevParams &lp, const float *params,
                             TrialPlanes tp, int batch, float *act, uint32_t *chk, float *tape_d,
                             float *tape_u, const double *prior, float *loss, float *grad, int *conv,
                             int *counters, bool backward, hipStream_t st);

void launch_adam(double *p, const double *grad, double *m, double *v, float *pf, int count,
                 double lr, int step, hipStream_t st);
void launch_to_double(const float *src, double *dst, int count, hipStream_t st);

// Learned inference on the LDS-resident decoder: T learned iterations with
// the float parameter vector, then plain BP restarted from the channel
// prior for up to max_iter iterations on trials that did not converge.
struct LearnedInfer {
    // device, learned_param_count(nb, T) floats, iteration-major: for each
    // iteration t, wc X[nb] | wc Z[nb] | wv X[nb] | wv Z[nb]; then mlp[65]
    const float *params = nullptr;
    int T = 0;
    int use_mlp = 1;
    int kind = kLearnedSP;
    // min-sum kind: per variable pass k = 0..T one record of kMsRec words,
    // for each column block cc: (wc X, wc Z) of iteration k-1 (1 at k = 0)
    // for r = 0..J-1, then (wv X, wv Z) of iteration k (1 at k = T), as FP16
    // pairs, two words of padding (device, (T+1) x kMsRec uint32).
    // Layered min-sum kind: record k = 0..T-1 holds iteration k+1's own
    // weights as one 32-word group per block row r: (wc X, wc Z) for column
    // blocks 0..11, then (wv X, wv Z) for 0..11, then 8 words of padding.
    const uint32_t *ms_rec = nullptr;
};
// Words of one min-sum weight record: 12 column blocks x (3 + 3 + 2).
constexpr int kMsRecCC = 8;
constexpr int kMsRec = 12 * kMsRecCC;

}  // namespace jbp
