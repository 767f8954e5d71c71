// Definitions shared by the host orchestration (jbp_api.cpp) and the HIP
// kernels of libjbp.  Plain structs of device pointers and sizes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jbp {

// Device copy of one Tanner graph side (check-major CSR, the edge order of
// the reference's build_cuda_graph, jointbp_ets.cpp:937-1016).
struct DevSide {
    int m = 0;                 // checks
    int E = 0;                 // edges
    const int *chk_off = nullptr;   // m+1
    const int *edge_var = nullptr;  // E
    const int *var_off = nullptr;   // n+1
    const int *var_edge = nullptr;  // E: edge ids per variable, ascending check
    // FP64 message rows, position-major: the i-th edge of check c lives in
    // row i*m + c, so a wavefront walking checks c..c+63 at position i reads
    // one contiguous span (rows = max check degree x m; unused rows idle)
    int cap = 0;
    const int *vrow = nullptr;      // E: row of each entry of the variable lists (var_edge order)
};

struct DevGraph {
    int n = 0;
    int wn = 0;                // 32-bit words of an n-bit vector
    int wm[2] = {0, 0};        // 32-bit words of an m-bit vector per side
    int max_cdeg = 0, max_vdeg = 0;
    int min_cdeg = 0;
    DevSide side[2];
    // variable -> edge lists padded to vpad_w entries (4, 8 or 16) with the
    // side's edge count E, which the generic FP32 decoder stores as an
    // always-zero message: [n][vpad_w] per side
    int vpad_w = 0;
    const int *vpad[2] = {nullptr, nullptr};
    // check -> variable lists padded to cpad_w entries (8, 12, 16 or 32)
    // with the variable index n, for the syndrome gathers: [m][cpad_w] per
    // side (rows 16-byte aligned)
    int cpad_w = 0;
    const int *cpad[2] = {nullptr, nullptr};
    // quasi-cyclic codes (null otherwise): per side, the block maps a, b,
    // ainv of check r*P + x meeting variable cc*P + (a*x + b mod P), as
    // [side][a | b | ainv][r*L + cc]; the FP64 kernels form a variable's
    // message rows from them instead of loading row tables
    const int *qc_tab = nullptr;
    int qc_P = 0, qc_J = 0, qc_L = 0;
};

// Quasi-cyclic structure: check (r*P + x) of side g meets variable
// cc*P + (a*x + b mod P) for every column block cc (jointbp_ets.cpp:636-645).
// ainv/b are kept per (g, r, cc) for the variable-side inverse map
// x = ainv*(y - b) mod P.
constexpr int kQcMaxBlocks = 64;
// Largest LDS staging area of the FP64 kernels' quasi-cyclic variable pass:
// two workgroups still share a compute unit's 160 KB.
constexpr size_t kF64MaxLds = 78 * 1024;
struct QcParams {
    int P = 0, J = 0, L = 0;
    int a[2][kQcMaxBlocks];     // [side][r*L + cc]
    int b[2][kQcMaxBlocks];
    int ainv[2][kQcMaxBlocks];
};

// Dense reduced-row-echelon basis of one side's row space, for the
// stabilizer-residual test (membership is basis independent).
struct DevBasis {
    int rank = 0;
    int words = 0;                     // 32-bit words per row (= wn)
    const uint32_t *rows = nullptr;    // rank x words
    const int *pivot = nullptr;        // rank: pivot column of each row
};

// Logical-operator representatives of one side as sparse rows: a residual
// part r with zero syndrome is a stabilizer iff every row has even overlap
// with r.
struct DevLogicals {
    int64_t rows = 0;
    const int64_t *off = nullptr;  // rows + 1
    const int *qubit = nullptr;    // off[rows]
};

// Per-trial bit planes resident in HBM (one row per trial in a chunk).
struct TrialPlanes {
    uint32_t *err_x = nullptr;   // chunk x wn
    uint32_t *err_z = nullptr;
    uint32_t *syn_x = nullptr;   // chunk x wm[0]
    uint32_t *syn_z = nullptr;   // chunk x wm[1]
    uint32_t *est_x = nullptr;   // chunk x wn
    uint32_t *est_z = nullptr;
    int *iters = nullptr;        // chunk
    uint8_t *syn_ok = nullptr;   // chunk
    uint8_t *success = nullptr;  // chunk
    uint8_t *exact = nullptr;    // chunk
    // decode order: the k-th trial a decoder pulls is order[k] (heaviest
    // syndrome first); null = index order.  Results do not depend on it.
    const int *order = nullptr;  // chunk
};

// Channel constants derived from the prior over (I, X, Z, Y).
struct Prior {
    double p[4];
};

// Sampler classification of one raw 64-bit draw d: state I when d < t[0],
// X when d < t[1], Z when d < t[2], else Y (all_below bit k: every draw is
// below threshold k).  Computed on the host for a given p.
struct DrawThresholds {
    uint64_t t[3];
    uint32_t all_below;
};
DrawThresholds draw_thresholds(double p);

// FP32 decoder channel constants of one prior: log2 prior (I, X, Z, Y), then
// (pi_I, pi_Z, pi_X, pi_Y) for the X-side extrinsic term, then
// (pi_I, pi_X, pi_Z, pi_Y) for the Z side.
constexpr int kPrior32 = 12;
void prior32_from(const Prior &pr, float out[kPrior32]);

// Channel of each trial of a chunk: slot i uses entry (phase + i) % n of the
// device tables.  One entry for a single p; entry idx % n for an interleaved
// list of p values (phase = first trial index % n); one per slot for a
// training batch (n = batch, phase 0).
struct NoiseTab {
    const DrawThresholds *th = nullptr;  // sampler thresholds [n]
    const double *prior = nullptr;       // normalized prior [n][4]
    const float *prior32 = nullptr;      // [n][kPrior32]
    int n = 1;
    int phase = 0;
    __host__ __device__ int entry(int slot) const { return (phase + slot) % n; }
};

// Launchers (defined in the .hip files).
// mt_state: scratch of mt_cap x 312 64-bit words for the generators'
// seeded states (count above mt_cap runs in pieces).
void launch_sample(const DevGraph &g, TrialPlanes tp, uint64_t base_seed, int64_t first_trial,
                   int count, const NoiseTab &nt, hipStream_t st, uint64_t *mt_state, int mt_cap);
constexpr int kMtWords = 312;  // 64-bit words of one generator state
void launch_syndromes(const DevGraph &g, TrialPlanes tp, int count, hipStream_t st);
// order[0..count) = trial indices by decreasing syndrome weight (a counting
// sort on a 12-bit key; keys: count ints of scratch).  Long decodes tend to
// be the heavy ones, so handing those out first shortens a launch's tail.
void launch_order_by_weight(const DevGraph &g, TrialPlanes tp, int count, int *keys, int *order,
                            hipStream_t st);
void launch_residual(const DevGraph &g, DevBasis bx, DevBasis bz, TrialPlanes tp, int count,
                     hipStream_t st);
// Success test with logical operators: lx against the X part of est ^ err,
// lz against the Z part.
void launch_logical(const DevGraph &g, DevLogicals lx, DevLogicals lz, TrialPlanes tp, int count,
                    hipStream_t st);
// freeze: the reference's --freeze-syn
void launch_decode_f64(const DevGraph &g, TrialPlanes tp, int count, const NoiseTab &nt, int max_iter,
                       double damping, int slots, int block, double *scratch, int *counter,
                       hipStream_t st, bool freeze = false);
size_t decode_f64_scratch_bytes(const DevGraph &g, int slots);
// d_tables: device ints [2][J][L] ainv followed by [2][J][L] b.
struct LearnedInfer;
// learned == nullptr: plain BP.  Otherwise learned->T learned iterations,
// then plain BP restarted on trials that did not converge.  The min-sum
// kind runs as two launches (the learned phase queues its unconverged
// trials in fb, count + 4 ints of device scratch, for the plain one).  Plain
// BP with 0 < cap_iter < max_iter (and fb) also runs as two: every trial up
// to cap_iter iterations, then the ones still unconverged from the prior up
// to max_iter.  With fb_st the second launch goes on that stream, after
// phase_end (recorded on st when the first launch ends; required then).
void launch_decode_qc_f32(const DevGraph &g, const QcParams &qc, const int *d_tables,
                          TrialPlanes tp, int count, const NoiseTab &nt, int max_iter, int check_warmup,
                          int check_interval, int slots, int *counter, hipStream_t st,
                          const LearnedInfer *learned = nullptr, int *fb = nullptr,
                          hipEvent_t phase_end = nullptr, hipStream_t fb_st = nullptr,
                          int cap_iter = 0);
bool qc_f32_supported(const QcParams &qc, int device, char *why, size_t len);
void launch_fill_i32(int *ptr, int value, int count, hipStream_t st);
void launch_unpack_states(const uint32_t *xbits, const uint32_t *zbits, int words, int n,
                          int count, int *out, hipStream_t st);
void launch_pack_bits(const int *vals, int len, int words, int count, uint32_t *out,
                      hipStream_t st);

}  // namespace jbp
