/*
 * jbp.h — C ABI of the MI355X joint belief-propagation decoder library
 * (libjbp.so).  Plain C types only: pointers, sizes, status codes.
 *
 * It replaces the reference's C++ GPU boundary jointbp_cuda.h:1-93
 * (CudaBPGraph / CudaBPResult / cuda_bp_create / cuda_bp_decode /
 * cuda_bp_destroy) and moves the Monte-Carlo inner loop of
 * jointbp_ets.cpp:5931-6207 (sampling, syndromes, decode, stabilizer-residual
 * check) onto the device as a batched call.
 *
 * Threading: a context owns a HIP stream and device buffers and is not
 * thread-safe; use one context per host thread.
 */
#ifndef JBP_H_
#define JBP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Every entry point returns one of these; on failure the
 * caller's error buffer (when given) holds a NUL-terminated message, the way
 * the reference fills *error_out (jointbp_cuda.cu:565-573). */
#define JBP_OK 0
#define JBP_ERR_ARG 1         /* invalid argument / size mismatch           */
#define JBP_ERR_DEVICE 2      /* HIP runtime failure                        */
#define JBP_ERR_UNSUPPORTED 3 /* option this build does not run on the GPU  */
#define JBP_ERR_NOMEM 4       /* device allocation failed                   */

/* Arithmetic of the decoder. */
#define JBP_FP32 32 /* production: FP32 log-likelihood-ratio messages       */
#define JBP_FP64 64 /* parity: FP64 probability messages, reference order   */

typedef struct jbp_ctx jbp_ctx;

/* Tanner-graph description; field for field the reference's CudaBPGraph
 * (jointbp_cuda.h:10-30) as built by build_cuda_graph (jointbp_ets.cpp:937-1016):
 * check-major edge numbering; x_var_edges lists each variable's edges in
 * ascending check order.  The arrays are copied during jbp_create; the
 * caller keeps ownership. */
typedef struct jbp_graph_desc {
    int nvars, mX, mZ, x_edges, z_edges;
    const int *x_check_offsets; /* mX+1 */
    const int *x_check_edges;   /* x_edges (identity in the reference)       */
    const int *x_edge_var;      /* x_edges                                   */
    const int *x_edge_check;    /* x_edges                                   */
    const int *x_edge_pos;      /* x_edges                                   */
    const int *x_var_offsets;   /* nvars+1                                   */
    const int *x_var_edges;     /* x_edges                                   */
    const int *z_check_offsets;
    const int *z_check_edges;
    const int *z_edge_var;
    const int *z_edge_check;
    const int *z_edge_pos;
    const int *z_var_offsets;
    const int *z_var_edges;
} jbp_graph_desc;

/* Per-decode result; the scalar fields of CudaBPResult (jointbp_cuda.h:32-43).
 * The estimate goes to the caller's est_out array instead of a std::vector. */
typedef struct jbp_decode_result {
    int iterations;     /* k at first syndrome match, else max_iter         */
    int check_count;    /* syndrome tests performed                         */
    int syndrome_match; /* 1 when both syndromes of est match the input     */
    double kernel_ms, memcpy_ms, check_kernel_ms, check_memcpy_ms, init_kernel_ms, host_ms;
} jbp_decode_result;

/* cuda_bp_create (jointbp_cuda.h:48, jointbp_cuda.cu:589-642): validate the
 * graph, select the device, upload the graph and allocate working buffers.
 * Codes whose checks are built from affine permutation blocks of one size P
 * (the params-file codes, jointbp_ets.cpp:98-160) are detected from the CSR
 * and get the LDS-resident FP32 decoder; any other code runs the generic
 * CSR decoders.  On failure *out is NULL. */
int jbp_create(const jbp_graph_desc *graph, int device_id, jbp_ctx **out, char *err, size_t errlen);

/* cuda_bp_destroy (jointbp_cuda.h:49, jointbp_cuda.cu:644-660); NULL is a no-op.
 * Learned models created on the context must be destroyed first. */
void jbp_destroy(jbp_ctx *ctx);

/* cuda_bp_decode (jointbp_cuda.h:50-63, jointbp_cuda.cu:681-1104): decode one
 * syndrome pair (sx: mX ints, sz: mZ ints, nonzero = 1) with the given prior
 * over (I, X, Z, Y).  est_out receives nvars states 0..3.  precision is
 * JBP_FP32 or JBP_FP64.  check_warmup / check_interval skip syndrome tests as
 * the reference's flags do (FP32 only; FP64 tests every iteration).
 * freeze_syn (the reference's --freeze-syn, jointbp_ets.cpp:1884-1901,
 * 2526-2535) runs on the FP64 decoder; with FP32 it returns
 * JBP_ERR_UNSUPPORTED so that the caller falls back, as the reference
 * caller does on any decode failure (jointbp_ets.cpp:5961-5985). */
int jbp_decode(jbp_ctx *ctx, const int *sx, const int *sz, const double prior[4], int max_iter,
               int check_warmup, int check_interval, int measure_costs, int use_graph,
               int freeze_syn, double damping, int precision, jbp_decode_result *out,
               int *est_out, char *err, size_t errlen);

/* Batched form of jbp_decode: `count` independent syndrome pairs, packed as
 * bits (sx_bits: count x ceil(mX/32) uint32 words, bit c of word c/32; sz
 * likewise).  Outputs per trial: iterations, syndrome_match and, when
 * est_out is non-NULL, count x nvars states. */
int jbp_decode_batch(jbp_ctx *ctx, int count, const uint32_t *sx_bits, const uint32_t *sz_bits,
                     const double prior[4], int max_iter, double damping, int precision,
                     int *iterations_out, uint8_t *syn_match_out, int *est_out, char *err,
                     size_t errlen);

/* Allocation-free form of jbp_decode_batch (FP32, LDS-resident decoder;
 * permutation-block codes only).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory when they decode (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs and outputs are the caller's DEVICE buffers: d_sx_bits / d_sz_bits
 * count x ceil(mX/32) / ceil(mZ/32) uint32 words (bit c of word c/32);
 * d_iterations, d_syn_match count entries; d_est_x / d_est_z the estimate's
 * X and Z bit planes, count x ceil(nvars/32) words.  Per decode the results
 * equal jbp_decode_batch's (damping 0).  Work is queued on `stream` (a
 * hipStream_t; NULL = the null stream) and the call returns without waiting.
 * It calls no allocator: the work counter, channel constants and decode order
 * live in a per-device static array (512 B + 8 B per trial, set up by
 * jbp_create), so count is at most 65,536 (larger: JBP_ERR_ARG).  Calls on
 * one GPU share that array: calls that overlap on one GPU race and must not
 * be issued. */
int jbp_decode_batch_noalloc(jbp_ctx *ctx, int count, const uint32_t *d_sx_bits,
                             const uint32_t *d_sz_bits, const double prior[4], int max_iter,
                             int *d_iterations, uint8_t *d_syn_match, uint32_t *d_est_x,
                             uint32_t *d_est_z, void *stream, char *err, size_t errlen);

/* Monte-Carlo configuration: the inner loop of the reference's --simulate
 * path (jointbp_ets.cpp:5931-6207), post-processing off or on. */
typedef struct jbp_sim_cfg {
    uint64_t base_seed;  /* --seed                                            */
    int64_t first_trial; /* trial index of the first trial (reference: 1)     */
    int64_t n_trials;    /* trials to run                                     */
    double p;            /* depolarizing rate: X, Y, Z each with p/3          */
    int max_iter;        /* --max-iter                                        */
    double damping;      /* --damping                                         */
    int precision;       /* JBP_FP32 or JBP_FP64                              */
    int check_warmup;    /* --cuda-check-warmup (FP32)                        */
    int check_interval;  /* --cuda-check-interval (FP32)                      */
    int slots;           /* trials resident at once; 0 = library default      */
    int block;           /* threads per block for the generic decoders; 0 = tuned */
    int64_t min_fails;   /* >0: stop at the first trial index where the
                            failure count reaches this (reference 6409-6411) */
    const double *p_list; /* n_p > 0: trial index idx uses p_list[idx % n_p]
                             instead of p (interleaved mixed-noise runs)     */
    int n_p;              /* 0 = single p; at most JBP_MAX_P_LIST            */
    int engine;           /* FP32 decoder: JBP_ENGINE_AUTO, _LDS or _HBM     */
    int schedule;         /* 0: decode trials heaviest syndrome first (shorter
                             launch tails), 1: in index order; per-trial
                             results are the same either way                 */
    int graph;            /* 1: replay the HBM engine's steps from captured
                             graphs (0: launch them; measured 2% faster)     */
    int pp;               /* 1: the reference's post-processing (ETS / flip /
                             OSD, default on there; off = --no-pp) with the
                             tables installed by jbp_set_pp_tables          */
    int flip_hist;        /* --flip-hist (reference default 5)              */
    int chunk;            /* LDS engine, runs longer than one launch: trials
                             per persistent launch (0 = 8192, at most 65536);
                             launches are pipelined over three streams, and
                             plain BP decodes a long run's stragglers again
                             on the third (per-trial results unchanged)      */
    int freeze_syn;       /* --freeze-syn (jointbp_ets.cpp:1884-1901,
                             2526-2535): FP64, post-processing off           */
} jbp_sim_cfg;

#define JBP_MAX_P_LIST 4096

/* FP32 decoders.  LDS: one trial per compute unit with all messages on chip,
 * for the 768 x (3 x 12) permutation-block codes.  HBM: any code; 64 trials
 * interleaved per wavefront, messages in device memory (slots = trials
 * resident, a multiple of 64).  AUTO picks LDS when it applies.  Damping
 * (the reference's scaled_mix of each variable-to-check message with the
 * previous one, jointbp_ets.cpp:1903-1911) runs on the HBM decoder, which
 * keeps the previous messages; AUTO picks it when damping > 0. */
#define JBP_ENGINE_AUTO 0
#define JBP_ENGINE_LDS 1
#define JBP_ENGINE_HBM 2

/* Per-trial outcome (structure of arrays; any pointer may be NULL). */
typedef struct jbp_trial_out {
    int *iterations;          /* n_trials                                    */
    uint8_t *syn_match;       /* n_trials                                    */
    uint8_t *success;         /* syndrome match and residual a stabilizer    */
    uint8_t *exact;           /* est == err                                  */
    int *err;                 /* n_trials x nvars, sampled error (optional)  */
    int *est;                 /* n_trials x nvars, final estimate (optional) */
    uint8_t *pp_flags;        /* n_trials, JBP_PPF_* bits (optional)         */
} jbp_trial_out;

/* Per-trial post-processing flags (jbp_trial_out.pp_flags). */
#define JBP_PPF_BP_OK 1        /* BP alone matched the syndrome             */
#define JBP_PPF_PP_SUCCESS 2   /* success owed to a repair                  */
#define JBP_PPF_ETS_SAVED 4    /* ... by ETS                                */
#define JBP_PPF_FLIP_SAVED 8   /* ... by a flip set                         */
#define JBP_PPF_OSD_SAVED 16   /* ... by OSD                                */
#define JBP_PPF_ETS_USED 32    /* an ETS repair was applied                 */
#define JBP_PPF_ETS6 64        /* ... from the 6-variable table             */
#define JBP_PPF_ETS12 128      /* ... from the 12-variable table            */

/* Aggregates in the reference's Run Summary terms (jointbp_ets.cpp:6414-6449). */
typedef struct jbp_run_stats {
    int64_t done, failures, bp_failures, exact, total_iters;
    double device_ms;  /* device time of the decode launches (HIP events) */
    double elapsed_ms; /* wall time of the whole call                      */
    /* post-processing counters, as the Run Summary names them
     * (jointbp_ets.cpp:6113-6228); zero with pp off */
    int64_t pp_success, flip_pp_success, osd_pp_success, stab_success;
    int64_t ets_used, ets_saved, ets6_used, ets6_saved, ets12_used, ets12_saved;
    int64_t pp_trials; /* trials handed to the post-processing stage     */
    double pp_ms;      /* wall time of that stage                         */
} jbp_run_stats;

/* Sample, decode and judge trials first_trial .. first_trial+n_trials-1 on
 * the device: per-trial seed splitmix64(base ^ idx*0x9e3779b97f4a7c15),
 * mt19937_64 + libstdc++ generate_canonical draws, depolarizing thresholds,
 * syndromes, decode, stabilizer-residual test.  Bit-exact with the reference
 * for the sampled errors, syndromes and residual tests. */
int jbp_simulate(jbp_ctx *ctx, const jbp_sim_cfg *cfg, jbp_trial_out *per_trial,
                 jbp_run_stats *stats, char *err, size_t errlen);

/* jbp_simulate for up to two codes in one engine: code c's trials (cfgs[c],
 * context ctxs[c], one context per code, all on one device) are sampled, then
 * decoded together by the HBM-resident FP32 decoder, each engine step's
 * kernels covering every code's resident trials, then judged by each code's
 * success test.  Per trial the outcome equals jbp_simulate of that code with
 * engine JBP_ENGINE_HBM.  FP32 only, at most 65,536 trials per code, no
 * post-processing, freeze-syn or min_fails.  per_trial (may be NULL) and
 * stats are arrays of n_codes; device_ms is the shared decode's time.  No
 * counterpart in the reference (its decoder is generic over check lists,
 * jointbp_ets.cpp:1943-1948, but runs one code per process). */
int jbp_simulate_mixed(jbp_ctx *const *ctxs, const jbp_sim_cfg *cfgs, int n_codes,
                       jbp_trial_out *per_trial, jbp_run_stats *stats, char *err, size_t errlen);

/* ---------------------------------------------------------------------------
 * Post-processing (PP) of unconverged decodes: the reference's ETS lookup,
 * flip-set solve and ordered-statistics repair (jointbp_ets.cpp:2736-3074 in
 * the decode loop from iteration 100, 5993-6197 after it).  With cfg.pp = 1,
 * jbp_simulate decodes as usual, then re-decodes the trials still
 * unconverged after min(99, max_iter - 1) iterations with FP64 BP on the
 * device one iteration per launch, running the reference's trigger logic and
 * repairs on host threads between launches; per trial the outcome is the
 * reference CPU decoder's.
 * ------------------------------------------------------------------------- */
typedef struct jbp_pp_tables jbp_pp_tables;

/* Empty set of trapping-set tables (host memory, no device needed). */
int jbp_pp_tables_create(jbp_pp_tables **out);
void jbp_pp_tables_free(jbp_pp_tables *t);

/* Adds one table group, read from the reference's ETS text files (the X
 * side from path_x, the Z side from path_z, either may be NULL):
 * "size"-variable trapping sets of the code described by graph, variables
 * "Vc:i" = c*P + i, checks "Cr:i" = r*P + i (read_ets_file,
 * jointbp_ets.cpp:3199-3258).  counts[s] = entries loaded for side s, 0 for
 * an empty file, -1 for a missing one.  The group is kept when either side
 * has entries; groups are tried in the order they were added (the
 * reference's order is 6, 8, 8p4, 10, 12, ..., 22, jointbp_ets.cpp:
 * 5309-5327). */
int jbp_pp_tables_add(jbp_pp_tables *t, const jbp_graph_desc *graph, int64_t P, int size,
                      const char *label, const char *path_x, const char *path_z, int counts[2],
                      char *err, size_t errlen);

/* Installs a copy of the tables on the context (NULL: none). */
int jbp_set_pp_tables(jbp_ctx *ctx, const jbp_pp_tables *t, char *err, size_t errlen);

/* One repair on one estimate, on the host (no device): method 0 = ETS over
 * the groups of t, 1 = flip set `vars` (n_vars entries), 2 = OSD with the
 * per-qubit |LLR| abs_llr, 3 = the same OSD with abs_llr holding the
 * normalised marginals [nvars][4] the |LLR| come from (the form the device
 * stage hands over); side 0 repairs the X part against the X-check
 * syndrome, 1 the Z part.  est (nvars states 0..3) is updated in place when
 * the repair applies; *applied = 1 then (for ETS, the group's size). */
int jbp_pp_repair(const jbp_graph_desc *graph, const jbp_pp_tables *t, int method, int side,
                  const int *syndrome, const double *abs_llr, const int *vars, int n_vars,
                  int *est, int *applied, char *err, size_t errlen);

/* jbp_decode_batch with post-processing (the single-decode modes of the
 * reference with PP on, jointbp_ets.cpp:5487-5710): same inputs plus
 * flip_hist; pp_out (may be NULL) receives per decode the JBP_PPF_* bits,
 * where the "saved" bits mean that the repair produced the syndrome match
 * (whether that is a success depends on the caller's residual test). */
int jbp_decode_batch_pp(jbp_ctx *ctx, int count, const uint32_t *sx_bits,
                        const uint32_t *sz_bits, const double prior[4], int max_iter,
                        double damping, int precision, int flip_hist, int *iterations_out,
                        uint8_t *syn_match_out, uint8_t *pp_out, int *est_out, char *err,
                        size_t errlen);

/* FP64 decode of count syndrome pairs (per decode the reference CPU
 * decoder's iterations and estimate) that also returns the per-qubit |LLR|
 * of the X and Z parts of the last iteration's marginals (abs_llr_x /
 * abs_llr_z, jointbp_ets.cpp:2483-2493: the ordering OSD uses), formed with
 * the C library's log from the device's FP64 marginals, so they are the
 * reference's values bit for bit.  llr_out: count x 2 x nvars doubles (X
 * then Z per decode); any output may be NULL.  max_iter >= 1. */
int jbp_decode_llr(jbp_ctx *ctx, int count, const uint32_t *sx_bits, const uint32_t *sz_bits,
                   const double prior[4], int max_iter, double damping, int *iterations_out,
                   uint8_t *syn_match_out, int *est_out, double *llr_out, char *err,
                   size_t errlen);

/* Sample only: errors of trials first..first+count-1 (count x nvars states)
 * and their syndromes (count x mX, count x mZ ints; either may be NULL). */
int jbp_sample(jbp_ctx *ctx, uint64_t base_seed, int64_t first_trial, int count, double p,
               int *err_out, int *sx_out, int *sz_out, char *err, size_t errlen);

/* Autotuner.  Runs cfg once per candidate launch configuration of the
 * decoder cfg selects and keeps the fastest (device time) for later calls on
 * this context that leave slots / block at 0:
 *   FP32 HBM engine: resident trials S = 64, 128, ... (up to n_trials and an
 *   8 GB state cap; the best S depends on the workload: small runs with
 *   long-running failures favour small S);
 *   FP64: threads per block {64, 128, 256, 512} x resident trials {1, 2, 4}
 *   per compute unit.
 * The LDS engine has no free parameter (one 768-thread workgroup per compute
 * unit is fixed by the code); it returns best = {0, 0}, *n_cand = 0.
 * best[0] = slots, best[1] = block; cand_ms (optional, *n_cand entries on
 * input) receives each candidate's device ms, *n_cand their number.  No
 * counterpart in the reference (its kernels use fixed 256-thread blocks,
 * jointbp_cuda.cu:681-1104). */
int jbp_tune(jbp_ctx *ctx, const jbp_sim_cfg *cfg, int best[2], double *cand_ms, int *n_cand,
             char *err, size_t errlen);

/* Success test by logical operators instead of the row basis built at
 * create (which is skipped for codes too large to eliminate densely on the
 * host): side 0 rows span ker(Hz) modulo rowspace(Hx) and are checked
 * against the X part of est ^ err, side 1 rows span ker(Hx) modulo
 * rowspace(Hz) against the Z part; a syndrome-matching estimate succeeds
 * iff every row has even overlap.  Row i = qubits[row_off[i] ..
 * row_off[i+1]).  Used once both sides are set; side = -1 clears both.
 * Replaces the in_row_space test of jointbp_ets.cpp:904-920 (membership in
 * a row space is orthogonality to its dual, so both tests agree). */
int jbp_set_logicals(jbp_ctx *ctx, int side, int64_t n_rows, const int64_t *row_off,
                     const int *qubits, char *err, size_t errlen);

/* Row-space membership of count bit vectors (count x nvars ints, nonzero = 1)
 * in the row space of Hx (side 0) or Hz (side 1): out[i] = 1 when vector i
 * is a sum of rows (the in_row_space test of jointbp_ets.cpp:904-920). */
int jbp_in_row_space(jbp_ctx *ctx, int side, int count, const int *bits, uint8_t *out, char *err,
                     size_t errlen);

/* Device time of the most recent launch of each kernel family, for
 * profiling: which = 0 sample, 1 decode, 2 residual check.  Returns ms and
 * the launch count since the context was created.  which = 3: the decode
 * time again, with the number of engine steps (one flooding iteration of
 * every resident trial) of the last generic-decoder call in *launches.
 * which = 4: the learned min-sum decoder's own launch within the last
 * decode (without the plain-BP launch over the trials it left), and the
 * count of such launches. */
double jbp_last_kernel_ms(jbp_ctx *ctx, int which, int64_t *launches);

/* Context facts: out[0] = 1 when the LDS-resident FP32 decoder applies,
 * out[1] = permutation size P, out[2] = J, out[3] = L, out[4] = device id,
 * out[5] = compute units, out[6] = tuned threads per block. */
int jbp_info(jbp_ctx *ctx, int64_t out[8]);

/* ---------------------------------------------------------------------------
 * Learned (neural-weighted) BP.  No counterpart in the reference program; it
 * extends the reference's joint BP with a weight per (side, permutation
 * block, iteration) on every variable-to-check and every check-to-variable
 * message and a 2-16-1 MLP that damps the check-to-variable messages from
 * (iteration / T, unsatisfied-check fraction).  Parameter vector (doubles):
 *   wc[2][nb][T] | wv[2][nb][T] | mlp[65] (W1[16][2], b1[16], W2[16], b2),
 * nb = J*L blocks per side; the block of the edge at position cc of check
 * r*P + x is r*L + cc.  All weights 1 and damping off is plain BP.
 * Needs a permutation-block code (jbp_info out[0] == 1).
 * ------------------------------------------------------------------------- */
typedef struct jbp_learned jbp_learned;

/* T learned iterations (1..64); parameters start at weights 1 and an MLP
 * whose damping is close to 0.  The kind is JBP_LEARNED_SUM_PRODUCT. */
int jbp_learned_create(jbp_ctx *ctx, int T, jbp_learned **out, char *err, size_t errlen);

/* Model kinds.  SUM_PRODUCT: the reference's joint BP (jointbp_ets.cpp:
 * 2373-2519) with the weights and damping above.  MIN_SUM: the same model
 * with max-product updates (check: sign product times the minimum
 * magnitude over the other edges, clamped at 64 bits; joint variable term
 * max(ln a + L, ln b) - max(ln c + L, ln d) for ln((a e^L + b) / (c e^L +
 * d)); hard decisions from the signs of the X / Z marginals): the weights
 * act as learned normalisation, and the inference decoder runs it on FP16
 * pairs with no transcendental function, several times cheaper per
 * iteration than sum-product BP.  LAYERED_MIN_SUM: the min-sum model on a
 * layered schedule (block row r of both sides is layer r; an iteration
 * updates the layers in turn, each from the messages the layers before it
 * left), which converges in fewer iterations; no damping (use_mlp must be
 * 0); an iteration counts as converged when its layers all start from the
 * same estimate and every check accepts it.  Plain BP stays the fallback. */
#define JBP_LEARNED_SUM_PRODUCT 0
#define JBP_LEARNED_MIN_SUM 1
#define JBP_LEARNED_LAYERED_MIN_SUM 2
int jbp_learned_create_kind(jbp_ctx *ctx, int T, int kind, jbp_learned **out, char *err,
                            size_t errlen);
/* The model's kind (JBP_LEARNED_*), -1 for NULL. */
int jbp_learned_kind(const jbp_learned *model);
/* Destroy a model before the context it was created on. */
void jbp_learned_destroy(jbp_learned *model);
int jbp_learned_param_count(const jbp_learned *model);
/* use_mlp = 0 switches damping off (d = 0). */
int jbp_learned_set_params(jbp_learned *model, const double *params, int use_mlp, char *err,
                           size_t errlen);
int jbp_learned_get_params(const jbp_learned *model, double *params);

/* Loss (mean binary cross-entropy of the per-qubit X/Z marginals against
 * the error, averaged over iterations 1..T) and its gradient for explicit
 * errors (batch x nvars states 0..3, one p per trial), computed on the
 * device by the forward and adjoint kernels in FP32 or FP64.  conv_out
 * (may be NULL): first iteration whose estimate satisfies both syndromes
 * (0 = none).  grad (may be NULL): jbp_learned_param_count values. */
int jbp_learned_loss_grad(jbp_learned *model, int precision, int batch, const int *err_states,
                          const double *p, double *loss, double *grad, int *conv_out, char *err,
                          size_t errlen);

/* One training step on device-sampled trials first..first+batch-1 (trial
 * seeds as in jbp_simulate) with p uniform in [p_lo, p_hi] per trial:
 * FP32 forward + adjoint, then an Adam update (betas 0.9 / 0.999). */
int jbp_learned_train_step(jbp_learned *model, uint64_t base_seed, int64_t first_trial, int batch,
                           double p_lo, double p_hi, double lr, double *loss, char *err,
                           size_t errlen);

/* jbp_simulate with the learned decoder for the first T iterations; a trial
 * whose estimate has not satisfied both syndromes by then is decoded again
 * by plain FP32 BP from the channel prior for up to cfg->max_iter
 * iterations (its iteration count is T plus those).  FP32 only. */
int jbp_simulate_learned(jbp_ctx *ctx, const jbp_learned *model, const jbp_sim_cfg *cfg,
                         jbp_trial_out *per_trial, jbp_run_stats *stats, char *err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* JBP_H_ */
