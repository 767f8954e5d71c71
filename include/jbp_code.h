/*
 * jbp_code.h — host-side code construction for libjbp (plain C ABI).
 *
 * Builds the Tanner graphs that jbp_create (jbp.h) consumes:
 *   - from a params file, the reference's format and construction
 *     (read_params jointbp_ets.cpp:162-180; build_block_matrix_hx/hz 98-160;
 *     build_check_neighbors 628-648; build_var_adjacency 922-935;
 *     build_cuda_graph 937-1016);
 *   - from explicit check lists;
 *   - a seeded hypergraph-product code of a random regular classical code
 *     (no reference counterpart: the mixed-workload code family).
 * Also the startup diagnostics the reference prints (rank and k
 * 5140-5165, degree stats 1025-1060, orthogonality 1070-1105, girth
 * 1107-1171).
 */
#ifndef JBP_CODE_H_
#define JBP_CODE_H_

#include <stddef.h>
#include <stdint.h>

#include "jbp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jbp_code jbp_code;

/* read_params + APM block construction.  Fails on a missing file, a bad
 * header or an odd L (reference read_params semantics). */
int jbp_code_from_params(const char *path, jbp_code **out, char *err, size_t errlen);

/* Check lists: `xdeg` entries per X check (row-major), -1 = absent entry,
 * likewise for Z. */
int jbp_code_from_checks(int nvars, int mx, int xdeg, const int *xflat, int mz, int zdeg,
                         const int *zflat, jbp_code **out, char *err, size_t errlen);

/* Hypergraph product of a seeded random (col_w, row_w)-regular m x n
 * classical parity-check matrix H with itself:
 *   Hx = [H (x) I_n | I_m (x) H^T],  Hz = [I_n (x) H | H^T (x) I_m],
 * n_q = n^2 + m^2 qubits, m*n checks per side. */
int jbp_code_hgp(int m, int n, int col_w, int row_w, uint64_t seed, jbp_code **out, char *err,
                 size_t errlen);

/* Logical-operator representatives of a hypergraph-product code, for the
 * success test without a row basis: side 0 spans ker(Hz) modulo
 * rowspace(Hx) (a residual X part r with Hx r = 0 is in rowspace(Hz) iff it
 * has even overlap with every one), side 1 spans ker(Hx) modulo
 * rowspace(Hz).  Sparse rows: row i is qubits[row_off[i] .. row_off[i+1]).
 * Call with row_off = qubits = NULL for the sizes.  JBP_ERR_UNSUPPORTED for
 * codes not built by jbp_code_hgp. */
int jbp_code_logicals(const jbp_code *code, int side, int64_t *n_rows, int64_t *n_entries,
                      int64_t *row_off, int *qubits);

void jbp_code_free(jbp_code *code);

/* out: n, mX, mZ, x_edges, z_edges, P, J, L (P/J/L are 0 for codes not read
 * from a params file). */
int jbp_code_dims(const jbp_code *code, int64_t out[8]);

/* Pointers into the code's own storage (valid while the code lives). */
int jbp_code_graph(const jbp_code *code, jbp_graph_desc *out);

/* The params-file affine pairs: a, b, c, d (L/2 each) into out[4*L/2]. */
int jbp_code_params(const jbp_code *code, int64_t *out);

/* Check lists of one side (0 = X, 1 = Z) as stored: m x deg, -1 = absent.
 * Returns deg (entries per check) or -1. */
int jbp_code_checks(const jbp_code *code, int side, int *flat);

/* Diagnostics. out: rank_x, rank_z, var_deg_min/max (X), check_deg_min/max
 * (X), var_deg_min/max (Z), check_deg_min/max (Z), odd_pairs, girth_x,
 * girth_z (girth only when with_girth, else 0). */
int jbp_code_diagnostics(const jbp_code *code, int with_girth, int64_t out[15]);

#ifdef __cplusplus
}
#endif

#endif /* JBP_CODE_H_ */
