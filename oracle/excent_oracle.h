/*
 * excent_oracle.h -- CPU restatement of the ExCent delta-push path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (bin/excent, the HIP
 * library, the host library) links, loads or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker or as the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed in the
 * header comment of excent_oracle.c.
 */
#ifndef EXCENT_ORACLE_H
#define EXCENT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PR = 0, ORC_KATZ = 1, ORC_EV = 2 };

/* ---- EGR binary CSR (README.md:136-144; include/StaticGraph.h:29-55) ---- */
/* Loads the whole file.  Returns 0 on success, -1 bad file, -2 bad header, -3 short file. */
int orc_egr_load(const char* path, int* n, uint64_t* m, uint64_t** idx, int32_t** adj);
void orc_free(void* p);
/* Writes an EGR: narrow (int32 offsets) when wide == 0, wide (sentinel -1, int64) otherwise. */
int orc_egr_write(const char* path, int n, uint64_t m, const uint64_t* idx, const int32_t* adj, int wide);

/* ---- chunk planner (include/StaticGraph.h:128-139; src/excent.cu:321-326) ---- */
uint64_t orc_chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges);
/* Writes chunk v0 values into v0_out (capacity max_chunks); returns the chunk count. */
int orc_plan_chunks(int n, const uint64_t* idx, uint64_t cap, int* v0_out, int max_chunks);

/* ---- init (src/excent.cu:41-47) ---- */
void orc_init(int n, const uint64_t* idx, int kind, float beta, float lambda, float* x, float* r);

/* ---- demand snapshot (src/excent.cu:49-98) ----
 * r[k] points to measure k's residual array (float32[n]).  hot may be NULL
 * (no hot skip); trace may be NULL.  bitmap has ceil(n/32) words; cnt has 2*nc entries
 * and is overwritten (not accumulated). */
void orc_snapshot(int n, int nm, const float* const* r, const float* eps,
                  const uint64_t* idx, const int* cv0, int nc, const uint32_t* hot,
                  uint32_t* bitmap, uint64_t* cnt, uint64_t* hotpend, uint16_t* trace);

/* ---- fetch plan (src/excent.cu:666-680) ----
 * mode[c]: 0 skip, 1 resident, 2 stream, 3 gather.  Returns need_bitmap. */
int orc_fetch_plan(int nc, const uint64_t* cnt, const uint8_t* resident, const uint8_t* loaded,
                   double tau, const uint64_t* chunk_edges, int hot_mode, int pass,
                   const int* hlo, const int* hhi, int8_t* mode, uint64_t* pend_v, uint64_t* demand);
/* pick_group (src/excent.cu:238) and the ids-vs-range test (src/excent.cu:639). */
int orc_pick_group(double avg);
int orc_use_list(uint64_t k, int nv, double list_frac);

/* ---- hot-set selection (src/excent.cu:428-513) ----
 * Degree threshold rule (no --hot-rank) when score == NULL; otherwise the
 * ranked rule.  ids/off (off has nh+1 entries) and bitmask (ceil(n/32)) are
 * outputs sized by the caller (n entries is always enough).  Returns nh. */
int orc_hot_select(int n, const uint64_t* idx, uint64_t budget, const uint16_t* score,
                   int32_t* ids, uint64_t* off, uint32_t* bitmask, uint64_t* deg_min);

/* ---- gather prefetch page/run counts (include/StaticGraph.h:90-114) ---- */
void orc_prefetch_counts(const int32_t* vids, int k, const uint64_t* idx, uint64_t nlist_off,
                         uint64_t page, int io_threads, uint64_t* pages, uint64_t* runs);

/* ---- tau from calibration timings (src/excent.cu:566-572) ---- */
double orc_tau(double t_stream, const double* t_g, const double* f_got);

/* ---- CPU asynchronous delta push, whole graph in memory (src/excent.cu:100-176, 656-760) ----
 * kinds[nm]; x[k], r[k] float32[n] already initialised.  Runs passes until
 * nothing is pending or max_passes; returns the pass count.  *pushed gets
 * the adjacency entries pushed (--count-pushes).  threads <= 0: OpenMP default.
 * time_limit > 0 stops after the pass during which that many seconds elapsed. */
int orc_run_push(int n, const uint64_t* idx, const int32_t* adj, int nc, const int* cv0,
                 int nm, float* const* x, float* const* r, const float* eps, const float* scale,
                 const int* divdeg, double list_frac, int max_passes, int threads,
                 double time_limit, uint64_t* pushed, double* seconds);

/* ---- FP64 synchronous oracle (tools/centrality_reference.cu:32-59, 184-214) ----
 * Pulls over in-neighbours (the transpose of the CSR) and divides by the
 * out-degree of the source, which equals the reference on symmetric graphs
 * and matches the push semantics on directed ones.  kind: ORC_PR, ORC_KATZ
 * (c = beta, s = alpha) or ORC_EV (power iteration, unit L2, sign by sum).
 * Returns the iteration count; *rel gets the final relative change. */
int orc_fp64_reference(int n, const uint64_t* idx, const int32_t* adj, int kind,
                       double alpha, double beta, int max_iter, double tol,
                       double* out, double* rel, double* lambda_out);

/* FP64 Jacobi with a per-vertex constant: y = b + s * sum_{u->w} x[u] (/ outdeg(u) when
 * divdeg) from x = b, until max|dy|/max|y| < tol.  Personalised PageRank is b = (1-d) t,
 * s = d, divdeg = 1 (no reference counterpart: the reference has no personalised mode). */
int orc_fp64_solve_b(int n, const uint64_t* idx, const int32_t* adj, const double* b, double s, int divdeg,
                     int max_iter, double tol, double* out, double* rel);

/* Weighted FP64 Jacobi: y[u] = b[u] + s * sum over edges e = (v -> u) of w[e] x[v], the
 * weights aligned with adj (CSR order).  The differentiable personalised PageRank's
 * forward is (idx, adj, w/S); its adjoint is the same call on the transpose with the
 * weights carried along.  No reference counterpart. */
int orc_fp64_solve_w(int n, const uint64_t* idx, const int32_t* adj, const double* w, const double* b, double s,
                     int max_iter, double tol, double* out, double* rel);

/* ---- constants (tools/excent_params.cu:105-170) ----
 * lambda1 by power iteration on the in-neighbour pull from `start`
 * (NULL = all ones, the reference's start), Katz mean with alpha = afac/lambda1,
 * eigenvector mean (v1.1)^2/n.  Returns power iterations used; *symmetric_bad =
 * sampled edges without a reverse edge (mt19937_64 seed 42 sample is not
 * reproduced; every edge is checked). */
int orc_params(int n, const uint64_t* idx, const int32_t* adj, const double* start, double afac,
               double* lambda, double* alpha, double* mean_katz, double* mean_ev, long* symmetric_bad);

/* ---- comparison metrics (tools/centrality_reference.cu:71-139) ---- */
void orc_compare(int n, const double* ref, const float* x, int unit_vectors,
                 double* l1_rel, double* l2, double* max_rel, int* top100);
void orc_rank(int n, const double* ref, const float* x, int unit_vectors,
              double* p_at /* [5]: P@10,100,1000,10000,100000 (-1 when k > n) */, double* tau1000);

/* FP64 residual of a state: sum_v |b_v + sum_{u->v} w_uv x_u - x_v| (the invariant
 * x + r = b + W x, so this is the true residual L1 of a push state). */
double orc_residual_l1(int n, const uint64_t* idx, const int32_t* adj, int kind,
                       double scale, double beta, const float* x);

#ifdef __cplusplus
}
#endif
#endif
