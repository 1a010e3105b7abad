/*
 * excent.h -- C ABI of the ExCent host library (C++17/OpenMP, libexcent.so).
 *
 * The drop-in boundary of the reference is its process interface: the bin/excent
 * command line (src/excent.cu:242-316, paths relative to the reference root), the EGR
 * file format (README.md:136-144) and the float32 --dump files (src/excent.cu:767-777).
 * excent_main is that command line as a callable entry point; bin/excent is a
 * three-line main() around it.  The remaining functions expose the host planner's
 * pure decisions so tests can hold them against the oracle.
 */
#ifndef EXCENT_H
#define EXCENT_H

#include <stddef.h>
#include <stdint.h>

#include "excent_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bin/excent <graph.egr> [flags] (src/excent.cu:242-805).  Returns the process exit code. */
int excent_main(int argc, const char* const* argv);

/* Measurements of one pass loop (the reference's T_setup -> T_iter window,
 * src/excent.cu:581, 761) and its kernels. */
typedef struct excent_run_stats {
    double time_passes_s, time_snapshot_s, time_io_s, time_wait_s;
    int passes;
    uint64_t pushed_edges, bytes_stream, bytes_gather, launches;
    exc_stats kernel;
    /* FP64 sum of |r| per measure at the end of the loop (exc_residual_l1; 0 for batched runs) */
    double residual_l1[EXC_MAXM];
    /* why the loop ended: 0 nothing pending (the reference's rule), 1 every measure's
     * sum of |r| <= n * eps (--stop global), 2 the pass limit */
    int stop_reason;
    /* setup of the session, outside the pass loop's window: opening the graph file, the rest of
     * setup (offsets upload, caches, tau, the transpose for dense passes), and of that the
     * transpose build alone (0 when it has not been built) */
    double time_open_s, time_setup_s, time_transpose_s;
} excent_run_stats;

/* Session: bin/excent's setup once (graph open, offsets, caches, tau), then any number of
 * pass loops on the same resident state (no reference counterpart; for benchmarks that
 * time converged runs with the graph already in HBM).  reinit != 0 re-initialises the
 * measures and clears the kernel statistics first.  results() writes --dump / the report. */
int excent_session_open(int argc, const char* const* argv, void** session);
int excent_session_run(void* session, int reinit, excent_run_stats* stats);
int excent_session_results(void* session);
/* Sessions opened with --measure pr --ppr-column J: restart from column `column`'s teleport
 * set (fresh state, statistics cleared); the next run(reinit = 0) solves that column. */
int excent_session_set_column(void* session, int column);
/* The session's state after a run: x (which = 0) or r (which = 1) of measure slot / batched
 * column `column`, n floats into out (what --dump writes, without the file). */
int excent_session_values(void* session, int column, int which, float* out);
/* Number of vertices of the session's graph. */
int excent_session_vertices(void* session);
int excent_session_close(void* session);
const char* excent_session_error(void);
/* bin/excent_rmat: seeded R-MAT graph writer (no reference counterpart). */
int excent_rmat_main(int argc, const char* const* argv);
/* bin/egr_convert <in.mtx> <out.egr> (tools/efficient_converter.py:5-113). */
int excent_convert_main(int argc, const char* const* argv);
/* bin/centrality_reference (tools/centrality_reference.cu:141-222). */
int excent_reference_main(int argc, const char* const* argv);
/* bin/excent_params <graph.egr> <out.env> (tools/excent_params.cu:73-174). */
int excent_params_main(int argc, const char* const* argv);
/* bin/egr_info <graph.egr> (tools/egr_info.cpp:10-29): n, m, largest degree, a sampled
 * reverse-edge and self-loop count. */
int excent_egr_info_main(int argc, const char* const* argv);
/* bin/evict_file [--report] <file>... (tools/evict_file.c:36-65): drop files from the page
 * cache, or report their resident share. */
int excent_evict_main(int argc, const char* const* argv);
/* bin/gpu_ballast <budget_GB> [--device D] (tools/gpu_ballast.cu:13-41): hold device memory
 * until only the budget stays free, until SIGINT / SIGTERM. */
int excent_ballast_main(int argc, const char* const* argv);

/* Differentiable personalised PageRank (no reference counterpart).  x = (1-d) s + d A x with
 * A[u][v] = w_vu / S_v (raw positive weights w in CSR order, S_v = v's out-weight sum).
 * forward: the weighted delta push over the CSR from r = (1-d) s; backward: for
 * L = sum g_v x_v, the adjoint delta push over the transpose, then dL/ds = (1-d) lambda,
 * dL/dw (raw weights, CSR order) and dL/d(w/S) (normalised weights); any output may be
 * NULL.  Both pushes stop when every |r| <= eps. */
int excent_dpr_create(int device, int n, uint64_t m, const uint64_t* idx, const int32_t* adj, const float* w_raw,
                      float damp, float eps, void** handle);
int excent_dpr_forward(void* handle, const float* s, float* x_out);
int excent_dpr_backward(void* handle, const float* g, float* grad_s, float* grad_w, float* grad_wn);
int excent_dpr_passes(void* handle, int* forward, int* backward);
int excent_dpr_destroy(void* handle);
const char* excent_dpr_error(void);

/* Last error text of the excent_* helpers below on this thread. */
const char* excent_last_error(void);

/* src/excent.cu:321-326 */
uint64_t excent_chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges);
/* include/StaticGraph.h:128-139 on the graph file: fills up to max chunks; returns nc or < 0. */
int excent_plan_chunks(const char* path, uint64_t cap, int* v0, int* v1, uint64_t* e0, uint64_t* e1, int max);
/* EGR header and offsets through the product reader (include/StaticGraph.h:29-68). */
int excent_egr_header(const char* path, int* n, uint64_t* m, int* wide);
int excent_egr_offsets(const char* path, uint64_t* idx_out /* n+1 */);
int excent_egr_edges(const char* path, uint64_t e0, uint64_t e1, int32_t* out);
/* src/excent.cu:666-677; mode[nc] out; returns need_bitmap; pend_v/demand out. */
int excent_fetch_plan(int nc, const uint64_t* cnt, const uint8_t* resident, const uint8_t* loaded, double tau,
                      const uint64_t* chunk_edges, int hot_mode, int pass, const int* hlo, const int* hhi,
                      int8_t* mode, uint64_t* pend_v, uint64_t* demand);
/* src/excent.cu:238 with the given cap (32 in the reference). */
int excent_pick_group(double avg, int cap);
/* src/excent.cu:432-491: ids/off/bitmask sized by the caller (n, n+1, ceil(n/32)). Returns nh. */
int excent_hot_select(int n, const uint64_t* off, uint64_t budget, const uint16_t* score,
                      int32_t* ids, uint64_t* hoff, uint32_t* bitmask, uint64_t* deg_min);
/* src/excent.cu:566-572 */
double excent_tau(double t_stream, const double* t_g, const double* f_got);
/* include/StaticGraph.h:90-114 page/run counts of a gather (issues the prefetch). */
int excent_prefetch_counts(const char* path, const int32_t* vids, int k, int io_threads,
                           uint64_t* pages, uint64_t* runs);
/* The gather path's host staging (src/excent.cu:709-740) on the graph file: the lists of
 * vids[0..k) packed into edges_out with their offsets off_out[0..k], readahead counted
 * when dedup. */
int excent_stage_lists(const char* path, const int32_t* vids, int k, int dedup, int threads, uint64_t* off_out,
                       int32_t* edges_out, uint64_t* pages, uint64_t* runs);
/* The vertices with the most in-edges (up to count), estimated from `sample` of the graph
 * file's columns (all of them below 16M); returns how many were written to out. */
int excent_top_in_degree(const char* path, int count, double sample, int threads, int32_t* out);
/* src/excent.cu:711-720 */
int excent_gather_list(const uint32_t* bitmap, int v0, int v1, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif
