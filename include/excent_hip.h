/*
 * excent_hip.h -- C ABI of the MI355X (gfx950) device layer of ExCent.
 *
 * Plain pointers and sizes only; no HIP or torch types cross this boundary.
 * The host orchestrator (C++17/OpenMP, bin/excent), the Python test/bench
 * bindings and any FFI caller use exactly these entry points.  Every call
 * returns 0 on success or a negative EXC_E* code; exc_last_error() gives the
 * text of the most recent failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference root): the
 * reference has no library API; its device work is a set of __global__ kernels
 * launched from main() (src/excent.cu:41-220, 599-760).  Each entry point below
 * cites the kernel(s) and launch site(s) it stands for.
 *
 * Threading: a context is driven from one host thread (the orchestrator).
 * Staging host buffers returned by exc_slot_buffers may be filled by any number
 * of host threads between exc_slot_acquire and the submit call.
 */
#ifndef EXCENT_HIP_H
#define EXCENT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXC_MAXM 3
#define EXC_BURST_MAX 16     /* passes of one exc_run_passes call */
enum { EXC_PR = 0, EXC_KATZ = 1, EXC_EV = 2 };

enum {
    EXC_OK = 0,
    EXC_EINVAL = -1,   /* bad argument (size, index, null pointer) */
    EXC_ENOMEM = -2,   /* device or pinned allocation failed (recoverable for caches) */
    EXC_EHIP = -3,     /* HIP runtime error */
    EXC_ESTATE = -4    /* call out of order (e.g. push before offsets) */
};

/* Push kernel families.  EXC_KERNEL_LB: block-level load-balanced scatter (each lane
 * takes one adjacency slot of the block's pushed vertices).  EXC_KERNEL_GROUP: G lanes
 * per vertex, G in {1,2,4,8,16,32,64}, the width chosen from the average degree
 * (src/excent.cu:238, 629-633).  Both split vertices of degree above the hub
 * threshold into pieces pushed by whole waves.  EXC_KERNEL_TUNE: an autotuner picks per
 * degree class (the launch's group hint) among the load-balanced kernel and two group
 * widths by timing each alone on the class's first launches that push enough edges. */
enum { EXC_KERNEL_LB = 0, EXC_KERNEL_GROUP = 1, EXC_KERNEL_TUNE = 2 };

/* Per-run measure table (src/excent.cu:28-37, 340-351). */
typedef struct exc_measures {
    int nm;                 /* 1..3 measures in this run */
    int kind[EXC_MAXM];     /* EXC_PR / EXC_KATZ / EXC_EV */
    float eps[EXC_MAXM];    /* pending iff |r| > eps */
    float scale[EXC_MAXM];  /* push weight numerator: 0.85f, (float)alpha, (float)(1/lambda) */
    int divdeg[EXC_MAXM];   /* 1: weight = scale / outdeg (PageRank) */
} exc_measures;

typedef struct exc_config {
    int device;                 /* HIP device ordinal */
    int n;                      /* vertices */
    uint64_t m;                 /* edges */
    int nc;                     /* chunks */
    const int* cv0;             /* host array [nc]: first vertex of each chunk */
    uint64_t max_chunk_edges;   /* staging slot capacity (edges) */
    int max_chunk_verts;        /* staging slot capacity (vertices) */
    int trace;                  /* allocate the --trace-pending counters */
    int zero_copy;              /* staging buffers mapped into the device (read in place) */
    int push_kernel;            /* EXC_KERNEL_LB, EXC_KERNEL_GROUP or EXC_KERNEL_TUNE */
    int timing;                 /* record HIP events around kernels (exc_stats): 1 every kernel, 2 only the
                                   dense pull (k_pull / k_pull_sorted, the cheapest way to time it) */
    int batch_k;                /* 0: plain measures; 1..64: batched personalised PageRank with
                                   K columns ([n][K] state, damping M.scale[0], eps M.eps[0]) */
} exc_config;

typedef struct exc_chunk { int v0, v1; uint64_t e0, e1; } exc_chunk;

/* Kernel statistics gathered with timing on (HIP events on the launch streams). */
typedef struct exc_stats {
    double snapshot_ms;  uint64_t snapshot_launches;  double snapshot_bytes;
    double push_ms;      uint64_t push_launches;      double push_bytes;
    double piece_ms;     uint64_t piece_launches;     double piece_bytes;
    double other_ms;     uint64_t other_launches;
    uint64_t pushed_vertices;    /* vertices that pushed at least one measure */
    uint64_t pushed_edges;       /* adjacency entries pushed (the --count-pushes count) */
    uint64_t pushed_edge_measures;   /* sum over pushed edges of measures pushed */
    uint64_t pushed_vertex_measures; /* sum over pushed vertices of measures pushed */
    /* the push kernel split by vertex source: 0 range, 1 snapshot bitmap, 2 gathered
     * list, 3 hot list (bytes exclude hub pieces, which k_pieces pushes) */
    double src_ms[4];
    uint64_t src_launches[4];
    double src_bytes[4];
    /* dense passes run as fold + pull (exc_enable_pull) */
    double fold_ms, pull_ms, pull_hub_ms;
    uint64_t pull_passes;
    double fold_bytes, pull_bytes, pull_hub_bytes;
    /* batched pull: rows in flight per wave in use, and the tuner's time per demand edge
     * (ns) with 8 and with 16 (0 until that trial pass ran) */
    int pull_rows_in_flight;
    double pull_tune_ns_per_edge[2];
    /* plain dense pull: destinations per tile of the source-sorted pull (k_pull_sorted), or 0
     * when the pull runs over the in-lists in vertex order (k_pull) */
    int pull_tile;
    /* the dense pull split by width: entry w-1 for passes over the first w measure slots
     * (exc_set_active_measures; entry nm-1 holds the passes over every slot, and a batched
     * context counts all its passes in entry 0): time, launches, algorithmic bytes, and the
     * sorted tile it ran with (0: k_pull).  pull_ms and pull_bytes are their sums. */
    double pull_w_ms[3];
    uint64_t pull_w_launches[3];
    double pull_w_bytes[3];
    int pull_w_tile[3];
    /* push bursts (exc_run_passes) replayed from recorded HIP graphs */
    uint64_t burst_graph_launches;
    /* EXC_KERNEL_TUNE: per degree class c (the launch's group hint 2^c, c = 0..6), the three
     * candidates' measured time per pushed edge (ns; 0 = not timed yet) -- the load-balanced
     * kernel, G = 2^c and G = max(1, 2^c / 4) -- and the one kept (-1 while undecided) */
    double push_tune_ns_edge[7][3];
    int push_tune_best[7];
    /* batched: snapshots that read only the vertices the pushes since the last snapshot
     * added to (every chunk with pending vertices pushed in between; EXC_BATCH_FRONT=0 off) */
    uint64_t front_snapshots;
} exc_stats;

const char* exc_last_error(void);
int exc_device_count(int* count);
int exc_mem_info(int device, size_t* free_bytes, size_t* total_bytes);
int exc_device_info(int device, int* cus, int* integrated, char* name, int name_len);
/* bin/gpu_ballast (tools/gpu_ballast.cu:20-40): takes device memory on `device` until at most
 * keep_free bytes stay free (or the device refuses more), and holds it until
 * exc_release_memory(handle).  held_bytes (may be NULL) receives the amount taken. */
int exc_hold_memory(int device, size_t keep_free, void** handle, size_t* held_bytes);
int exc_release_memory(void* handle);
/* Waits for all work on the device (every stream of every context). */
int exc_device_synchronize(int device);

/* Context: device state for one run (src/excent.cu:335-398).  x and r live in HBM in a
 * vertex-major [n][S] layout (S = 1, 2 or 4 for 1, 2 or 3 measures) so the fused
 * push touches one 16-byte row per neighbour. */
int exc_create(const exc_config* cfg, const exc_measures* M, void** ctx);
int exc_destroy(void* ctx);

/* Offsets upload (src/excent.cu:360-369): host slice [v0, v0+cnt) of the n+1 offsets. */
int exc_upload_offsets(void* ctx, const uint64_t* h_idx, size_t v0, size_t cnt);

/* k_init (src/excent.cu:41-47, 370-372) for measure slot i of kind M.kind[i]; on a
 * batched context it zeroes the whole [n][K] state. */
int exc_init_measure(void* ctx, int i, float beta, float lambda);
/* Host-supplied state for measure slot i (any of x, r may be NULL to leave it). */
int exc_set_state(void* ctx, int i, const float* h_x, const float* h_r);
/* Copy measure slot i's x (which = 0) or r (which = 1), vertices [v0, v0+cnt), to host
 * (src/excent.cu:763-778). */
int exc_read_state(void* ctx, int i, int which, float* h_out, size_t v0, size_t cnt);

/* Saves (restore = 0) or restores (restore = 1) every measure's x and r with a device
 * copy, so a run can restart from its initial state without host work (no reference
 * counterpart; the reference initialises once per process, src/excent.cu:370-372). */
int exc_state_checkpoint(void* ctx, int restore);

/* Demand snapshot (k_snapshot, src/excent.cu:69-98, 660-665).  hot_skip != 0 excludes
 * pending hot-cached vertices (counted into *h_hotpend).  Blocks until the 2*nc counts
 * and hotpend are on the host.  The bitmap stays on the device. */
int exc_snapshot(void* ctx, int hot_skip, uint64_t* h_cnt, uint64_t* h_hotpend);
/* FP64 global residual sums of the last snapshot: out[i] = sum over v of |r_i[v]| for each
 * of the nm measures, each |r| widened to double and summed in FP64 (per wave or block, then
 * over 64 partial-sum slots on the host), for the state the snapshot tested -- after a dense
 * pass that also folded the next pass, the state before that fold.  No reference
 * counterpart: the reference stops on the per-vertex test alone (src/excent.cu:680); this
 * is the quantity a global L1 stop and the PageRank bound ||x - x*||_1 <= ||r||_1 / (1 - d)
 * are stated in.  EXC_ESTATE before the first snapshot of a state, and on batched contexts. */
int exc_residual_l1(void* ctx, double* out);
/* The largest |r_i[v]| per measure of the same snapshot state (the per-vertex stop holds for
 * measure i iff out[i] <= eps_i).  No reference counterpart; EXC_ESTATE as above. */
int exc_residual_max(void* ctx, float* out);
/* Allocation-free demand snapshot.  No counterpart in the reference (whose snapshot is
 * k_snapshot, src/excent.cu:69-98, on buffers main() allocates): it is exc_snapshot -- same
 * kernel, same outputs, bit for bit -- for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No context: the caller
 * passes device buffers -- r as [n][S] rows (S = 1, 2, 4 for nm = 1, 2, 3), the n+1 offsets,
 * cv0[nc], an optional hot bitmask of ceil(n/32) words (NULL: none), the bitmap output of
 * ceil(n/32) words, counts[2 nc + 1] (per-chunk pending count and out-degree sum, then
 * hot-pending) and l1[nm] (FP64 sum of |r| per measure) -- and eps[nm] on the host.  The
 * work is enqueued on `stream` (NULL: the null stream) and not waited for.  The count
 * accumulators and residual partial sums live in a static device array of 72 KiB per GPU,
 * held until the process ends: nc is limited to 1..4096 (EXC_EINVAL beyond).  Calls
 * clear it first; calls that overlap on one GPU share it and race.  Calls no hipMalloc,
 * hipFree or other allocating routine. */
int exc_snapshot_noalloc(int device, void* stream, int n, int nm, const float* d_r, const float* eps,
                         const uint64_t* d_idx, const int* d_cv0, int nc, const uint32_t* d_hot, uint32_t* d_bitmap,
                         unsigned long long* d_counts, double* d_l1);
/* Bitmap readback (src/excent.cu:678): ceil(n/32) words. */
int exc_bitmap_to_host(void* ctx, uint32_t* h_bitmap);
int exc_read_trace(void* ctx, uint16_t* h_out, size_t v0, size_t cnt);

/* GPU graph cache, chunks mode (src/excent.cu:410-421): one arena holding edges
 * [0, resident_edges) -- the resident chunks are always a prefix in file order.
 * Returns EXC_ENOMEM (and no arena) when the allocation fails. */
int exc_cache_chunks(void* ctx, uint64_t resident_edges);
/* Upload into the arena directly (setup-time preload), edges [e0, e0+cnt). */
int exc_cache_upload(void* ctx, const int32_t* h_edges, uint64_t e0, uint64_t cnt);
/* GPU graph cache, hot mode (src/excent.cu:492-512): ids[nh], off[nh+1], bitmask
 * ceil(n/32) words.  Returns EXC_ENOMEM when it cannot be allocated. */
int exc_cache_hot(void* ctx, int nh, const int32_t* ids, const uint64_t* off, const uint32_t* bitmask);
/* Upload hot lists directly (setup): hot list entries [j0, j0+cnt) flattened. */
int exc_cache_hot_upload(void* ctx, const int32_t* h_edges, uint64_t off0, uint64_t cnt);

/* Staging slots (src/excent.cu:224-233, 377-390, 599-601). */
int exc_slot_buffers(void* ctx, int slot, int32_t** h_edges, int32_t** h_vids, uint64_t** h_off);
/* Waits for the slot's previous work (cudaEventSynchronize at 600); *waited_s gets the wait. */
int exc_slot_acquire(void* ctx, int slot, double* waited_s);
/* Times a synchronous H2D copy of `bytes` from the slot's host edge buffer to its device
 * buffer.  tau calibration adds this PCIe leg to the storage timings, which the reference
 * measures storage-only (src/excent.cu:534-576). */
int exc_slot_copy_probe(void* ctx, int slot, uint64_t bytes, double* seconds);

/* Destinations whose pushed deltas are summed per block in LDS before they reach r (no
 * reference counterpart; the reference adds every delta with its own atomic,
 * src/excent.cu:118-123): up to 1024 vertex ids, normally those with the most in-edges,
 * whose float atomics would otherwise serialise at the memory side.  Applies to the
 * pieces of long out-lists (k_pieces).  count = 0 turns it off.  Any set is correct: a
 * destination's deltas arrive by the end of the launch either way. */
int exc_set_hub_destinations(void* ctx, const int32_t* ids, int count);

/* Weighted graphs (no reference counterpart): per-edge coefficients c_e aligned with the
 * cached edges (the whole graph must be cached).  A push then adds p * c_e along edge e
 * with p = scale * d (set divdeg = 0 when the coefficients are already normalised).  Only
 * for contexts with one measure (nm = 1) and no batch columns; EXC_EINVAL otherwise. */
int exc_set_edge_weights(void* ctx, const float* h_w, uint64_t cnt);

/* Dense passes by pull (no reference counterpart): builds the transpose of the cached
 * graph on the device (the whole graph must be in the edge cache) and a weight buffer.
 * Afterwards exc_push_resident over all vertices runs as fold + pull when the last
 * snapshot's demand is at least threshold * m: every pending vertex folds its residual
 * and publishes its weight, then every vertex adds the weights of its in-neighbours --
 * the same deltas as the push, all folds before all adds.  EXC_ENOMEM leaves the push. */
int exc_enable_pull(void* ctx, double threshold);
/* 1 when the last dense pass also folded the next one (its pending vertices have moved r into
 * x and published their weights): the next pass must then be exc_push_resident over the
 * whole vertex range, which runs as a dense pass whatever its demand; 0 otherwise.  No
 * reference counterpart (the reference folds inside each push). */
int exc_dense_required(void* ctx);
/* allow = 0 keeps later dense passes from folding the pass after them (the default, 1, lets
 * them while the demand is high).  A caller that will stop after the next pass -- a pass
 * limit, or a stop test it will apply to the next snapshot -- turns it off, so the state it
 * stops on holds every delta it has folded (x + r = b + W x).  No reference counterpart. */
int exc_fold_ahead(void* ctx, int allow);
/* Measure slots from `count` on frozen (1 <= count <= nm; nm, every slot, after exc_create
 * and after any call that replaces the state): their eps becomes infinite for every kernel,
 * so no vertex is pending in them -- they are neither pushed nor folded, their x and r stay
 * as they are, and snapshots still sum their |r| -- and dense passes gather rows of the first
 * `count` measures only (one measure: the source-sorted pull).  The caller decides when a
 * slot is done (bin/excent: nothing pending in it, or its FP64 residual sum within n eps
 * under the global rule).  Needs a snapshot; narrowing is refused (EXC_ESTATE) while the
 * last dense pass has folded the next one and a slot being frozen had a vertex above eps
 * (its folded deltas would be lost), and widening while any pass is folded.  No reference
 * counterpart (the reference pushes every measure of a pending vertex). */
int exc_set_active_measures(void* ctx, int count);
/* A burst: `passes` (1..EXC_BURST_MAX) passes over the whole cached graph enqueued back to back
 * with one host synchronisation at the end -- dense passes (dense = 1; exc_enable_pull) or
 * bitmap-driven pushes of every pending vertex (dense = 0; refused while the last dense pass
 * has folded the next one).  Each pass starts from the snapshot before it (the first from the
 * last exc_snapshot) and takes the snapshot of the state it leaves: its per-chunk counts go to
 * h_cnt[i * 2 nc ...] and, when the arrays are given, its FP64 residual sums and largest |r|
 * per measure to h_l1[i * nm + k], h_linf[i * nm + k]; the next exc_snapshot returns the last
 * pass's without a launch.  A pass that starts from a snapshot with nothing pending changes
 * nothing (no vertex folds or pushes), so the caller ends its run at the first snapshot of the
 * burst with nothing pending, and the state is the one that snapshot saw.  Plain contexts
 * with the whole graph in the edge cache, no hot set and no trace.  No reference counterpart
 * (the reference synchronises with the host at every pass, src/excent.cu:664-665, 750). */
int exc_run_passes(void* ctx, int passes, int dense, uint64_t* h_cnt, double* h_l1, float* h_linf);
/* exc_run_passes in two halves, so that the next burst runs while the host reads this one:
 * exc_burst_submit enqueues a burst (same arguments and checks) and returns at once; at most
 * two bursts are in flight, the second starting from the state the first leaves.
 * exc_burst_collect waits for the oldest one and writes its passes' counts, sums and largest
 * |r| as exc_run_passes does (*passes: how many).  A burst submitted after the run's last
 * needed pass changes nothing once a pass started from nothing pending, so the caller may
 * submit the next burst before it knows whether the current one converged.  While a burst is
 * in flight the entry points that read or change the state refuse (EXC_ESTATE); after the
 * last one is collected the next exc_snapshot returns its last pass's snapshot.  No
 * reference counterpart. */
int exc_burst_submit(void* ctx, int passes, int dense);
int exc_burst_collect(void* ctx, uint64_t* h_cnt, double* h_l1, float* h_linf, int* passes);

/* Push a resident chunk run [ch.v0, ch.v1) from the arena on the resident stream
 * (launch_chunk with dst = resident[c], src/excent.cu:634-652, 684-688).  use_list:
 * push only the vertices whose snapshot bit is set (the ids launch, 639-645),
 * otherwise every vertex with a live check (k_push_range, 127-137).  group: lanes per
 * vertex for EXC_KERNEL_GROUP.  hot_skip: skip hot-cached vertices (range only). */
int exc_push_resident(void* ctx, const exc_chunk* ch, int use_list, int group, int hot_skip);

/* Stream a chunk (src/excent.cu:696-707): H2D copy of the slot's host edges for
 * [ch.e0, ch.e1) into the slot (or into the arena when to_resident), optional pass-0
 * hot-list fill of hot entries [hot_lo, hot_hi) (k_fill_hot, 212-220, 704-705),
 * then the push as in exc_push_resident, all on the slot's stream. */
int exc_stream_chunk(void* ctx, int slot, const exc_chunk* ch, int to_resident,
                     int hot_lo, int hot_hi, int use_list, int group, int hot_skip);

/* Gather (src/excent.cu:709-740): the slot's host vids[k], off[k+1] and packed edges
 * (off[k] of them) go to the device and are pushed (k_push_list, 140-148). */
int exc_gather_chunk(void* ctx, int slot, int k, int group);

/* Hot push (launch_hot, src/excent.cu:615-628; k_compact_hot + k_push_hot_ids):
 * hot entries [lo, hi) pending now are compacted and pushed on the resident stream
 * with the count kept on the device (no host round trip).  after_slots != 0 makes
 * the resident stream wait for both slots' last work first (746). */
int exc_push_hot(void* ctx, int lo, int hi, int group, int after_slots);

/* Batched personalised PageRank (no reference counterpart).  K teleport vectors share
 * every list read: x and r are [n][K] rows, one wave per vertex, lane j = column j, so a
 * neighbour update is one wave-wide atomic on a contiguous row.  A context created with
 * batch_k = K starts from x = r = 0; exc_batch_seed adds value to r[seed][col] (the
 * teleport mass (1-d) s_j); the snapshot sets a vertex pending when any active column is
 * above its eps; exc_batch_colpend gives the per-column pending counts of the last
 * snapshot and exc_batch_set_active drops converged columns from the pushes. */
int exc_batch_eps(void* ctx, const float* eps);
int exc_batch_seed(void* ctx, int col, const int32_t* seeds, int count, float value);
int exc_batch_colpend(void* ctx, uint32_t* out);
int exc_batch_set_active(void* ctx, uint64_t mask);

/* End of pass (src/excent.cu:750-753): waits for all streams; *waited_s gets the wait. */
int exc_pass_sync(void* ctx, double* waited_s);

/* Pushed-edge total (the --count-pushes counter, src/excent.cu:117, 795). Always kept,
 * as one atomic per block. */
int exc_pushed(void* ctx, uint64_t* edges);
int exc_get_stats(void* ctx, exc_stats* out);
int exc_reset_stats(void* ctx);

/* In-degrees of the resident arena (all chunks resident): out[n] (uint32).  Used for the
 * eigenvector initial residual on directed graphs. */
int exc_indegree(void* ctx, uint32_t* h_out);

/* ---------------------------------------------------------------------------------
 * FP64 tools (tools/centrality_reference.cu:32-59, tools/excent_params.cu:27-60): the
 * synchronous pull over in-neighbours on a device copy of the transpose.
 * ------------------------------------------------------------------------------- */
/* Builds a device FP64 solver for the graph (CSR on host); computes the transpose on
 * the device. */
int exc_fp64_create(int device, int n, uint64_t m, const uint64_t* h_idx, const int32_t* h_adj, void** solver);
int exc_fp64_destroy(void* solver);
/* kind: EXC_PR (c=0.15, s=0.85, /outdeg), EXC_KATZ (c=beta, s=alpha), EXC_EV (power
 * iteration, unit L2, sign by sum).  x0: NULL = 0.15 (pr) / 1 (katz, ev), else the start.
 * Iterates until rel < tol or max_iter.  Returns iterations via *iters. */
int exc_fp64_solve(void* solver, int kind, double alpha, double beta, int max_iter, double tol,
                   const double* h_x0, double* h_out, double* rel, double* lambda, int* iters);
/* FP64 global residual ||c + s W x - x||_1 of a float32 state (W = the transpose's pull,
 * scaled by 1/outdeg when divdeg; the constant is c, or the vector h_c[n] when non-NULL),
 * the quantity the PageRank bound ||x - x*||_1 <= ||r||_1 / (1 - d) is stated in
 * (src/excent.cu:100-124 keeps the invariant x + r = c + s W x;
 * tools/centrality_reference.cu has no residual). */
int exc_fp64_residual(void* solver, double c, const double* h_c, double s, int divdeg, const float* h_x, double* l1);
/* Jacobi x = c + s W x with a per-vertex constant h_c[n] (personalised PageRank: c =
 * (1 - d) * teleport), from x = c until max|x' - x| / max|x'| < tol or max_iter
 * (tools/centrality_reference.cu:184-214 with a vector constant). */
int exc_fp64_solve_vec(void* solver, const double* h_c, double s, int divdeg, int max_iter, double tol, double* h_out,
                       double* rel, int* iters);
/* lambda1 by power iteration from h_start (NULL = ones), stop at rel change < tol. */
int exc_fp64_lambda(void* solver, const double* h_start, int max_iter, double tol,
                    double* lambda, double* mean_ev, int* iters);
/* Katz fixed-point mean with alpha (tools/excent_params.cu:149-163). */
int exc_fp64_katz_mean(void* solver, double alpha, double beta, double* mean, int* iters);
/* Per-edge FP64 coefficients (no reference counterpart; the FP64 check of the weighted
 * forward and adjoint solves): h_adj is the graph's adjacency as given to exc_fp64_create,
 * h_coef one coefficient per edge in the same order. */
int exc_fp64_set_coef(void* solver, const int32_t* h_adj, const double* h_coef);
/* Jacobi with the coefficients, from x = c until max|x' - x| / max|x'| < tol or max_iter:
 * transpose = 0: x = c + s W x, (W x)[u] = sum over edges v -> u of coef * x[v] (the
 * forward weighted personalised PageRank); transpose = 1: x = c + s W^T x, the sum over the
 * out-edges u -> w of coef * x[w] (its adjoint). */
int exc_fp64_solve_coef(void* solver, int transpose, const double* h_c, double s, int max_iter, double tol,
                        double* h_out, double* rel, int* iters);

#ifdef __cplusplus
}
#endif
#endif
