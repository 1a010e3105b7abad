/*
 * excent_oracle.c -- CPU restatement of the ExCent delta-push path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/excent.cu            init, pending, chunk_of, k_snapshot, push_vertex_g,
 *                            launch choice, pick_group, fetch plan, hot-set selection, tau
 *   include/StaticGraph.h    EGR header/offsets, plan_chunks, prefetch_lists
 *   tools/centrality_reference.cu   FP64 Jacobi / power iteration, compare, rank
 *   tools/excent_params.cu   lambda1, Katz mean, eigenvector mean
 *
 * How each function is pinned (checked against something other than itself):
 *   orc_egr_load / orc_egr_write : byte equality with the reference's sample.egr
 *       (written from sample.mtx in file order), narrow/wide round trips, and the
 *       reference's own StaticGraph reader compiled into oracle/_ref.
 *   orc_plan_chunks / orc_chunk_cap : StaticGraph::plan_chunks from oracle/_ref, plus
 *       brute-force properties (every chunk <= cap unless a single vertex).
 *   orc_snapshot : brute-force popcount/degree sums on tiny inputs, per-lane
 *       recomputation in numpy, and the closed-form "everything pending" case
 *       (count = n, demand = m).
 *   orc_fetch_plan / orc_pick_group / orc_use_list : exhaustive small tables written
 *       from the reference's decision rules by hand in the tests.
 *   orc_hot_select : brute-force top-cost prefix property (hot set = vertices of
 *       highest degree whose cost fits the budget, ties in vertex order).
 *   orc_prefetch_counts : StaticGraph::prefetch_lists page/run counts from oracle/_ref
 *       (madvise counted through the reference code) and a brute-force page-set union.
 *   orc_tau : hand-computed interpolation cases (end extrapolation, middle, clamp).
 *   orc_init : closed forms (0.15, beta, deg/lambda - 1).
 *   orc_run_push : converges to the FP64 oracle within the bound
 *       ||x - x*||_1 <= ||r||_1 / (1 - d) for PageRank; FP64 invariant x + r = b + W x;
 *       closed forms on a directed cycle, a star and an edgeless graph.
 *   orc_fp64_reference : dense numpy.linalg.solve / eig on the sample (which
 *       reproduce sample.env's LAMBDA/ALPHA/EPS digits), closed forms on cycles.
 *   orc_params : sample.env digits (LAMBDA 9.1443499824, ALPHA 0.0929536,
 *       EPS_KATZ 0.058337, EPS_EV 0.0084467).
 *   orc_compare / orc_rank : hand-computed vectors (identity, permutations).
 */
#include "excent_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned long long ull;

void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------- EGR */

/* include/StaticGraph.h:32-55: int32 n; int32 m or -1 then int64 m; offsets int32
 * (read as uint32) or int64; int32 columns. */
int orc_egr_load(const char* path, int* n_out, uint64_t* m_out, uint64_t** idx_out, int32_t** adj_out) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2) { fclose(f); return -3; }
    int n = hdr[0];
    uint64_t m;
    int wide = hdr[1] == -1;
    if (wide) { if (fread(&m, 8, 1, f) != 1) { fclose(f); return -3; } }
    else m = (uint64_t)(uint32_t)hdr[1];
    if (n < 1) { fclose(f); return -2; }
    uint64_t* idx = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    if (wide) {
        if (fread(idx, 8, (size_t)n + 1, f) != (size_t)n + 1) { free(idx); fclose(f); return -3; }
    } else {
        int32_t* t = (int32_t*)malloc(((size_t)n + 1) * 4);
        if (fread(t, 4, (size_t)n + 1, f) != (size_t)n + 1) { free(t); free(idx); fclose(f); return -3; }
        for (size_t i = 0; i <= (size_t)n; i++) idx[i] = (uint64_t)(uint32_t)t[i];
        free(t);
    }
    int32_t* adj = (int32_t*)malloc((m ? m : 1) * 4);
    if (m && fread(adj, 4, m, f) != m) { free(adj); free(idx); fclose(f); return -3; }
    fclose(f);
    *n_out = n; *m_out = m; *idx_out = idx; *adj_out = adj;
    return 0;
}

/* README.md:136-144 layout; tools/efficient_converter.py:70-84 writes the wide form. */
int orc_egr_write(const char* path, int n, uint64_t m, const uint64_t* idx, const int32_t* adj, int wide) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int32_t h0 = n;
    fwrite(&h0, 4, 1, f);
    if (wide) {
        int32_t s = -1; fwrite(&s, 4, 1, f); fwrite(&m, 8, 1, f);
        fwrite(idx, 8, (size_t)n + 1, f);
    } else {
        int32_t e = (int32_t)m; fwrite(&e, 4, 1, f);
        for (size_t i = 0; i <= (size_t)n; i++) { int32_t t = (int32_t)(uint32_t)idx[i]; fwrite(&t, 4, 1, f); }
    }
    if (m) fwrite(adj, 4, m, f);
    return fclose(f) == 0 ? 0 : -1;
}

/* ---------------------------------------------------------------- chunks */

/* src/excent.cu:321-326 */
uint64_t orc_chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges) {
    if (chunk_medges > 0) return (uint64_t)(chunk_medges * 1e6);
    uint64_t lo = (uint64_t)(min_chunk_medges * 1e6), hi = 64000000ull;
    uint64_t c = (uint64_t)ceil((double)m / target_chunks);
    uint64_t a = lo > c ? lo : c;
    return a < hi ? a : hi;
}

/* include/StaticGraph.h:128-139: v1 = upper_bound(off[v+1..n], off[v] + cap) - 1, at least v + 1. */
int orc_plan_chunks(int n, const uint64_t* idx, uint64_t cap, int* v0_out, int max_chunks) {
    int nc = 0, v = 0;
    while (v < n) {
        uint64_t limit = idx[v] + cap;
        /* upper_bound over idx[v+1 .. n] */
        long lo = v + 1, hi = (long)n + 1;
        while (lo < hi) { long mid = (lo + hi) / 2; if (idx[mid] <= limit) lo = mid + 1; else hi = mid; }
        int v1 = (int)lo - 1;
        if (v1 <= v) v1 = v + 1;
        if (nc < max_chunks) v0_out[nc] = v;
        nc++;
        v = v1;
    }
    return nc;
}

/* ---------------------------------------------------------------- init */

/* src/excent.cu:41-47 (lambda arrives as float, 372). */
void orc_init(int n, const uint64_t* idx, int kind, float beta, float lambda, float* x, float* r) {
    for (int v = 0; v < n; v++) {
        if (kind == ORC_PR) { x[v] = 0.f; r[v] = 1.f - 0.85f; }
        else if (kind == ORC_KATZ) { x[v] = 0.f; r[v] = beta; }
        else { x[v] = 1.f; r[v] = (float)((double)(idx[v + 1] - idx[v]) / (double)lambda - 1.0); }
    }
}

/* ---------------------------------------------------------------- snapshot */

/* src/excent.cu:49-56 */
static int chunk_of(int v, const int* cv0, int nc) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (cv0[mid] <= v) lo = mid; else hi = mid - 1; }
    return lo;
}

/* src/excent.cu:58-64 */
static int pending_v(int nm, const float* const* r, const float* eps, long v) {
    int p = 0;
    for (int k = 0; k < nm; k++) p |= fabsf(r[k][v]) > eps[k];
    return p;
}

/* src/excent.cu:69-98, serially: the counts are integers, so any order gives the same result. */
void orc_snapshot(int n, int nm, const float* const* r, const float* eps,
                  const uint64_t* idx, const int* cv0, int nc, const uint32_t* hot,
                  uint32_t* bitmap, uint64_t* cnt, uint64_t* hotpend, uint16_t* trace) {
    size_t words = ((size_t)n + 31) / 32;
    memset(bitmap, 0, words * 4);
    memset(cnt, 0, 2 * (size_t)nc * sizeof(uint64_t));
    uint64_t hp = 0;
    for (int v = 0; v < n; v++) {
        int pend = pending_v(nm, r, eps, v);
        if (trace && pend) trace[v]++;
        if (hot && pend && ((hot[v >> 5] >> (v & 31)) & 1u)) { hp++; pend = 0; }
        if (!pend) continue;
        bitmap[v >> 5] |= 1u << (v & 31);
        int c = chunk_of(v, cv0, nc);
        cnt[2 * c] += 1;
        cnt[2 * c + 1] += idx[v + 1] - idx[v];
    }
    *hotpend = hp;
}

/* ---------------------------------------------------------------- plan */

/* src/excent.cu:666-677 */
int orc_fetch_plan(int nc, const uint64_t* cnt, const uint8_t* resident, const uint8_t* loaded,
                   double tau, const uint64_t* chunk_edges, int hot_mode, int pass,
                   const int* hlo, const int* hhi, int8_t* mode, uint64_t* pend_v, uint64_t* demand) {
    int need = 0;
    uint64_t pv = 0, de = 0;
    for (int c = 0; c < nc; c++) {
        uint64_t cv = cnt[2 * c], ce = cnt[2 * c + 1];
        pv += cv; de += ce;
        mode[c] = 0;
        if (cv == 0) continue;
        if (resident[c] && loaded[c]) mode[c] = 1;
        else if (resident[c] || (double)ce >= tau * (double)chunk_edges[c]) mode[c] = 2;
        else if (hot_mode && pass == 0 && hhi[c] > hlo[c]) mode[c] = 2;
        else { mode[c] = 3; need = 1; }
    }
    *pend_v = pv; *demand = de;
    return need;
}

/* src/excent.cu:238 */
int orc_pick_group(double avg) { int g = 1; while (g < 32 && g < avg) g <<= 1; return g; }
/* src/excent.cu:639 */
int orc_use_list(uint64_t k, int nv, double list_frac) { return (double)k < list_frac * nv; }

/* ---------------------------------------------------------------- hot set */

static uint64_t hot_cost(uint64_t d) { return d * 4 + 2 * 4 + 8; }   /* src/excent.cu:433 */
static int cmp_desc_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? 1 : x > y ? -1 : 0;
}

/* src/excent.cu:432-491 */
int orc_hot_select(int n, const uint64_t* idx, uint64_t budget, const uint16_t* score,
                   int32_t* ids, uint64_t* off, uint32_t* bitmask, uint64_t* deg_min) {
    const uint64_t CAP = 1ull << 20;
    uint64_t* hist = (uint64_t*)calloc(CAP, sizeof(uint64_t));
    uint64_t* big = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    size_t nb = 0;
    for (int v = 0; v < n; v++) { uint64_t d = idx[v + 1] - idx[v]; if (d >= CAP) big[nb++] = d; else hist[d]++; }
    qsort(big, nb, sizeof(uint64_t), cmp_desc_u64);
    uint64_t used = 0, T = 0, quota = 0; int done = 0;
    for (size_t i = 0; i < nb; i++) {
        uint64_t d = big[i];
        if (used + hot_cost(d) > budget) { T = d; quota = 0; done = 1; break; }
        used += hot_cost(d); T = d; quota++;
    }
    if (!done) { quota = 0; T = CAP - 1; }
    for (uint64_t d = CAP - 1; !done && d >= 1; d--) {
        if (!hist[d]) continue;
        uint64_t take = (budget - used) / hot_cost(d);
        if (take > hist[d]) take = hist[d];
        used += take * hot_cost(d);
        T = d; quota = take;
        if (take < hist[d]) done = 1;
    }
    int S = 0; uint64_t s_left = 0;
    if (score) {   /* src/excent.cu:466-472 */
        double* cs = (double*)calloc(65536, sizeof(double));
        for (int v = 0; v < n; v++) { uint64_t d = idx[v + 1] - idx[v]; if (d && score[v]) cs[score[v]] += (double)hot_cost(d); }
        double u = 0;
        for (int s = 65535; s >= 1; s--) {
            if (u + cs[s] <= (double)budget) { u += cs[s]; continue; }
            S = s; s_left = budget - (uint64_t)u; break;
        }
        free(cs);
    }
    size_t words = ((size_t)n + 31) / 32;
    memset(bitmask, 0, words * 4);
    int nh = 0;
    off[0] = 0;
    for (int v = 0; v < n; v++) {   /* src/excent.cu:477-490 */
        uint64_t d = idx[v + 1] - idx[v];
        if (d == 0) continue;
        if (score) {
            if (score[v] == 0 || score[v] < S) continue;
            if (score[v] == S) { if (hot_cost(d) > s_left) continue; s_left -= hot_cost(d); }
        } else {
            int take = d > T || (d == T && quota > 0);
            if (!take) continue;
            if (d == T) quota--;
        }
        ids[nh] = v; off[nh + 1] = off[nh] + d; nh++;
        bitmask[v >> 5] |= 1u << (v & 31);
    }
    *deg_min = score ? 0 : T;
    free(hist); free(big);
    return nh;
}

/* ---------------------------------------------------------------- prefetch */

/* include/StaticGraph.h:90-114, counting the madvise calls instead of issuing them. */
void orc_prefetch_counts(const int32_t* vids, int k, const uint64_t* idx, uint64_t nlist_off,
                         uint64_t P, int io_threads, uint64_t* pages, uint64_t* runs) {
    uint64_t np = 0, nr = 0;
    int parts = k / 1024 < io_threads ? k / 1024 : io_threads;
    if (parts < 1) parts = 1;
    for (int t = 0; t < parts; t++) {
        int i0 = (int)((long long)k * t / parts), i1 = (int)((long long)k * (t + 1) / parts);
        uint64_t r0 = 0, r1 = 0; int open = 0;
        for (int i = i0; i < i1; i++) {
            uint64_t d = idx[vids[i] + 1] - idx[vids[i]];
            if (!d) continue;
            uint64_t b0 = nlist_off + idx[vids[i]] * 4, b1 = b0 + d * 4;
            uint64_t p0 = b0 / P, p1 = (b1 + P - 1) / P;
            if (open && p0 <= r1) { if (p1 > r1) r1 = p1; }
            else { if (open) { np += r1 - r0; nr++; } r0 = p0; r1 = p1; open = 1; }
        }
        if (open) { np += r1 - r0; nr++; }
    }
    *pages = np; *runs = nr;
}

/* ---------------------------------------------------------------- tau */

/* src/excent.cu:566-572 */
double orc_tau(double t_stream, const double* t_g, const double* f_got) {
    double s;
    if (t_g[0] >= t_stream) s = f_got[0] * t_stream / t_g[0];
    else if (t_g[2] < t_stream) s = f_got[2] * t_stream / t_g[2];
    else {
        int j = t_g[1] >= t_stream ? 1 : 2;
        s = f_got[j - 1] + (f_got[j] - f_got[j - 1]) * (t_stream - t_g[j - 1]) / (t_g[j] - t_g[j - 1]);
    }
    if (s < 0.005) s = 0.005;
    if (s > 1.0) s = 1.0;
    return s;
}

/* ---------------------------------------------------------------- CPU push */

static inline float atomic_exch_f(float* p, float v) {
    uint32_t nv; memcpy(&nv, &v, 4);
    uint32_t old = __atomic_exchange_n((uint32_t*)p, nv, __ATOMIC_RELAXED);
    float o; memcpy(&o, &old, 4);
    return o;
}
static inline void atomic_add_f(float* p, float v) {
    uint32_t* q = (uint32_t*)p;
    uint32_t old = __atomic_load_n(q, __ATOMIC_RELAXED);
    for (;;) {
        float f; memcpy(&f, &old, 4);
        f += v;
        uint32_t nv; memcpy(&nv, &f, 4);
        if (__atomic_compare_exchange_n(q, &old, nv, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
    }
}

/* src/excent.cu:100-124: live check, exchange, fold, scatter of scale*d[/deg]. */
static uint64_t push_vertex(int v, const uint64_t* idx, const int32_t* adj, int nm, float* const* x,
                            float* const* r, const float* eps, const float* scale, const int* divdeg) {
    float p[3] = {0.f, 0.f, 0.f};
    int any = 0;
    uint64_t deg = idx[v + 1] - idx[v];
    for (int k = 0; k < nm; k++) {
        float d = 0.f;
        uint32_t bits = __atomic_load_n((uint32_t*)&r[k][v], __ATOMIC_RELAXED);
        float cur; memcpy(&cur, &bits, 4);
        if (fabsf(cur) > eps[k]) {
            d = atomic_exch_f(&r[k][v], 0.f);
            x[k][v] += d;
        }
        if (d != 0.f) { any = 1; p[k] = divdeg[k] ? scale[k] * d / (float)deg : scale[k] * d; }
    }
    if (!any || deg == 0) return 0;
    const int32_t* nb = adj + idx[v];
    for (uint64_t i = 0; i < deg; i++) {
        int u = nb[i];
        for (int k = 0; k < nm; k++) if (p[k] != 0.f) atomic_add_f(&r[k][u], p[k]);
    }
    return deg;
}

static double now_s(void) { struct timespec t; clock_gettime(CLOCK_MONOTONIC, &t); return t.tv_sec + 1e-9 * t.tv_nsec; }

/* Pass loop of src/excent.cu:656-760 with every chunk resident in host memory:
 * snapshot, plan (all resident), ids-vs-range choice per chunk (639), push. */
int orc_run_push(int n, const uint64_t* idx, const int32_t* adj, int nc, const int* cv0,
                 int nm, float* const* x, float* const* r, const float* eps, const float* scale,
                 const int* divdeg, double list_frac, int max_passes, int threads,
                 double time_limit, uint64_t* pushed_out, double* seconds) {
    if (threads > 0) omp_set_num_threads(threads);
    size_t words = ((size_t)n + 31) / 32;
    uint32_t* bitmap = (uint32_t*)malloc(words * 4);
    uint64_t* cnt = (uint64_t*)malloc(2 * (size_t)nc * 8);
    int32_t* list = (int32_t*)malloc((size_t)n * 4);
    uint64_t pushed = 0, hp = 0;
    const float* const* rc = (const float* const*)r;
    double t0 = now_s();
    int pass = 0;
    for (; pass < max_passes; pass++) {
        orc_snapshot(n, nm, rc, eps, idx, cv0, nc, NULL, bitmap, cnt, &hp, NULL);
        uint64_t pend = 0;
        for (int c = 0; c < nc; c++) pend += cnt[2 * c];
        if (pend == 0) break;
        for (int c = 0; c < nc; c++) {
            uint64_t k = cnt[2 * c];
            if (!k) continue;
            int v0 = cv0[c], v1 = c + 1 < nc ? cv0[c + 1] : n;
            uint64_t pc = 0;
            if (orc_use_list(k, v1 - v0, list_frac)) {
                int kk = 0;
                for (int v = v0; v < v1; v++) if ((bitmap[v >> 5] >> (v & 31)) & 1u) list[kk++] = v;
                #pragma omp parallel for schedule(dynamic, 64) reduction(+ : pc)
                for (int i = 0; i < kk; i++) pc += push_vertex(list[i], idx, adj, nm, x, r, eps, scale, divdeg);
            } else {
                #pragma omp parallel for schedule(dynamic, 256) reduction(+ : pc)
                for (int v = v0; v < v1; v++) pc += push_vertex(v, idx, adj, nm, x, r, eps, scale, divdeg);
            }
            pushed += pc;
        }
        if (time_limit > 0 && now_s() - t0 >= time_limit) { pass++; break; }
    }
    if (seconds) *seconds = now_s() - t0;
    if (pushed_out) *pushed_out = pushed;
    free(bitmap); free(cnt); free(list);
    return pass;
}

/* ---------------------------------------------------------------- FP64 oracle */

/* In-neighbour lists (CSC) of a CSR, counting sort, lists in ascending source order. */
static void transpose(int n, const uint64_t* idx, const int32_t* adj, uint64_t** tidx_out, int32_t** tadj_out) {
    uint64_t m = idx[n];
    uint64_t* tidx = (uint64_t*)calloc((size_t)n + 1, 8);
    int32_t* tadj = (int32_t*)malloc((m ? m : 1) * 4);
    for (uint64_t e = 0; e < m; e++) tidx[adj[e] + 1]++;
    for (int v = 0; v < n; v++) tidx[v + 1] += tidx[v];
    uint64_t* pos = (uint64_t*)malloc(((size_t)n + 1) * 8);
    memcpy(pos, tidx, ((size_t)n + 1) * 8);
    for (int v = 0; v < n; v++)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) tadj[pos[adj[e]]++] = v;
    free(pos);
    *tidx_out = tidx; *tadj_out = tadj;
}

/* tools/centrality_reference.cu:71-76 */
static void unit_vec(int n, double* v) {
    double s = 0, t = 0;
    for (int i = 0; i < n; i++) { s += v[i] * v[i]; t += v[i]; }
    double f = (t < 0 ? -1.0 : 1.0) / sqrt(s);
    for (int i = 0; i < n; i++) v[i] *= f;
}

/* y[w] = c + s * sum_{u -> w} x[u] * (divdeg ? 1/L(u) : 1)   (centrality_reference.cu:32-49 on the CSC) */
static void pull_step(int n, const uint64_t* idx, const uint64_t* tidx, const int32_t* tadj,
                      const double* x, double* y, double c, double s, int divdeg, double* mxd, double* mxy) {
    double a = 0, b = 0;
    #pragma omp parallel for schedule(dynamic, 1024) reduction(max : a, b)
    for (int w = 0; w < n; w++) {
        double acc = 0;
        for (uint64_t e = tidx[w]; e < tidx[w + 1]; e++) {
            int u = tadj[e];
            acc += divdeg ? x[u] / (double)(idx[u + 1] - idx[u]) : x[u];
        }
        double nv = c + s * acc;
        y[w] = nv;
        double d = fabs(nv - x[w]);
        if (d > a) a = d;
        if (fabs(nv) > b) b = fabs(nv);
    }
    *mxd = a; *mxy = b;
}

/* tools/centrality_reference.cu:184-214 */
int orc_fp64_reference(int n, const uint64_t* idx, const int32_t* adj, int kind,
                       double alpha, double beta, int max_iter, double tol,
                       double* out, double* rel_out, double* lambda_out) {
    uint64_t* tidx; int32_t* tadj;
    transpose(n, idx, adj, &tidx, &tadj);
    double* x = (double*)malloc((size_t)n * 8);
    double* y = (double*)malloc((size_t)n * 8);
    for (int i = 0; i < n; i++) x[i] = kind == ORC_PR ? 0.15 : 1.0;
    double c = 0, s = 1; int divdeg = 0;
    if (kind == ORC_PR) { c = 0.15; s = 0.85; divdeg = 1; }
    else if (kind == ORC_KATZ) { c = beta; s = alpha; }
    int it = 0; double rel = 1, lam = 0;
    for (; it < max_iter && rel > tol; it++) {
        double mxd, mxy;
        pull_step(n, idx, tidx, tadj, x, y, c, s, divdeg, &mxd, &mxy);
        if (kind == ORC_EV) {
            double ss = 0;
            for (int i = 0; i < n; i++) ss += y[i] * y[i];
            lam = sqrt(ss);
            for (int i = 0; i < n; i++) y[i] /= lam;
            double md = 0;
            for (int i = 0; i < n; i++) { double d = fabs(x[i] - y[i]); if (d > md) md = d; }
            rel = md;
            if (it == 0) { unit_vec(n, y); rel = 1; }
        } else {
            rel = mxd / mxy;
        }
        double* t = x; x = y; y = t;
    }
    memcpy(out, x, (size_t)n * 8);
    if (rel_out) *rel_out = rel;
    if (lambda_out) *lambda_out = lam;
    free(x); free(y); free(tidx); free(tadj);
    return it;
}

int orc_fp64_solve_b(int n, const uint64_t* idx, const int32_t* adj, const double* b, double s, int divdeg,
                     int max_iter, double tol, double* out, double* rel_out) {
    uint64_t* tidx; int32_t* tadj;
    transpose(n, idx, adj, &tidx, &tadj);
    double* x = (double*)malloc((size_t)n * 8);
    double* y = (double*)malloc((size_t)n * 8);
    memcpy(x, b, (size_t)n * 8);
    int it = 0;
    double rel = 1;
    for (; it < max_iter && rel > tol; it++) {
        double a = 0, mx = 0;
        #pragma omp parallel for schedule(dynamic, 1024) reduction(max : a, mx)
        for (int w = 0; w < n; w++) {
            double acc = 0;
            for (uint64_t e = tidx[w]; e < tidx[w + 1]; e++) {
                int u = tadj[e];
                acc += divdeg ? x[u] / (double)(idx[u + 1] - idx[u]) : x[u];
            }
            double nv = b[w] + s * acc;
            y[w] = nv;
            if (fabs(nv - x[w]) > a) a = fabs(nv - x[w]);
            if (fabs(nv) > mx) mx = fabs(nv);
        }
        rel = mx > 0 ? a / mx : 0;
        double* t = x; x = y; y = t;
    }
    memcpy(out, x, (size_t)n * 8);
    if (rel_out) *rel_out = rel;
    free(x); free(y); free(tidx); free(tadj);
    return it;
}

int orc_fp64_solve_w(int n, const uint64_t* idx, const int32_t* adj, const double* w, const double* b, double s,
                     int max_iter, double tol, double* out, double* rel_out) {
    uint64_t m = idx[n];
    uint64_t* tidx = (uint64_t*)calloc((size_t)n + 1, 8);
    int32_t* tsrc = (int32_t*)malloc((m ? m : 1) * 4);
    double* tw = (double*)malloc((m ? m : 1) * 8);
    for (uint64_t e = 0; e < m; e++) tidx[adj[e] + 1]++;
    for (int v = 0; v < n; v++) tidx[v + 1] += tidx[v];
    uint64_t* pos = (uint64_t*)malloc(((size_t)n + 1) * 8);
    memcpy(pos, tidx, ((size_t)n + 1) * 8);
    for (int v = 0; v < n; v++)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) { uint64_t q = pos[adj[e]]++; tsrc[q] = v; tw[q] = w[e]; }
    free(pos);
    double* x = (double*)malloc((size_t)n * 8);
    double* y = (double*)malloc((size_t)n * 8);
    memcpy(x, b, (size_t)n * 8);
    int it = 0;
    double rel = 1;
    for (; it < max_iter && rel > tol; it++) {
        double a = 0, mx = 0;
        #pragma omp parallel for schedule(dynamic, 1024) reduction(max : a, mx)
        for (int u = 0; u < n; u++) {
            double acc = 0;
            for (uint64_t e = tidx[u]; e < tidx[u + 1]; e++) acc += tw[e] * x[tsrc[e]];
            double nv = b[u] + s * acc;
            y[u] = nv;
            if (fabs(nv - x[u]) > a) a = fabs(nv - x[u]);
            if (fabs(nv) > mx) mx = fabs(nv);
        }
        rel = mx > 0 ? a / mx : 0;
        double* t = x; x = y; y = t;
    }
    memcpy(out, x, (size_t)n * 8);
    if (rel_out) *rel_out = rel;
    free(x); free(y); free(tidx); free(tsrc); free(tw);
    return it;
}

/* tools/excent_params.cu:105-163, with the power iteration on the in-neighbour pull. */
int orc_params(int n, const uint64_t* idx, const int32_t* adj, const double* start, double afac,
               double* lambda, double* alpha, double* mean_katz, double* mean_ev, long* symmetric_bad) {
    uint64_t m = idx[n];
    long bad = 0;
    for (int v = 0; v < n; v++)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) {
            int u = adj[e], found = 0;
            for (uint64_t f = idx[u]; f < idx[u + 1]; f++) if (adj[f] == v) { found = 1; break; }
            if (!found) bad++;
        }
    (void)m;
    *symmetric_bad = bad;
    uint64_t* tidx; int32_t* tadj;
    transpose(n, idx, adj, &tidx, &tadj);
    double* x = (double*)malloc((size_t)n * 8);
    double* y = (double*)malloc((size_t)n * 8);
    for (int i = 0; i < n; i++) x[i] = start ? start[i] : 1.0;
    double nn = 0; for (int i = 0; i < n; i++) nn += x[i] * x[i];
    nn = sqrt(nn);
    for (int i = 0; i < n; i++) x[i] /= nn;
    double lam = 0, lamPrev = 0, a, b;
    int pit = 0;
    for (; pit < 5000; pit++) {
        pull_step(n, idx, tidx, tadj, x, y, 0.0, 1.0, 0, &a, &b);
        double ss = 0; for (int i = 0; i < n; i++) ss += y[i] * y[i];
        lam = sqrt(ss);
        for (int i = 0; i < n; i++) y[i] /= lam;
        double* t = x; x = y; y = t;
        if (pit > 10 && fabs(lam - lamPrev) / lam < 1e-12) break;
        lamPrev = lam;
    }
    double v1sum = 0; for (int i = 0; i < n; i++) v1sum += x[i];
    *mean_ev = v1sum * v1sum / n;
    *lambda = lam;
    double al = afac / lam;
    *alpha = al;
    for (int i = 0; i < n; i++) x[i] = 1.0;
    double meanK = 0;
    for (int it = 0; it < 5000; it++) {
        pull_step(n, idx, tidx, tadj, x, y, 1.0, al, 0, &a, &b);
        double* t = x; x = y; y = t;
        double sum = 0; for (int i = 0; i < n; i++) sum += x[i];
        double nm = sum / n;
        if (fabs(nm - meanK) / nm < 1e-12) { meanK = nm; break; }
        meanK = nm;
    }
    *mean_katz = meanK;
    free(x); free(y); free(tidx); free(tadj);
    return pit;
}

/* ---------------------------------------------------------------- compare */

static const double* g_key;
static int cmp_key_desc(const void* a, const void* b) {
    int i = *(const int*)a, j = *(const int*)b;
    if (g_key[i] > g_key[j]) return -1;
    if (g_key[i] < g_key[j]) return 1;
    return i < j ? -1 : i > j ? 1 : 0;
}
static int cmp_int(const void* a, const void* b) { int x = *(const int*)a, y = *(const int*)b; return (x > y) - (x < y); }

static int* top_order(int n, const double* key) {
    int* o = (int*)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) o[i] = i;
    g_key = key;
    qsort(o, n, sizeof(int), cmp_key_desc);
    return o;
}
static int overlap(const int* a, const int* b, int k) {
    int* x = (int*)malloc(k * sizeof(int)); int* y = (int*)malloc(k * sizeof(int));
    memcpy(x, a, k * sizeof(int)); memcpy(y, b, k * sizeof(int));
    qsort(x, k, sizeof(int), cmp_int); qsort(y, k, sizeof(int), cmp_int);
    int i = 0, j = 0, c = 0;
    while (i < k && j < k) { if (x[i] == y[j]) { c++; i++; j++; } else if (x[i] < y[j]) i++; else j++; }
    free(x); free(y);
    return c;
}

/* tools/centrality_reference.cu:78-105 (ties broken by vertex id) */
void orc_compare(int n, const double* ref, const float* xf, int unit_vectors,
                 double* l1_rel, double* l2, double* max_rel, int* top100) {
    double* r = (double*)malloc((size_t)n * 8);
    double* x = (double*)malloc((size_t)n * 8);
    memcpy(r, ref, (size_t)n * 8);
    for (int i = 0; i < n; i++) x[i] = xf[i];
    if (unit_vectors) { unit_vec(n, r); unit_vec(n, x); }
    double l1 = 0, ref1 = 0, s2 = 0, mr = 0;
    for (int i = 0; i < n; i++) {
        double d = fabs(x[i] - r[i]);
        l1 += d; ref1 += fabs(r[i]); s2 += d * d;
        if (fabs(r[i]) > 1e-12 && d / fabs(r[i]) > mr) mr = d / fabs(r[i]);
    }
    *l1_rel = l1 / ref1; *l2 = sqrt(s2); *max_rel = mr;
    int K = n < 100 ? n : 100;
    int* a = top_order(n, r); int* b = top_order(n, x);
    *top100 = overlap(a, b, K);
    free(a); free(b); free(r); free(x);
}

/* tools/centrality_reference.cu:109-138 */
void orc_rank(int n, const double* ref, const float* xf, int unit_vectors, double* p_at, double* tau1000) {
    double* r = (double*)malloc((size_t)n * 8);
    double* x = (double*)malloc((size_t)n * 8);
    memcpy(r, ref, (size_t)n * 8);
    for (int i = 0; i < n; i++) x[i] = xf[i];
    if (unit_vectors) { unit_vec(n, r); unit_vec(n, x); }
    int* a = top_order(n, r); int* b = top_order(n, x);
    const int ks[5] = {10, 100, 1000, 10000, 100000};
    int kmax = n < 100000 ? n : 100000;
    for (int j = 0; j < 5; j++) p_at[j] = ks[j] > kmax ? -1.0 : (double)overlap(a, b, ks[j]) / ks[j];
    int T = n < 1000 ? n : 1000;
    long long conc = 0, disc = 0;
    for (int i = 0; i < T; i++)
        for (int j = i + 1; j < T; j++) {
            double dr = r[a[i]] - r[a[j]], dx = x[a[i]] - x[a[j]];
            if (dr * dx > 0) conc++; else if (dr * dx < 0) disc++;
        }
    *tau1000 = T > 1 ? (double)(conc - disc) / ((double)T * (T - 1) / 2) : 1.0;
    free(a); free(b); free(r); free(x);
}

double orc_residual_l1(int n, const uint64_t* idx, const int32_t* adj, int kind,
                       double scale, double beta, const float* x) {
    double* acc = (double*)calloc((size_t)n, 8);
    for (int v = 0; v < n; v++) {
        uint64_t deg = idx[v + 1] - idx[v];
        if (!deg) continue;
        double w = kind == ORC_PR ? scale / (double)deg : scale;
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) acc[adj[e]] += w * (double)x[v];
    }
    double b = kind == ORC_PR ? 0.15 : kind == ORC_KATZ ? beta : 0.0, s = 0;
    for (int v = 0; v < n; v++) s += fabs(b + acc[v] - (double)x[v]);
    free(acc);
    return s;
}
