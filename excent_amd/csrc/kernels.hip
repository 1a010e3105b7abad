// kernels.hip -- ExCent device layer for MI355X (gfx950, CDNA4, wave64).
//
// Written for CDNA4 directly: 64-lane wavefronts (ballots are 64-bit), 256-thread
// workgroups (4 waves), LDS-staged block-level load balancing, hardware float
// atomics (global_atomic_add_f32, no return) into HBM-resident residuals.
// The reference's kernels live in src/excent.cu:41-220 (paths relative to the
// reference root); each kernel here cites what it stands for.
//
// State layout: x and r of all measures of a run are interleaved vertex-major,
// row v = [x or r of measure 0, 1, 2, pad] with a row stride S of 1, 2 or 4 floats.
// The fused push of 3 measures then touches one 16-byte row per neighbour (one
// 64-byte line) instead of three lines in three arrays.

#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "excent_hip.h"
#include "pull_sort.hpp"

namespace {

constexpr int BLOCK = 256;            // 4 waves
constexpr int WAVE = 64;
constexpr int NWAVES = BLOCK / WAVE;
constexpr uint32_t HUB = 1024;        // push vertices with more edges are split into pieces
constexpr uint32_t PIECE = 4096;      // edges per piece (one wave each)
constexpr int BITMAP_WINDOW = BLOCK * 32;   // vertices per block in bitmap mode (one word per lane)

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(EXC_EHIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                          \
    } while (0)

// ------------------------------------------------------------------ device helpers

struct DevMeasures {
    float* x;        // [n][S]
    float* r;        // [n][S]
    float eps[EXC_MAXM];
    float scale[EXC_MAXM];
    int divdeg[EXC_MAXM];
};

struct Piece {           // 32 bytes: a slice of one hub vertex's list
    uint64_t start;      // index into the launch's edge array
    uint32_t len;
    uint32_t pad;
    float p[4];
};

// Counters kept per context (one atomic per block per field).
// Per push source (4) x 8 fields, then the overflow flag.
enum { C_VERT = 0, C_VERT_NM, C_EDGES, C_EDGES_NM, C_PEDGES, C_PEDGES_NM, C_FIELDS = 8 };
// Row 4 holds the fold kernel of dense (pull) passes.
// rows 0-3: push sources; row 4: dense passes; row 5: dense passes that used the live-row
// bitmap (their fold counts give the rows the pull read)
enum { C_DENSE = 4, C_DENSE_LIVE = 5, C_OVERFLOW = 6 * C_FIELDS };
// Counters updated once per block by kernels with one block per tile (k_pull): spread over
// SPREAD_SLOTS slots on separate 128-byte lines (slot = block mod SPREAD_SLOTS), summed on
// the host.  Fields: rows k_pull read; the fused fold's vertices, vertex-measures, edges and
// edge-measures; rows read by dense passes over the first one and the first two measures
// only (S_ROWS counts the passes over every measure).
enum { SPREAD_SLOTS = 64, SPREAD_STRIDE = 16, S_ROWS = 0, S_VERT, S_VERT_NM, S_EDGES, S_EDGES_NM, S_ROWS_A1, S_ROWS_A2 };
enum { C_SPREAD = C_OVERFLOW + 16, C_NCOUNTERS = C_SPREAD + SPREAD_SLOTS * SPREAD_STRIDE };
// the dense-pass rows have no piece edges; a spare field counts k_pull_hub's row reads, two
// more those of hub passes over the first one / two measures
enum { C_PULL_HUB_LIVE = C_PEDGES, C_PULL_HUB_LIVE_A1 = 6, C_PULL_HUB_LIVE_A2 = 7 };

__device__ __forceinline__ void atomic_add_f32(float* p, float v) {
    // No-return hardware float atomic on coarse-grained HBM (global_atomic_add_f32).
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float atomic_exch_f32(float* p, float v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

// Piece counters: one 256-byte line per counter (count, p-row count), two phases, so a
// counter being cleared never shares a line with one being incremented.
constexpr int CNT_STRIDE = 64;   // uint32 slots between counters
__device__ __forceinline__ void clear_counters(uint32_t* c) {
    __hip_atomic_exchange(c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_exchange(c + CNT_STRIDE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive block scan of one uint32 per thread; *total gets the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) s_wave[wid] = x;
    __syncthreads();
    uint32_t wofs = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; w++) {
        uint32_t t = s_wave[w];
        wofs += w < wid ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return wofs + x - v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        uint32_t l2 = __shfl_xor(lo, o, WAVE), h2 = __shfl_xor(hi, o, WAVE);
        uint32_t s = lo + l2;
        hi = hi + h2 + (s < lo ? 1u : 0u);
        lo = s;
    }
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// FP64 global residual sums sum_v |r_k[v]| of a snapshot: each wave (k_snapshot) or block
// (k_pull's fused snapshot) adds its partial sums into one of L1_SLOTS slots on separate
// 128-byte lines (slot = wave or block index mod L1_SLOTS), so no address takes more than
// a few thousand FP64 atomics per pass; the host adds the slots.  The slots sit behind the
// per-chunk counts in the counts buffer and are cleared with them.
constexpr int L1_SLOTS = 64, L1_STRIDE = 16;   // doubles per slot line
constexpr int LINF_AT = 8;                       // per slot: the max |r| per measure, as float bits, from double 8 on
__device__ __forceinline__ void l1_flush(double* l1, long long who, int k, double v) {
    if (v != 0.0) __hip_atomic_fetch_add(l1 + (who % L1_SLOTS) * L1_STRIDE + k, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The largest |r| per measure: |r| as float bits (non-negative floats order as unsigned
// integers), maximised per wave or block, then one atomic max per slot.
__device__ __forceinline__ void linf_flush(double* l1, long long who, int k, uint32_t bits) {
    if (bits) atomicMax(reinterpret_cast<unsigned int*>(l1 + (who % L1_SLOTS) * L1_STRIDE + LINF_AT) + k, bits);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, WAVE));
    return v;
}

// src/excent.cu:49-56: last chunk c with cv0[c] <= v.
__device__ __forceinline__ int chunk_of(int v, const int* __restrict__ cv0, int nc) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cv0[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int S>
struct Row;
template <> struct Row<1> { float v[1]; };
template <> struct Row<2> { float v[2]; };
template <> struct Row<4> { float v[4]; };

template <int S>
__device__ __forceinline__ Row<S> load_row(const float* __restrict__ r, long long v) {
    Row<S> o;
    if constexpr (S == 1) { o.v[0] = r[v]; }
    else if constexpr (S == 2) { float2 t = reinterpret_cast<const float2*>(r)[v]; o.v[0] = t.x; o.v[1] = t.y; }
    else { float4 t = reinterpret_cast<const float4*>(r)[v]; o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w; }
    return o;
}

// Rows of the dense pass's weights p hold exactly NM floats (12 bytes for the fused
// three-measure pass, not the state's padded 16): 10.7 rows share a 128-byte line instead
// of 8, so more of the most-read rows stay cached.  The unused lanes of the Row read 0.
template <int NM, int S>
__device__ __forceinline__ Row<S> load_prow(const float* __restrict__ p, long long q) {
    Row<S> o;
    if constexpr (NM == S) {
        o = load_row<S>(p, q);
    } else {
        const float* a = p + q * NM;
#pragma unroll
        for (int k = 0; k < NM; k++) o.v[k] = a[k];
#pragma unroll
        for (int k = NM; k < S; k++) o.v[k] = 0.f;
    }
    return o;
}

// ------------------------------------------------------------------ init / state

// k_init (src/excent.cu:41-47) for measure slot k of a [n][S] state.
__global__ void k_init(int n, const uint64_t* __restrict__ idx, float* x, float* r, int S, int k,
                       int kind, float beta, float lambda) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        float xv, rv;
        if (kind == EXC_PR) { xv = 0.f; rv = 1.f - 0.85f; }
        else if (kind == EXC_KATZ) { xv = 0.f; rv = beta; }
        else { xv = 1.f; rv = (float)((double)(idx[v + 1] - idx[v]) / (double)lambda - 1.0); }
        x[v * S + k] = xv;
        r[v * S + k] = rv;
    }
}

__global__ void k_scatter_col(int n, const float* __restrict__ src, float* dst, int S, int k) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x)
        dst[v * S + k] = src[v];
}
__global__ void k_gather_col(long long cnt, const float* __restrict__ src, float* dst, int S, int k) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < cnt;
         v += (long long)gridDim.x * blockDim.x)
        dst[v] = src[v * S + k];
}
__global__ void k_zero_u32(uint32_t* p, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0u;
}
__global__ void k_zero_u64(unsigned long long* p, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0ull;
}

// ------------------------------------------------------------------ snapshot

struct SnapArgs {
    int n;
    const float* r;
    float eps[EXC_MAXM];
    const uint64_t* idx;
    const int* cv0;
    int nc;
    const uint32_t* hot;        // hot bitmask or null
    uint32_t* bitmap;           // ceil(n/32) words
    unsigned long long* cnt;    // [2*nc]
    unsigned long long* hotpend;
    uint16_t* trace;            // or null
    int segs_per_wave;          // 64-vertex segments per wave
    double* l1;                 // FP64 residual sum slots (L1_SLOTS x L1_STRIDE), or null
};

// k_snapshot (src/excent.cu:69-98).  One wave takes segs_per_wave consecutive 64-vertex
// segments.  The 64-bit ballot of a segment is the bitmap's two 32-bit words (same bit
// layout as the reference's warp ballots).  Pending counts and degree sums accumulate
// in registers while the segment lies inside the wave's current chunk and are flushed
// with one pair of atomics per wave and chunk; a segment that straddles a chunk start
// falls back to per-lane atomics for the lanes outside the current chunk.
template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_snapshot(SnapArgs a) {
    const int lane = lane_id();
    const long long gwave = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE;
    const long long nseg = ((long long)a.n + WAVE - 1) / WAVE;
    long long s0 = gwave * a.segs_per_wave;
    if (s0 >= nseg) return;
    long long s1 = s0 + a.segs_per_wave;
    if (s1 > nseg) s1 = nseg;
    const long long words = ((long long)a.n + 31) / 32;

    int cc = chunk_of((int)(s0 * WAVE), a.cv0, a.nc);
    int clo = a.cv0[cc], chi = cc + 1 < a.nc ? a.cv0[cc + 1] : a.n;
    uint32_t acc_n = 0;
    uint64_t acc_d = 0, hp = 0;
    double l1[NM];
    uint32_t mx[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) { l1[k] = 0.0; mx[k] = 0u; }

    auto flush = [&]() {
        uint32_t tn = wave_sum_u32(acc_n);
        uint64_t td = wave_sum_u64(acc_d);
        if (lane == 0 && tn) {
            atomicAdd(&a.cnt[2 * cc], (unsigned long long)tn);
            atomicAdd(&a.cnt[2 * cc + 1], (unsigned long long)td);
        }
        acc_n = 0;
        acc_d = 0;
    };

    constexpr int U = 4;
    for (long long sb = s0; sb < s1; sb += U) {
        Row<S> rows[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            long long v = (sb + u) * WAVE + lane;
            if (sb + u < s1 && v < a.n) rows[u] = load_row<S>(a.r, v);
            else {
#pragma unroll
                for (int k = 0; k < S; k++) rows[u].v[k] = 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {   // rows past n or past the wave's range are zero
#pragma unroll
            for (int k = 0; k < NM; k++) {
                l1[k] += fabs((double)rows[u].v[k]);
                mx[k] = max(mx[k], __float_as_uint(fabsf(rows[u].v[k])));
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long seg = sb + u;
            if (seg >= s1) break;
            const long long v = seg * WAVE + lane;
            const bool in = v < a.n;
            bool pend = false;
#pragma unroll
            for (int k = 0; k < NM; k++) pend |= in && fabsf(rows[u].v[k]) > a.eps[k];
            if (a.trace && pend) a.trace[v]++;
            if (a.hot) {
                bool ph = pend && ((a.hot[v >> 5] >> (v & 31)) & 1u);
                hp += ph ? 1 : 0;
                pend = pend && !ph;
            }
            const uint64_t mask = __ballot(pend);
            if (lane == 0) {
                const long long w0 = seg * 2;
                a.bitmap[w0] = (uint32_t)mask;
                if (w0 + 1 < words) a.bitmap[w0 + 1] = (uint32_t)(mask >> 32);
            }
            if (!mask) continue;
            const uint64_t d = pend ? a.idx[v + 1] - a.idx[v] : 0;
            long long vlast = seg * WAVE + WAVE - 1;
            if (vlast >= a.n) vlast = a.n - 1;
            const long long vfirst = seg * WAVE;
            if (vfirst >= clo && vlast < chi) {
                acc_n += pend ? 1u : 0u;
                acc_d += d;
            } else {
                flush();
                const int c = in ? chunk_of((int)v, a.cv0, a.nc) : -1;
                cc = chunk_of((int)vlast, a.cv0, a.nc);
                clo = a.cv0[cc];
                chi = cc + 1 < a.nc ? a.cv0[cc + 1] : a.n;
                if (pend) {
                    if (c == cc) { acc_n += 1u; acc_d += d; }
                    else {
                        atomicAdd(&a.cnt[2 * c], 1ull);
                        atomicAdd(&a.cnt[2 * c + 1], (unsigned long long)d);
                    }
                }
            }
        }
    }
    flush();
    if (a.hot) {
        uint64_t t = wave_sum_u64(hp);
        if (lane == 0 && t) atomicAdd(a.hotpend, (unsigned long long)t);
    }
    if (a.l1) {
#pragma unroll
        for (int k = 0; k < NM; k++) {
            const double t = wave_sum_f64(l1[k]);
            const uint32_t m = wave_max_u32(mx[k]);
            if (lane == 0) {
                l1_flush(a.l1, gwave, k, t);
                linf_flush(a.l1, gwave, k, m);
            }
        }
    }
}

// ------------------------------------------------------------------ push

enum { SRC_RANGE = 0, SRC_BITMAP = 1, SRC_LIST = 2, SRC_HOT = 3 };

struct PushArgs {
    int v0, v1;                     // RANGE / BITMAP: vertex range
    const uint32_t* bitmap;         // BITMAP
    const uint32_t* hot;            // RANGE: skip hot-cached vertices (null: none)
    const int32_t* vids;            // LIST: gathered vertex ids
    const uint64_t* off;            // LIST: k+1 offsets into edges
    int k;                          // LIST: entries
    const int32_t* hlist;           // HOT: compacted hot entry indices
    const uint32_t* hcount;         // HOT: entry count (device)
    long long ntiles;               // RANGE / BITMAP / LIST: tiles of the launch (HOT: from hcount)
    const int32_t* hids;            // HOT: hot vertex ids
    const uint64_t* hoff;           // HOT: hot list offsets
    const uint64_t* idx;            // CSR offsets (RANGE / BITMAP)
    const int32_t* edges;           // edge array of this launch
    const float* wts;               // per-edge coefficients aligned with edges, or null
    uint64_t ebase;                 // global edge index of edges[0] (RANGE / BITMAP)
    Piece* pieces;
    uint32_t* piece_count;          // this launch's piece counter
    uint32_t* piece_zero;           // the other counter (and batched p-row counter): cleared
    uint32_t piece_cap;             //   here for the next launch on this stream
    unsigned long long* counters;   // this launch's source row of C_FIELDS counters
    unsigned long long* overflow;
};

// Delta push of one block of entries (src/excent.cu:100-124 semantics).
// Phase A, one lane per entry: live check |r| > eps, atomic exchange of the residual,
// fold into x (a vertex is pushed by one lane only, so x has one writer), weight
// p = scale*d[/deg].  Entries of degree above HUB become pieces for k_pieces.
// Phase B: scatter p to every neighbour's residual with hardware float atomics.
//   G == 0: load-balanced -- lane j of the block takes adjacency slot j, j+256, ...
//           of the concatenated lists (owner found by binary search in LDS).
//   G > 0:  G lanes per entry (the reference's vertex groups, G in 1..64).
template <int NM, int S, int G>
__device__ __forceinline__ void push_entries(bool valid, int v, uint64_t ptr, uint64_t deg,
                                             const PushArgs& a, const DevMeasures& M,
                                             uint32_t* s_start, uint64_t* s_ptr, float (*s_p)[BLOCK],
                                             uint32_t* s_wave, unsigned long long* s_cnt) {
    const int tid = threadIdx.x;
    float p[NM];
    int nmv = 0;
    if (valid) {
        Row<S> row = load_row<S>(M.r, v);
#pragma unroll
        for (int k = 0; k < NM; k++) {
            p[k] = 0.f;
            if (fabsf(row.v[k]) > M.eps[k]) {
                float d = atomic_exch_f32(&M.r[(long long)v * S + k], 0.f);
                M.x[(long long)v * S + k] += d;
                if (d != 0.f) {
                    p[k] = M.divdeg[k] ? M.scale[k] * d / (float)deg : M.scale[k] * d;
                    nmv++;
                }
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < NM; k++) p[k] = 0.f;
    }
    const bool pushes = nmv > 0 && deg > 0;
    const bool hub = pushes && deg > HUB;
    if (hub) {
        const uint32_t np = (uint32_t)((deg + PIECE - 1) / PIECE);
        const uint32_t base = atomicAdd(a.piece_count, np);
        if (base + np <= a.piece_cap) {
            for (uint32_t j = 0; j < np; j++) {
                Piece pc;
                pc.start = ptr + (uint64_t)j * PIECE;
                uint64_t rem = deg - (uint64_t)j * PIECE;
                pc.len = (uint32_t)(rem < PIECE ? rem : PIECE);
                pc.pad = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) pc.p[k] = k < NM ? p[k] : 0.f;
                a.pieces[base + j] = pc;
            }
        } else {
            atomicAdd(a.overflow, 1ull);
        }
    }
    // block statistics: vertices that folded, edges pushed inline / as pieces
    {
        const uint32_t vv = nmv > 0 ? 1u : 0u;
        uint32_t t0 = wave_sum_u32(vv), t1 = wave_sum_u32(nmv);
        uint64_t de = pushes && !hub ? deg : 0, dp = hub ? deg : 0;
        uint64_t t2 = wave_sum_u64(de), t3 = wave_sum_u64(de * nmv);
        uint64_t t4 = wave_sum_u64(dp), t5 = wave_sum_u64(dp * nmv);
        if (lane_id() == 0 && t0) {
            atomicAdd(&s_cnt[0], (unsigned long long)t0);
            atomicAdd(&s_cnt[1], (unsigned long long)t1);
            atomicAdd(&s_cnt[2], (unsigned long long)t2);
            atomicAdd(&s_cnt[3], (unsigned long long)t3);
            atomicAdd(&s_cnt[4], (unsigned long long)t4);
            atomicAdd(&s_cnt[5], (unsigned long long)t5);
        }
    }
    const uint32_t mydeg = pushes && !hub ? (uint32_t)deg : 0u;
    if constexpr (G == 0) {
        uint32_t total;
        const uint32_t excl = block_excl_scan(mydeg, s_wave, &total);
        s_start[tid] = excl;
        s_ptr[tid] = ptr;
#pragma unroll
        for (int k = 0; k < NM; k++) s_p[k][tid] = p[k];
        __syncthreads();
        // PU slots per lane per round: their owners (LDS searches), then their ids loaded
        // together, then the atomics -- a block's rounds are a chain of id loads
        constexpr int PU = 4;
        int lo = 0;
        for (uint32_t j0 = tid; j0 < total; j0 += BLOCK * PU) {
            int own[PU], uq[PU];
            float wq[PU];
#pragma unroll
            for (int q = 0; q < PU; q++) {
                const uint32_t j = j0 + (uint32_t)q * BLOCK;
                own[q] = -1;
                if (j < total) {   // j grows with q, so the owner does too
                    int hi = BLOCK - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (s_start[mid] <= j) lo = mid; else hi = mid - 1;
                    }
                    own[q] = lo;
                }
            }
#pragma unroll
            for (int q = 0; q < PU; q++) {
                const uint64_t e = own[q] >= 0 ? s_ptr[own[q]] + (j0 + (uint32_t)q * BLOCK - s_start[own[q]]) : 0;
                uq[q] = own[q] >= 0 ? a.edges[e] : 0;
                wq[q] = own[q] >= 0 && a.wts ? a.wts[e] : 1.f;
            }
#pragma unroll
            for (int q = 0; q < PU; q++) {
                if (own[q] < 0) continue;
#pragma unroll
                for (int k = 0; k < NM; k++) {
                    const float pk = s_p[k][own[q]];
                    if (pk != 0.f) atomic_add_f32(&M.r[(long long)uq[q] * S + k], a.wts ? pk * wq[q] : pk);
                }
            }
        }
        __syncthreads();
    } else {
        s_start[tid] = mydeg;
        s_ptr[tid] = ptr;
#pragma unroll
        for (int k = 0; k < NM; k++) s_p[k][tid] = p[k];
        __syncthreads();
        const int gl = tid % G;
        for (int ent = tid / G; ent < BLOCK; ent += BLOCK / G) {
            const uint32_t d = s_start[ent];
            if (!d) continue;
            const uint64_t b = s_ptr[ent];
            float pk[NM];
#pragma unroll
            for (int k = 0; k < NM; k++) pk[k] = s_p[k][ent];
            for (uint32_t e = gl; e < d; e += G) {
                const int u = a.edges[b + e];
                const float we = a.wts ? a.wts[b + e] : 1.f;
#pragma unroll
                for (int k = 0; k < NM; k++)
                    if (pk[k] != 0.f) atomic_add_f32(&M.r[(long long)u * S + k], a.wts ? pk[k] * we : pk[k]);
            }
        }
        __syncthreads();
    }
}

template <int NM, int S, int SRC, int G>
__global__ __launch_bounds__(BLOCK) void k_push(PushArgs a, DevMeasures M) {
    __shared__ uint32_t s_start[BLOCK];
    __shared__ uint64_t s_ptr[BLOCK];
    __shared__ float s_p[NM][BLOCK];
    __shared__ uint32_t s_wave[NWAVES];
    __shared__ unsigned long long s_cnt[6];
    const int tid = threadIdx.x;
    if (tid < 6) s_cnt[tid] = 0;
    if (blockIdx.x == 0 && tid == 0) clear_counters(a.piece_zero);
    __syncthreads();

    // persistent over tiles of BLOCK entries (a window of BITMAP_WINDOW vertices in bitmap
    // mode), so each block flushes its statistics once
    long long ntiles = a.ntiles;
    if constexpr (SRC == SRC_HOT) ntiles = ((long long)*a.hcount + BLOCK - 1) / BLOCK;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    if constexpr (SRC == SRC_RANGE) {
        const long long v = (long long)a.v0 + tile * BLOCK + tid;
        bool valid = v < a.v1;
        if (valid && a.hot && ((a.hot[v >> 5] >> (v & 31)) & 1u)) valid = false;
        uint64_t b = valid ? a.idx[v] : 0, d = valid ? a.idx[v + 1] - b : 0;
        push_entries<NM, S, G>(valid, (int)(valid ? v : 0), b - (valid ? a.ebase : 0), d, a, M,
                               s_start, s_ptr, s_p, s_wave, s_cnt);
    } else if constexpr (SRC == SRC_LIST) {
        const long long i = tile * BLOCK + tid;
        const bool valid = i < a.k;
        const int v = valid ? a.vids[i] : 0;
        const uint64_t b = valid ? a.off[i] : 0, d = valid ? a.off[i + 1] - b : 0;
        push_entries<NM, S, G>(valid, v, b, d, a, M, s_start, s_ptr, s_p, s_wave, s_cnt);
    } else if constexpr (SRC == SRC_HOT) {
        const long long k = *a.hcount;
        const long long e = tile * BLOCK + tid;
        const bool valid = e < k;
        const int i = valid ? a.hlist[e] : 0;
        const int v = valid ? a.hids[i] : 0;
        const uint64_t b = valid ? a.hoff[i] : 0, d = valid ? a.hoff[i + 1] - b : 0;
        push_entries<NM, S, G>(valid, v, b, d, a, M, s_start, s_ptr, s_p, s_wave, s_cnt);
    } else {   // SRC_BITMAP: a window of BITMAP_WINDOW vertices, one bitmap word per lane
        __shared__ int32_t s_list[BITMAP_WINDOW];
        const long long w0 = ((long long)a.v0 >> 5) + tile * BLOCK;
        const long long word = w0 + tid;
        uint32_t bits = 0;
        const long long wv = word * 32;
        if (wv < a.v1) {
            bits = a.bitmap[word];
            if (wv < a.v0) bits &= ~0u << (a.v0 - wv);
            if (wv + 32 > a.v1) bits &= (a.v1 - wv) >= 32 ? ~0u : ((1u << (a.v1 - wv)) - 1u);
        }
        uint32_t total;
        uint32_t pos = block_excl_scan((uint32_t)__popc(bits), s_wave, &total);
        while (bits) {
            const int t = __ffs(bits) - 1;
            s_list[pos++] = (int32_t)(wv + t);
            bits &= bits - 1;
        }
        __syncthreads();
        for (uint32_t r0 = 0; r0 < total; r0 += BLOCK) {   // block-uniform rounds
            const uint32_t i = r0 + tid;
            const bool valid = i < total;
            const int v = valid ? s_list[i] : 0;
            uint64_t b = valid ? a.idx[v] : 0, d = valid ? a.idx[v + 1] - b : 0;
            push_entries<NM, S, G>(valid, v, b - (valid ? a.ebase : 0), d, a, M, s_start, s_ptr,
                                   s_p, s_wave, s_cnt);
        }
    }
    __syncthreads();   // the tile's LDS (s_list) is reused by the next tile
    }
    if (tid < 6 && s_cnt[tid]) atomicAdd(&a.counters[tid], s_cnt[tid]);
}

// Destinations privatised in LDS by the push of hub pieces.  The vertices with the most
// in-edges take a large share of every pass's scatter, and float atomics on one address
// serialise at the memory side; a block accumulates its deltas for them in LDS and adds each
// sum once at its end.  The set is a static open-addressing table (linear probing, at most
// `probes` steps) built on the host, with a filter of 2^16 bits in front of it so the other
// destinations cost one LDS read.  The sums reach r when the launch ends, like every other
// delta of the pass.
constexpr int HUB_SLOTS = 2048, HUB_MAX = HUB_SLOTS / 2, HUB_FILTER_WORDS = 2048;
struct HubTable {
    const int32_t* keys;       // [HUB_SLOTS], -1 = empty
    const uint32_t* filter;    // [HUB_FILTER_WORDS]
    int probes;                // longest probe sequence of a key
    int count;                 // keys in the table (0: no privatisation)
};
__host__ __device__ __forceinline__ uint32_t hub_slot(uint32_t u) { return (u * 2654435761u) >> 21; }   // 11 bits
__host__ __device__ __forceinline__ uint32_t hub_fbit(uint32_t u) {   // 16 bits, independent of hub_slot
    uint32_t h = u * 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    return h >> 16;
}

// Hub pieces: one wave per piece, persistent over the device-side count.  Counters are
// double-buffered per stream: launch i counts into buffer i%2 and clears buffer (i+1)%2,
// whose previous user (launch i-1 and its pieces) finished before launch i started.
template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_pieces(const Piece* __restrict__ pcs, const uint32_t* count,
                                                   uint32_t cap, const int32_t* __restrict__ edges,
                                                   const float* __restrict__ wts, DevMeasures M, HubTable ht) {
    __shared__ int32_t s_key[HUB_SLOTS];
    __shared__ uint32_t s_fil[HUB_FILTER_WORDS];
    __shared__ float s_acc[NM][HUB_SLOTS];
    // the count may exceed cap after an overflow (pieces beyond cap were not written and
    // the overflow counter reports it); never read past the buffer
    uint32_t cnt = *count;
    if (cnt > cap) cnt = cap;
    const long long nw = (long long)gridDim.x * NWAVES;
    const long long w0 = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE;
    if ((long long)blockIdx.x * NWAVES >= cnt) return;   // block-uniform: no piece for this block
    const bool priv = ht.count > 0;
    if (priv) {
        for (int i = threadIdx.x; i < HUB_SLOTS; i += BLOCK) {
            s_key[i] = ht.keys[i];
#pragma unroll
            for (int k = 0; k < NM; k++) s_acc[k][i] = 0.f;
        }
        for (int i = threadIdx.x; i < HUB_FILTER_WORDS; i += BLOCK) s_fil[i] = ht.filter[i];
        __syncthreads();
    }
    const int lane = lane_id();
    constexpr int PU = 8;   // edge ids per lane loaded together (a short pass is the chain of its longest piece)
    for (long long w = w0; w < cnt; w += nw) {
        const Piece pc = pcs[w];
        for (uint32_t t0 = 0; t0 < pc.len; t0 += WAVE * PU) {
            int uq[PU];
            float wq[PU];
#pragma unroll
            for (int q = 0; q < PU; q++) {
                const uint32_t t = t0 + (uint32_t)q * WAVE + lane;
                uq[q] = t < pc.len ? edges[pc.start + t] : -1;
                wq[q] = wts && t < pc.len ? wts[pc.start + t] : 1.f;
            }
#pragma unroll
            for (int q = 0; q < PU; q++) {
                const int u = uq[q];
                if (u < 0) continue;
                const float we = wq[q];
                int slot = -1;
                if (priv) {
                    const uint32_t fb = hub_fbit((uint32_t)u);
                    if ((s_fil[fb >> 5] >> (fb & 31)) & 1u) {
                        uint32_t h = hub_slot((uint32_t)u);
                        for (int r = 0; r < ht.probes; r++, h = (h + 1) & (HUB_SLOTS - 1)) {
                            const int key = s_key[h];
                            if (key == u) { slot = (int)h; break; }
                            if (key < 0) break;
                        }
                    }
                }
                if (slot >= 0) {
#pragma unroll
                    for (int k = 0; k < NM; k++)
                        if (pc.p[k] != 0.f) atomicAdd(&s_acc[k][slot], wts ? pc.p[k] * we : pc.p[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < NM; k++)
                        if (pc.p[k] != 0.f) atomic_add_f32(&M.r[(long long)u * S + k], wts ? pc.p[k] * we : pc.p[k]);
                }
            }
        }
    }
    if (priv) {   // one atomic per privatised destination and measure with a nonzero sum
        __syncthreads();
        for (int i = threadIdx.x; i < HUB_SLOTS; i += BLOCK) {
            const int key = s_key[i];
            if (key < 0) continue;
#pragma unroll
            for (int k = 0; k < NM; k++) {
                const float v = s_acc[k][i];
                if (v != 0.f) atomic_add_f32(&M.r[(long long)key * S + k], v);
            }
        }
    }
}

// k_compact_hot (src/excent.cu:178-182): hot entries in [lo, hi) pending now -> list.
template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_compact_hot(int lo, int hi, const int32_t* __restrict__ hids,
                                                        DevMeasures M, int32_t* out, uint32_t* count) {
    const long long i = lo + blockIdx.x * (long long)BLOCK + threadIdx.x;
    bool p = false;
    if (i < hi) {
        Row<S> row = load_row<S>(M.r, hids[i]);
#pragma unroll
        for (int k = 0; k < NM; k++) p |= fabsf(row.v[k]) > M.eps[k];
    }
    const uint64_t b = __ballot(p);
    if (!b) return;
    const int lane = lane_id();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(b));
    base = __shfl(base, 0, WAVE);
    if (p) out[base + __popcll(b & ((1ull << lane) - 1ull))] = (int32_t)i;
}

// k_fill_hot (src/excent.cu:212-220): one wave per hot entry copies its list from a
// chunk staged at src (global edge e0) into the hot arena.
__global__ __launch_bounds__(BLOCK) void k_fill_hot(int lo, int hi, const int32_t* __restrict__ hids,
                                                     const uint64_t* __restrict__ hoff,
                                                     const uint64_t* __restrict__ idx, uint64_t e0,
                                                     const int32_t* __restrict__ src, int32_t* hedges) {
    const long long nw = (long long)gridDim.x * NWAVES;
    for (long long w = lo + (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE; w < hi; w += nw) {
        const int32_t* from = src + (idx[hids[w]] - e0);
        int32_t* to = hedges + hoff[w];
        const uint64_t d = hoff[w + 1] - hoff[w];
        for (uint64_t i = lane_id(); i < d; i += WAVE) to[i] = from[i];
    }
}

__global__ void k_indegree(uint64_t m, const int32_t* __restrict__ edges, uint32_t* indeg) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)m;
         e += (long long)gridDim.x * blockDim.x)
        atomicAdd(&indeg[edges[e]], 1u);
}

// ------------------------------------------------------------------ dense passes: fold + pull

// A dense pass over the whole resident graph runs as two kernels: every pending vertex
// folds its residual and publishes its per-edge weight p (k_fold), then every vertex sums
// the p of its in-neighbours from the transpose and adds it to its residual (k_pull).  This
// is the delta push executed with all folds before all adds -- one of the schedules the
// asynchronous kernel may produce -- so x + r = b + W x holds exactly as after a push, but
// the scatter's random atomics become gathers plus one plain read-modify-write per vertex.

struct FoldArgs {
    int n;
    const uint64_t* idx;
    float* p;                       // [sources][NM] weights of this pass (0 when v did not push)
    const int32_t* rank;            // row of v in p: vertices with out-edges in decreasing
                                    // out-degree order, so the most-read rows share lines
    unsigned long long* counters;   // C_FIELDS row
};

template <int S>
__device__ __forceinline__ void store_row(float* base, long long v, const Row<S>& r) {
    if constexpr (S == 1) base[v] = r.v[0];
    else if constexpr (S == 2) reinterpret_cast<float2*>(base)[v] = make_float2(r.v[0], r.v[1]);
    else reinterpret_cast<float4*>(base)[v] = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
}

// No other kernel touches r or x while a dense pass runs (it is the only work of the pass,
// on one stream), so the fold reads and writes whole rows with plain vector accesses.
template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_fold(FoldArgs a, DevMeasures M) {
    __shared__ unsigned long long s_cnt[4];
    const int tid = threadIdx.x;
    if (tid < 4) s_cnt[tid] = 0;
    __syncthreads();
    uint32_t c_vert = 0, c_vnm = 0;
    uint64_t c_e = 0, c_enm = 0;
    for (long long v = blockIdx.x * (long long)BLOCK + tid; v < a.n; v += (long long)gridDim.x * BLOCK) {
        uint32_t nmv = 0;
        const uint64_t deg = a.idx[v + 1] - a.idx[v];
        Row<S> row = load_row<S>(M.r, v);
        Row<S> pr;
#pragma unroll
        for (int k = 0; k < S; k++) pr.v[k] = 0.f;
        Row<S> dx;
#pragma unroll
        for (int k = 0; k < S; k++) dx.v[k] = 0.f;
#pragma unroll
        for (int k = 0; k < NM; k++) {
            if (fabsf(row.v[k]) > M.eps[k]) {
                const float d = row.v[k];
                row.v[k] = 0.f;
                dx.v[k] = d;
                nmv++;
                if (deg) pr.v[k] = M.divdeg[k] ? M.scale[k] * d / (float)deg : M.scale[k] * d;
            }
        }
        if (nmv) {
            store_row<S>(M.r, v, row);
            Row<S> xr = load_row<S>(M.x, v);
#pragma unroll
            for (int k = 0; k < NM; k++) xr.v[k] += dx.v[k];
            store_row<S>(M.x, v, xr);
            c_vert++;
            c_vnm += nmv;
            c_e += deg;
            c_enm += deg * nmv;
        }
        if (deg) {   // vertices without out-edges are never read
            if constexpr (NM == S) store_row<S>(a.p, a.rank[v], pr);
            else {
                float* q = a.p + (long long)a.rank[v] * NM;
#pragma unroll
                for (int k = 0; k < NM; k++) q[k] = pr.v[k];
            }
        }
    }
    const uint32_t t0 = wave_sum_u32(c_vert), t1 = wave_sum_u32(c_vnm);
    const uint64_t t2 = wave_sum_u64(c_e), t3 = wave_sum_u64(c_enm);
    if (lane_id() == 0 && t0) {
        atomicAdd(&s_cnt[0], (unsigned long long)t0);
        atomicAdd(&s_cnt[1], (unsigned long long)t1);
        atomicAdd(&s_cnt[2], (unsigned long long)t2);
        atomicAdd(&s_cnt[3], (unsigned long long)t3);
    }
    __syncthreads();
    if (tid < 4 && s_cnt[tid]) atomicAdd(&a.counters[tid], s_cnt[tid]);
}

struct PullArgs {
    int n;
    const uint64_t* tidx;       // transpose offsets [n+1]
    const int32_t* tadj;        // in-neighbours, as rows of p (FoldArgs::rank)
    const float* p;             // [sources][NM]
    const uint32_t* live;       // bit q set iff row q of p is nonzero this pass (k_pbits);
                                // nullptr: every row is read
    unsigned long long* live_count;   // k_pull_hub: rows read
    const float* tw;            // weighted graphs: each in-edge's coefficient, aligned with tadj
    // k_pull: the next pass's snapshot (bitmap + per-chunk counts), fused into the final
    // row update; bitmap == nullptr: not fused
    const uint64_t* idx;
    const int* cv0;
    int nc;
    uint32_t* bitmap;
    unsigned long long* cnt;
    // k_pull: the next pass's fold, fused after the snapshot test (p_next == nullptr: not
    // fused): pending vertices move r into x and publish their row of p_next at rank[u]
    float* p_next;
    const int32_t* rank;
    int nsrc;
    unsigned long long* spread;   // C_SPREAD area of the counters
    double* l1;                   // fused snapshot: FP64 residual sum slots, or null
    int rows_field;               // spread field that counts the rows read (S_ROWS, S_ROWS_A1/A2)
};

// Bitmap of the p rows the fold made nonzero: one wave per 64 consecutive rows, one ballot,
// two words.  Rows of sources that did not push (or pushed a zero weight) stay zero, so a
// gather of them would add nothing; the pull tests the bit and skips the row.  Late dense
// passes, where a few high-degree sources still push, then read the in-lists and the bitmap
// but almost no rows.
template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_pbits(const float* __restrict__ p, int nsrc, uint32_t* bits) {
    const int lane = lane_id();
    const long long waves = (long long)gridDim.x * NWAVES;
    for (long long w = blockIdx.x * (long long)NWAVES + threadIdx.x / WAVE; w * WAVE < nsrc; w += waves) {
        const long long q = w * WAVE + lane;
        bool nz = false;
        if (q < nsrc) {
            const Row<S> row = load_prow<NM, S>(p, q);
#pragma unroll
            for (int k = 0; k < NM; k++) nz |= row.v[k] != 0.f;
        }
        const uint64_t b = __ballot(nz);
        if (lane < 2) bits[2 * w + lane] = (uint32_t)(b >> (32 * lane));
    }
}

// Rows of p written with IN columns (a fold over IN measures) repacked to their first OUT
// columns, into another buffer.  A dense pass that drops measures right after a pull that
// folded it finds p in the wider layout; the dropped columns are zero there (their measures
// had nothing pending when the fold ran), so keeping the first OUT columns is exact.
template <int IN, int OUT>
__global__ __launch_bounds__(BLOCK) void k_prepack(const float* __restrict__ src, float* __restrict__ dst, int nsrc) {
    for (long long q = blockIdx.x * (long long)BLOCK + threadIdx.x; q < nsrc; q += (long long)gridDim.x * BLOCK) {
        float v[OUT];
#pragma unroll
        for (int k = 0; k < OUT; k++) v[k] = src[q * IN + k];
#pragma unroll
        for (int k = 0; k < OUT; k++) dst[q * OUT + k] = v[k];
    }
}

__device__ __forceinline__ bool row_live(const uint32_t* __restrict__ live, int q) {
    return !live || ((live[q >> 5] >> (q & 31)) & 1u);
}

// The end of a dense pull for BLOCK consecutive destinations u0 .. u0 + nt - 1, one per
// thread, with `acc` the thread's sums of in-edge rows: the row update of r, and when the
// next snapshot is fused (bitmap != nullptr) the pending test of the next pass -- the same
// predicate and outputs as k_snapshot without hot set or trace: two bitmap words per wave,
// per-chunk counts reduced per block, the FP64 residual of the snapshot state -- and when
// p_next is set, the next pass's fold on the row in registers.  Every row is final here
// (the hub rows were completed by k_pull_hub before the pull).  `who` spreads the residual
// slots.  Block-uniform: every thread of the block calls it.
template <int NM, int S>
__device__ void pull_finish(const PullArgs& a, const DevMeasures& M, long long u0, int nt, const float (&acc)[NM],
                            unsigned long long* slot, long long who) {
    __shared__ unsigned long long s_cnt[NWAVES][2];
    __shared__ unsigned long long s_fold[NWAVES][4];
    __shared__ double s_l1[NWAVES][NM];
    __shared__ uint32_t s_mx[NWAVES][NM];
    const int tid = threadIdx.x;
    if (!a.bitmap) {
        if (tid < nt) {
            const long long u = u0 + tid;
#pragma unroll
            for (int k = 0; k < NM; k++)
                if (acc[k] != 0.f) M.r[u * S + k] += acc[k];
        }
        return;
    }
    bool pend = false;
    uint64_t d = 0;
    uint32_t nmv = 0;
    double l1[NM];
    uint32_t mx[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) { l1[k] = 0.0; mx[k] = 0u; }
    if (tid < nt) {
        const long long u = u0 + tid;
        Row<S> row = load_row<S>(M.r, u);
        bool upd = false;
#pragma unroll
        for (int k = 0; k < NM; k++)
            if (acc[k] != 0.f) { row.v[k] += acc[k]; upd = true; }
        // the snapshot state's residual: the row after this pass, before the fused fold
#pragma unroll
        for (int k = 0; k < NM; k++) {
            l1[k] = fabs((double)row.v[k]);
            mx[k] = __float_as_uint(fabsf(row.v[k]));
        }
#pragma unroll
        for (int k = 0; k < NM; k++) pend |= fabsf(row.v[k]) > M.eps[k];
        if (pend) d = a.idx[u + 1] - a.idx[u];
        if (a.p_next) {   // the next pass's fold (k_fold's rule, on the row in registers)
            Row<S> pr, dx;
#pragma unroll
            for (int k = 0; k < S; k++) { pr.v[k] = 0.f; dx.v[k] = 0.f; }
            if (pend) {
#pragma unroll
                for (int k = 0; k < NM; k++) {
                    if (fabsf(row.v[k]) > M.eps[k]) {
                        const float dv = row.v[k];
                        row.v[k] = 0.f;
                        dx.v[k] = dv;
                        nmv++;
                        if (d) pr.v[k] = M.divdeg[k] ? M.scale[k] * dv / (float)d : M.scale[k] * dv;
                    }
                }
                Row<S> xr = load_row<S>(M.x, u);
#pragma unroll
                for (int k = 0; k < NM; k++) xr.v[k] += dx.v[k];
                store_row<S>(M.x, u, xr);
                upd = true;
            }
            const int q = a.rank[u];
            if (q < a.nsrc) {
                float* pq = a.p_next + (long long)q * NM;
#pragma unroll
                for (int k = 0; k < NM; k++) pq[k] = pr.v[k];
            }
        }
        if (upd) store_row<S>(M.r, u, row);
    }
    const uint64_t mask = __ballot(pend);
    const long long wb = u0 + (tid & ~(WAVE - 1));
    const int lane = lane_id();
    if (lane == 0 && wb < u0 + nt) {   // the waves holding this call's destinations
        const long long w0 = wb >> 5, words = ((long long)a.n + 31) / 32;
        a.bitmap[w0] = (uint32_t)mask;
        if (w0 + 1 < words) a.bitmap[w0 + 1] = (uint32_t)(mask >> 32);
    }
    if (a.l1) {   // block sums and maxima, one atomic each per measure and block
#pragma unroll
        for (int k = 0; k < NM; k++) {
            const double t = wave_sum_f64(l1[k]);
            const uint32_t m = wave_max_u32(mx[k]);
            if (lane == 0) { s_l1[tid / WAVE][k] = t; s_mx[tid / WAVE][k] = m; }
        }
        __syncthreads();
        if (tid < NM) {
            double t = 0.0;
            uint32_t m = 0u;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) { t += s_l1[w][tid]; m = max(m, s_mx[w][tid]); }
            l1_flush(a.l1, who, tid, t);
            linf_flush(a.l1, who, tid, m);
        }
    }
    const int c0 = chunk_of((int)u0, a.cv0, a.nc), c1 = chunk_of((int)(u0 + nt - 1), a.cv0, a.nc);
    if (c0 == c1) {   // the usual case: the destinations lie in one chunk
        const uint32_t tn = wave_sum_u32(pend ? 1u : 0u);
        const uint64_t td = wave_sum_u64(d);
        if (lane == 0) { s_cnt[tid / WAVE][0] = tn; s_cnt[tid / WAVE][1] = td; }
        __syncthreads();
        if (tid < 2) {   // thread 0: pending vertices, thread 1: their out-degree sum
            unsigned long long t = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) t += s_cnt[w][tid];
            if (t) atomicAdd(&a.cnt[2 * c0 + tid], t);
        }
    } else if (pend) {
        const int c = chunk_of((int)(u0 + tid), a.cv0, a.nc);
        atomicAdd(&a.cnt[2 * c], 1ull);
        atomicAdd(&a.cnt[2 * c + 1], (unsigned long long)d);
    }
    if (a.p_next) {   // fused fold statistics, one atomic per field and block (spread slots)
        const uint64_t f[4] = {nmv ? 1ull : 0ull, nmv, nmv ? d : 0ull, d * nmv};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint64_t t = wave_sum_u64(f[i]);
            if (lane == 0) s_fold[tid / WAVE][i] = t;
        }
        __syncthreads();
        if (tid < 4) {
            unsigned long long sum = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) sum += s_fold[w][tid];
            if (sum) atomicAdd(&slot[S_VERT + tid], sum);
        }
    }
    __syncthreads();   // a next call by the same block reuses the shared arrays
}

constexpr uint32_t PULL_HUB = 4096;   // in-lists longer than this are summed by whole blocks

// Sequential-slot pull.  A block owns BLOCK consecutive destinations; with hub in-lists
// stored apart (k_pull_hub), their in-lists are one contiguous run of tadj.  The run is
// staged through LDS in windows of PULL_WIN slots (coalesced global reads); lane t then
// takes PULL_SEQ consecutive slots of the window, issues all their p-row loads at once
// (memory-level parallelism), and sums them in registers, walking the destination
// boundaries; a partial sum goes to the destination's LDS accumulator when the boundary is
// crossed (a list spans few lanes, so the LDS float atomics rarely collide).  One binary
// search per lane per window replaces a search per slot, and no cross-lane reduction is
// needed.  Finally each destination adds its sum to r with one plain row update.
// live != nullptr: rows whose bit is clear are zero this pass and are not loaded.
constexpr int PULL_LD = BLOCK + 1;           // LDS row pitch: slot j at (j % SEQ) * LD + j / SEQ

// W: weighted graph -- each row is scaled by its in-edge's coefficient (staged with the ids).
// TILE: destinations per block (BLOCK, or 64 on graphs too small to give every CU four
// blocks of BLOCK destinations: a quarter of the lanes then finish the tile, but four times
// the blocks keep gathers in flight).
template <int NM, int S, int PULL_SEQ, bool W = false, int TILE = BLOCK>
__global__ __launch_bounds__(BLOCK, 8) void k_pull(PullArgs a, DevMeasures M) {
    static_assert(TILE % WAVE == 0 && TILE <= BLOCK, "a tile is whole waves of destinations");
    constexpr int PULL_WIN = BLOCK * PULL_SEQ;   // 2048 slots (8 KB of LDS) at PULL_SEQ 8
    __shared__ uint32_t s_start[BLOCK + 1];
    __shared__ float s_w[W ? PULL_SEQ * PULL_LD : 1];
    __shared__ int32_t s_src[PULL_SEQ * PULL_LD];
    __shared__ float s_acc[NM][BLOCK];
    const int tid = threadIdx.x;
    const long long u0 = (long long)blockIdx.x * TILE;
    const int nt = (int)((long long)a.n - u0 < TILE ? (long long)a.n - u0 : TILE);
    const uint64_t base = a.tidx[u0];
    for (int i = tid; i <= nt; i += BLOCK) s_start[i] = (uint32_t)(a.tidx[u0 + i] - base);
#pragma unroll
    for (int k = 0; k < NM; k++) s_acc[k][tid] = 0.f;
    __syncthreads();
    const uint32_t total = s_start[nt];
    uint32_t taken = 0;   // p rows read by this lane
    // Thread tid stages slots tid + k BLOCK (k < PULL_SEQ) of each window.  The next window's
    // ids (and coefficients) are loaded into registers while this window's rows are gathered,
    // so a window waits for one round trip to memory, not two.
    static_assert(BLOCK % PULL_SEQ == 0, "slot layout");
    int32_t nid[PULL_SEQ];
    float nwt[W ? PULL_SEQ : 1];
    auto fetch = [&](uint32_t w0) {
#pragma unroll
        for (int k = 0; k < PULL_SEQ; k++) {
            const uint32_t j = w0 + (uint32_t)tid + (uint32_t)k * BLOCK;
            nid[k] = j < total ? a.tadj[base + j] : 0;
            if constexpr (W) nwt[k] = j < total ? a.tw[base + j] : 0.f;
        }
    };
    fetch(0);
    for (uint32_t w0 = 0; w0 < total; w0 += PULL_WIN) {
        const uint32_t wn = total - w0 < (uint32_t)PULL_WIN ? total - w0 : (uint32_t)PULL_WIN;
#pragma unroll
        for (int k = 0; k < PULL_SEQ; k++) {
            const uint32_t j = (uint32_t)tid + (uint32_t)k * BLOCK;
            if (j < wn) {
                s_src[(j % PULL_SEQ) * PULL_LD + j / PULL_SEQ] = nid[k];
                if constexpr (W) s_w[(j % PULL_SEQ) * PULL_LD + j / PULL_SEQ] = nwt[k];
            }
        }
        __syncthreads();
        if (w0 + PULL_WIN < total) fetch(w0 + PULL_WIN);
        const uint32_t j0 = (uint32_t)tid * PULL_SEQ;
        if (j0 < wn) {
            const uint32_t jn = wn - j0 < (uint32_t)PULL_SEQ ? wn - j0 : (uint32_t)PULL_SEQ;
            const uint32_t g = w0 + j0;   // first slot of this lane, relative to base
            // ids of the lane's slots, then (live-row passes) their bitmap words, then the
            // rows: each stage is issued as one batch and used only after it, so a lane waits
            // once per stage, not once per slot
            int vid[PULL_SEQ];
#pragma unroll
            for (int q = 0; q < PULL_SEQ; q++) vid[q] = (uint32_t)q < jn ? s_src[q * PULL_LD + tid] : -1;
            bool lv[PULL_SEQ];
            if (a.live) {
                uint32_t wd[PULL_SEQ];
#pragma unroll
                for (int q = 0; q < PULL_SEQ; q++) wd[q] = vid[q] >= 0 ? a.live[vid[q] >> 5] : 0u;
#pragma unroll
                for (int q = 0; q < PULL_SEQ; q++) lv[q] = vid[q] >= 0 && ((wd[q] >> (vid[q] & 31)) & 1u);
            } else {
#pragma unroll
                for (int q = 0; q < PULL_SEQ; q++) lv[q] = vid[q] >= 0;
            }
            Row<S> rows[PULL_SEQ];
#pragma unroll
            for (int q = 0; q < PULL_SEQ; q++) {
#pragma unroll
                for (int k = 0; k < S; k++) rows[q].v[k] = 0.f;
                if (lv[q]) {
                    rows[q] = load_prow<NM, S>(a.p, vid[q]);
                    taken++;
                }
            }
            if constexpr (W) {   // rows of non-live slots are zero
#pragma unroll
                for (int q = 0; q < PULL_SEQ; q++) {
                    const float w = (uint32_t)q < jn ? s_w[q * PULL_LD + tid] : 0.f;
#pragma unroll
                    for (int k = 0; k < NM; k++) rows[q].v[k] *= w;
                }
            }
            // owner: the destination whose range holds slot g (searched while the rows above
            // are in flight)
            int lo = 0, hi = nt - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_start[mid] <= g) lo = mid; else hi = mid - 1;
            }
            int owner = lo;
            uint32_t next = s_start[owner + 1];
            float acc[NM];
#pragma unroll
            for (int k = 0; k < NM; k++) acc[k] = 0.f;
#pragma unroll
            for (int q = 0; q < PULL_SEQ; q++) {
                if ((uint32_t)q < jn) {
                    while (g + q >= next) {   // crossing into the next non-empty list
#pragma unroll
                        for (int k = 0; k < NM; k++)
                            if (acc[k] != 0.f) { atomicAdd(&s_acc[k][owner], acc[k]); acc[k] = 0.f; }
                        owner++;
                        next = s_start[owner + 1];
                    }
#pragma unroll
                    for (int k = 0; k < NM; k++) acc[k] += rows[q].v[k];
                }
            }
#pragma unroll
            for (int k = 0; k < NM; k++)
                if (acc[k] != 0.f) atomicAdd(&s_acc[k][owner], acc[k]);
        }
        __syncthreads();   // the next window rewrites s_src
    }
    __shared__ unsigned long long s_red[NWAVES];
    unsigned long long* slot = a.spread + (blockIdx.x % SPREAD_SLOTS) * SPREAD_STRIDE;
    {
        const uint32_t t = wave_sum_u32(taken);
        if (lane_id() == 0) s_red[tid / WAVE] = t;
        __syncthreads();
        if (tid == 0) {
            unsigned long long sum = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) sum += s_red[w];
            if (sum) atomicAdd(&slot[a.rows_field], sum);
        }
        __syncthreads();
    }
    float accv[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) accv[k] = tid < nt && s_start[tid + 1] > s_start[tid] ? s_acc[k][tid] : 0.f;
    pull_finish<NM, S>(a, M, u0, nt, accv, slot, blockIdx.x);
}

// Source-sorted dense pull.  A block owns T consecutive destinations; their non-hub in-edges
// are stored per tile sorted by source row (row id + the destination's offset in the tile),
// and the block streams them in that order: lane t takes entries t, t + BLOCK, ... of each
// step of BLOCK x U entries, so a wave gathers 64 consecutive entries of the sorted order --
// rows of the same or neighbouring sources, which share cache lines, are served by one
// request per line instead of one per lane.  Each entry's row is added to its destination's
// accumulator in LDS (ds_add_f32).  The destinations are then finished BLOCK at a time by
// pull_finish (row update, fused snapshot and fold), as in k_pull.
struct SortedTiles {
    const uint64_t* start;   // [tiles + 1]
    const int32_t* row;      // p row of each entry, ascending within a tile
    const uint16_t* dst;     // destination offset in the tile
    const float* w;          // weighted graphs: the entry's coefficient
};

// A tile's accumulators (NM x T floats) must fit the block's LDS budget of 64 KB.
constexpr bool sorted_fits(int nm, int T) { return (long long)nm * T * 4 <= 65536; }

template <int NM, int S, int T, bool W = false>
__global__ __launch_bounds__(BLOCK) void k_pull_sorted(PullArgs a, SortedTiles st, DevMeasures M) {
    static_assert(sorted_fits(NM, T), "tile accumulators exceed the LDS budget");
    constexpr int U = 8;   // entries per lane in flight
    __shared__ float s_acc[NM][T];
    __shared__ unsigned long long s_red[NWAVES];
    const int tid = threadIdx.x;
    const long long u0 = (long long)blockIdx.x * T;
    for (int i = tid; i < T; i += BLOCK) {
#pragma unroll
        for (int k = 0; k < NM; k++) s_acc[k][i] = 0.f;
    }
    __syncthreads();
    const uint64_t e0 = st.start[blockIdx.x], e1 = st.start[blockIdx.x + 1];
    uint32_t taken = 0;
    const uint64_t last = e1 ? e1 - 1 : 0;
    for (uint64_t b = e0; b < e1; b += (uint64_t)BLOCK * U) {
        // ids, destinations and coefficients of the lane's entries, then their live-bitmap
        // words, then the rows: each stage issued as one batch (the ids without branches: an
        // entry past the tile's end reads the tile's last entry and is not live)
        int vid[U];
        uint32_t dl[U];
        float wv[W ? U : 1];
        bool in[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t e = b + (uint64_t)u * BLOCK + tid;
            in[u] = e < e1;
            const uint64_t ec = in[u] ? e : last;
            vid[u] = st.row[ec];
            dl[u] = st.dst[ec];
            if constexpr (W) wv[u] = st.w[ec];
        }
        bool lv[U];
        if (a.live) {
            uint32_t wd[U];
#pragma unroll
            for (int u = 0; u < U; u++) wd[u] = a.live[vid[u] >> 5];
#pragma unroll
            for (int u = 0; u < U; u++) lv[u] = in[u] && ((wd[u] >> (vid[u] & 31)) & 1u);
        } else {
#pragma unroll
            for (int u = 0; u < U; u++) lv[u] = in[u];
        }
        Row<S> rows[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int k = 0; k < S; k++) rows[u].v[k] = 0.f;
            if (lv[u]) {
                rows[u] = load_prow<NM, S>(a.p, vid[u]);
                taken++;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int k = 0; k < NM; k++) {
                const float v = W ? rows[u].v[k] * wv[u] : rows[u].v[k];
                if (v != 0.f) atomicAdd(&s_acc[k][dl[u]], v);
            }
        }
    }
    __syncthreads();
    unsigned long long* slot = a.spread + (blockIdx.x % SPREAD_SLOTS) * SPREAD_STRIDE;
    {
        const uint32_t t = wave_sum_u32(taken);
        if (lane_id() == 0) s_red[tid / WAVE] = t;
        __syncthreads();
        if (tid == 0) {
            unsigned long long sum = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) sum += s_red[w];
            if (sum) atomicAdd(&slot[a.rows_field], sum);
        }
    }
    for (int j = 0; j < T / BLOCK; j++) {   // block-uniform
        const long long ub = u0 + (long long)j * BLOCK;
        if (ub >= a.n) break;
        const int nt = (int)(a.n - ub < BLOCK ? a.n - ub : BLOCK);
        float accv[NM];
#pragma unroll
        for (int k = 0; k < NM; k++) accv[k] = tid < nt ? s_acc[k][j * BLOCK + tid] : 0.f;
        pull_finish<NM, S>(a, M, ub, nt, accv, slot, (long long)blockIdx.x * (T / BLOCK) + j);
    }
}

// Destinations with more than PULL_HUB in-neighbours: their in-lists are cut into pieces of
// PULL_PIECE edges at setup; one block per piece reduces and adds its sum to r with one
// atomic per measure (pieces of one destination may run concurrently).  Runs before
// k_pull in the pass; the two kernels touch disjoint rows of r.
struct PullPiece { int32_t u; uint32_t len; uint64_t start; };
constexpr uint32_t PULL_PIECE = 2048;   // one step of BLOCK x HUB_UNROLL entries
// Batched contexts sum a piece with one wave (k_bpull_hub), so their pieces are shorter: a
// wave's rows are a serial chain, and more pieces keep more waves busy.
constexpr uint32_t BPULL_PIECE = 1024;
constexpr int HUB_UNROLL = 8;

template <int NM, int S>
__global__ __launch_bounds__(BLOCK) void k_pull_hub(PullArgs a, const PullPiece* __restrict__ pieces, int npieces,
                                                     DevMeasures M) {
    __shared__ float s_red[NM][NWAVES];
    __shared__ unsigned long long s_live;
    if (threadIdx.x == 0) s_live = 0;
    __syncthreads();
    uint32_t taken = 0;
    for (int h = blockIdx.x; h < npieces; h += gridDim.x) {
        const PullPiece pc = pieces[h];
        const int u = pc.u;
        const uint64_t b = pc.start, e = pc.start + pc.len;
        float acc[NM];
#pragma unroll
        for (int k = 0; k < NM; k++) acc[k] = 0.f;
        // HUB_UNROLL entries per lane per step: their ids, then their rows, are loaded
        // together so each lane keeps several gathers in flight
        for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += (uint64_t)BLOCK * HUB_UNROLL) {
            int vs[HUB_UNROLL];
            float ws[HUB_UNROLL];
#pragma unroll
            for (int q = 0; q < HUB_UNROLL; q++) {
                const uint64_t i = i0 + (uint64_t)q * BLOCK;
                vs[q] = i < e ? a.tadj[i] : -1;
                ws[q] = a.tw && i < e ? a.tw[i] : 1.f;
            }
            // the live-bitmap words, then the rows, each as one batch (see k_pull)
            bool lv[HUB_UNROLL];
            if (a.live) {
                uint32_t wd[HUB_UNROLL];
#pragma unroll
                for (int q = 0; q < HUB_UNROLL; q++) wd[q] = vs[q] >= 0 ? a.live[vs[q] >> 5] : 0u;
#pragma unroll
                for (int q = 0; q < HUB_UNROLL; q++) lv[q] = vs[q] >= 0 && ((wd[q] >> (vs[q] & 31)) & 1u);
            } else {
#pragma unroll
                for (int q = 0; q < HUB_UNROLL; q++) lv[q] = vs[q] >= 0;
            }
            Row<S> rows[HUB_UNROLL];
#pragma unroll
            for (int q = 0; q < HUB_UNROLL; q++) {
#pragma unroll
                for (int k = 0; k < S; k++) rows[q].v[k] = 0.f;
                if (lv[q]) {
                    rows[q] = load_prow<NM, S>(a.p, vs[q]);
                    taken++;
                }
            }
            if (a.tw) {   // rows of non-live slots are zero
#pragma unroll
                for (int q = 0; q < HUB_UNROLL; q++)
#pragma unroll
                    for (int k = 0; k < NM; k++) rows[q].v[k] *= ws[q];
            }
#pragma unroll
            for (int q = 0; q < HUB_UNROLL; q++)
#pragma unroll
                for (int k = 0; k < NM; k++) acc[k] += rows[q].v[k];
        }
#pragma unroll
        for (int k = 0; k < NM; k++) {
            float t = acc[k];
            for (int o = WAVE / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, WAVE);
            if (lane_id() == 0) s_red[k][threadIdx.x / WAVE] = t;
        }
        __syncthreads();
        if (threadIdx.x < NM) {
            float t = 0.f;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) t += s_red[threadIdx.x][w];
            if (t != 0.f) atomic_add_f32(&M.r[(long long)u * S + threadIdx.x], t);
        }
        __syncthreads();
    }
    const uint32_t t = wave_sum_u32(taken);
    if (lane_id() == 0 && t) atomicAdd(&s_live, (unsigned long long)t);
    __syncthreads();
    if (threadIdx.x == 0 && s_live) atomicAdd(a.live_count, s_live);
}

// Transpose build: in-degree counts, then a fill through per-vertex cursors.
__global__ void k_count_in64(uint64_t m, const int32_t* __restrict__ adj, unsigned long long* cnt) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)m; e += (long long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[adj[e]], 1ull);
}
// Transpose fill, one thread per edge: a block takes FILL_EDGES consecutive edges, finds the
// sources of its first and last edge by binary search in idx, stages the offsets of the
// sources in between in LDS (when they fit; else each edge searches idx itself) and each
// edge finds its source there -- the last v with idx[v] <= e, so sources without edges
// (equal offsets) are passed over.  A hub's out-list is spread over many blocks instead of
// one lane.  In-list order follows the atomics' order.
constexpr int FILL_EDGES = 4096, FILL_STAGE = 4096;
__device__ __forceinline__ long long last_le(const uint64_t* __restrict__ a, long long lo, long long hi, uint64_t key) {
    while (lo < hi) {   // the last i in [lo, hi] with a[i] <= key (a[lo] <= key)
        const long long mid = (lo + hi + 1) >> 1;
        if (a[mid] <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__global__ __launch_bounds__(BLOCK) void k_fill_in64(int n, uint64_t m, const uint64_t* __restrict__ idx,
                                                     const int32_t* __restrict__ adj, unsigned long long* cursor,
                                                     const int32_t* __restrict__ rank, int32_t* tadj,
                                                     const float* __restrict__ w, float* tw) {
    __shared__ uint64_t s_idx[FILL_STAGE + 1];
    __shared__ long long s_v[2];
    for (uint64_t e0 = (uint64_t)blockIdx.x * FILL_EDGES; e0 < m; e0 += (uint64_t)gridDim.x * FILL_EDGES) {
        const uint64_t e1 = e0 + FILL_EDGES < m ? e0 + FILL_EDGES : m;
        if (threadIdx.x < 2) s_v[threadIdx.x] = last_le(idx, 0, n, threadIdx.x == 0 ? e0 : e1 - 1);
        __syncthreads();
        const long long v0 = s_v[0], v1 = s_v[1];
        const long long span = v1 - v0 + 1;   // sources of this block's edges
        const bool staged = span <= FILL_STAGE;
        if (staged)
            for (long long i = threadIdx.x; i < span; i += BLOCK) s_idx[i] = idx[v0 + i];
        __syncthreads();
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += BLOCK) {
            const long long v = staged ? v0 + last_le(s_idx, 0, span - 1, e) : last_le(idx, v0, v1, e);
            const unsigned long long q = atomicAdd(&cursor[adj[e]], 1ull);
            tadj[q] = rank ? rank[v] : (int32_t)v;
            if (w) tw[q] = w[e];
        }
        __syncthreads();   // s_idx and s_v are rewritten by the next range
    }
}

// ------------------------------------------------------------------ batched PPR

// K personalised PageRank vectors share the graph: x and r are [n][K] rows (K <= 64),
// one wave per vertex with lane j owning column j.  Every neighbour update is then one
// wave-wide atomic on a contiguous K*4-byte row (coalesced), and one read of a list
// serves all K columns.  Columns with nothing pending drop out through `active`.
struct BatchArgs {
    float* x;
    float* r;
    const float* eps;     // [K]
    float damp;
    int K;
    uint64_t active;      // bit j: column j still has pending vertices
    uint32_t* colpend;    // [K] snapshot: pending vertices per column
    float* prows;         // hub p rows [cap][K]
    uint32_t* prow_count;
    uint32_t prow_cap;
    uint8_t* touch;       // [n] or null: pushes set touch[u] = 1 for every destination they add to
};

struct BPiece {          // 16 bytes: a slice of one hub vertex's list
    uint64_t start;
    uint32_t len;
    uint32_t prow;
};

constexpr uint32_t BHUB = 2048;    // batched: lists longer than this are split
constexpr uint32_t BPIECE = 2048;

// Snapshot of a [n][K] state: vertex v is pending iff any active column j has
// |r[v][j]| > eps[j].  Same bitmap / per-chunk counts / hot / trace outputs as
// k_snapshot, plus per-column pending counts.  One wave per 64-vertex segment.
// FRONT: the snapshot after a push pass that pushed every vertex the last snapshot found
// pending.  A pushed vertex's r ends at zero in the columns it pushed and at most eps in the
// others, and a vertex that was not pending is unchanged, unless a push added to it; so only
// the vertices the pass's pushes touched can be pending.  The kernel reads (and clears) the
// touch bytes and reads only those vertices' rows; every other bit is zero.  (Hot and trace
// outputs are off in this mode.)
template <bool FRONT>
__global__ __launch_bounds__(BLOCK) void k_bsnapshot(SnapArgs a, BatchArgs B) {
    __shared__ uint32_t s_col[NWAVES][WAVE];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const long long gwave = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE;
    const long long nseg = ((long long)a.n + WAVE - 1) / WAVE;
    const long long words = ((long long)a.n + 31) / 32;
    const bool col_on = lane < B.K && ((B.active >> lane) & 1ull);
    const float eps = lane < B.K ? B.eps[lane] : 0.f;
    uint32_t colcnt = 0;
    long long s0 = gwave * a.segs_per_wave, s1 = s0 + a.segs_per_wave;
    if (s1 > nseg) s1 = nseg;
    int cc = s0 < nseg ? chunk_of((int)(s0 * WAVE), a.cv0, a.nc) : 0;
    int clo = a.cv0[cc], chi = cc + 1 < a.nc ? a.cv0[cc + 1] : a.n;
    uint32_t acc_n = 0;
    uint64_t acc_d = 0, hp = 0;
    auto flush = [&]() {
        uint32_t tn = wave_sum_u32(acc_n);
        uint64_t td = wave_sum_u64(acc_d);
        if (lane == 0 && tn) {
            atomicAdd(&a.cnt[2 * cc], (unsigned long long)tn);
            atomicAdd(&a.cnt[2 * cc + 1], (unsigned long long)td);
        }
        acc_n = 0;
        acc_d = 0;
    };
    for (long long seg = s0; seg < s1; seg++) {
        const long long base = seg * WAVE;
        uint64_t mask = 0;
        constexpr int U = 8;
        if constexpr (FRONT) {
            const long long vt = base + lane;
            const bool t = vt < a.n && B.touch[vt] != 0;
            uint64_t tm = __ballot(t);
            if (t) B.touch[vt] = 0;   // clean for the next push pass
            while (tm) {   // the touched vertices, U rows at a time
                int off[U];
                float row[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    off[u] = tm ? __ffsll((long long)tm) - 1 : -1;
                    tm &= tm - 1;
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                    row[u] = (off[u] >= 0 && col_on) ? B.r[(base + off[u]) * B.K + lane] : 0.f;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool pj = col_on && off[u] >= 0 && fabsf(row[u]) > eps;
                    colcnt += pj ? 1u : 0u;
                    if (__ballot(pj)) mask |= 1ull << off[u];
                }
            }
        } else {
            for (int i0 = 0; i0 < WAVE; i0 += U) {
                float row[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const long long v = base + i0 + u;
                    row[u] = (v < a.n && col_on) ? B.r[v * B.K + lane] : 0.f;   // inactive columns: no load
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const bool pj = col_on && fabsf(row[u]) > eps && base + i0 + u < a.n;
                    colcnt += pj ? 1u : 0u;
                    if (__ballot(pj)) mask |= 1ull << (i0 + u);
                }
            }
        }
        const long long v = base + lane;
        const bool in = v < a.n;
        bool pend = in && ((mask >> lane) & 1ull);
        if (a.trace && pend) a.trace[v]++;
        if (a.hot) {
            bool ph = pend && ((a.hot[v >> 5] >> (v & 31)) & 1u);
            hp += ph ? 1 : 0;
            pend = pend && !ph;
        }
        const uint64_t bm = __ballot(pend);
        if (lane == 0) {
            a.bitmap[seg * 2] = (uint32_t)bm;
            if (seg * 2 + 1 < words) a.bitmap[seg * 2 + 1] = (uint32_t)(bm >> 32);
        }
        if (!bm) continue;
        const uint64_t d = pend ? a.idx[v + 1] - a.idx[v] : 0;
        long long vlast = base + WAVE - 1;
        if (vlast >= a.n) vlast = a.n - 1;
        if (base >= clo && vlast < chi) {
            acc_n += pend ? 1u : 0u;
            acc_d += d;
        } else {
            flush();
            const int c = in ? chunk_of((int)v, a.cv0, a.nc) : -1;
            cc = chunk_of((int)vlast, a.cv0, a.nc);
            clo = a.cv0[cc];
            chi = cc + 1 < a.nc ? a.cv0[cc + 1] : a.n;
            if (pend) {
                if (c == cc) { acc_n += 1u; acc_d += d; }
                else {
                    atomicAdd(&a.cnt[2 * c], 1ull);
                    atomicAdd(&a.cnt[2 * c + 1], (unsigned long long)d);
                }
            }
        }
    }
    flush();
    if (a.hot) {
        uint64_t t = wave_sum_u64(hp);
        if (lane == 0 && t) atomicAdd(a.hotpend, (unsigned long long)t);
    }
    s_col[wid][lane] = colcnt;
    __syncthreads();
    if (wid == 0 && lane < B.K) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; w++) t += s_col[w][lane];
        if (t) atomicAdd(&B.colpend[lane], t);
    }
}

// Push of one vertex by one wave (lanes = columns).  Returns true when it pushed.
__device__ __forceinline__ void bpush_vertex(int v, uint64_t ptr, uint64_t deg, const PushArgs& a,
                                             const BatchArgs& B, BPiece* pieces, uint64_t* st) {
    const int lane = lane_id();
    const bool col_on = lane < B.K && ((B.active >> lane) & 1ull);
    float* rp = B.r + (long long)v * B.K + lane;
    float p = 0.f, d = 0.f;
    if (col_on && fabsf(*rp) > B.eps[lane]) {
        d = atomic_exch_f32(rp, 0.f);
        B.x[(long long)v * B.K + lane] += d;
        if (d != 0.f && deg) p = B.damp * d / (float)deg;
    }
    const uint64_t pm = __ballot(p != 0.f);
    const uint32_t npush = (uint32_t)__popcll(pm);
    // a vertex that folded any column counts as pushed (offset pair + its x and r entries)
    const uint64_t fm = __ballot(d != 0.f);
    if (lane == 0 && fm) { st[0] += 1; st[1] += (uint32_t)__popcll(fm); }
    if (!pm || deg == 0) return;
    if (deg > BHUB) {
        const uint32_t np = (uint32_t)((deg + BPIECE - 1) / BPIECE);
        uint32_t base = 0, row = 0;
        if (lane == 0) {
            base = atomicAdd(a.piece_count, np);
            row = atomicAdd(B.prow_count, 1u);
        }
        base = __shfl(base, 0, WAVE);
        row = __shfl(row, 0, WAVE);
        if (base + np <= a.piece_cap && row < B.prow_cap) {
            if (lane < B.K) B.prows[(long long)row * B.K + lane] = p;
            for (uint32_t j = lane; j < np; j += WAVE) {
                BPiece pc;
                pc.start = ptr + (uint64_t)j * BPIECE;
                const uint64_t rem = deg - (uint64_t)j * BPIECE;
                pc.len = (uint32_t)(rem < BPIECE ? rem : BPIECE);
                pc.prow = row;
                pieces[base + j] = pc;
            }
        } else if (lane == 0) {
            atomicAdd(a.overflow, 1ull);
        }
        if (lane == 0) { st[4] += deg; st[5] += deg * npush; }
        return;
    }
    if (lane == 0) { st[2] += deg; st[3] += deg * npush; }
    for (uint64_t e0 = 0; e0 < deg; e0 += WAVE) {
        const int id = e0 + lane < deg ? a.edges[ptr + e0 + lane] : 0;
        const int cnt = (int)(deg - e0 < (uint64_t)WAVE ? deg - e0 : WAVE);
        if (B.touch && e0 + lane < deg) B.touch[id] = 1;
        for (int t = 0; t < cnt; t++) {
            const int u = __shfl(id, t, WAVE);
            if (p != 0.f) atomic_add_f32(&B.r[(long long)u * B.K + lane], p);
        }
    }
}

// Batched push.  RANGE / BITMAP: each wave takes 64 consecutive vertices of [v0, v1)
// (all of them with a live check, or only the snapshot's pending ones); LIST: gathered
// lists; HOT: the compacted hot list (count on the device).
template <int SRC>
__global__ __launch_bounds__(BLOCK) void k_bpush(PushArgs a, BatchArgs B) {
    const int lane = lane_id();
    const long long gwave = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE;
    uint64_t st[6] = {0, 0, 0, 0, 0, 0};   // meaningful in lane 0
    BPiece* pieces = reinterpret_cast<BPiece*>(a.pieces);
    if (blockIdx.x == 0 && threadIdx.x == 0) clear_counters(a.piece_zero);
    if constexpr (SRC == SRC_RANGE || SRC == SRC_BITMAP) {
        const long long vb = ((long long)a.v0 & ~63ll) + gwave * WAVE;
        if (vb < a.v1) {
            uint64_t mask;
            if constexpr (SRC == SRC_BITMAP) {
                const long long w = vb >> 5;
                mask = (uint64_t)a.bitmap[w];
                if (vb + 32 < a.v1) mask |= (uint64_t)a.bitmap[w + 1] << 32;
            } else {
                mask = ~0ull;
            }
            if (vb < a.v0) mask &= ~0ull << (a.v0 - vb);
            if (vb + WAVE > a.v1) mask &= (1ull << (a.v1 - vb)) - 1ull;
            while (mask) {
                const int t = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const long long v = vb + t;
                if constexpr (SRC == SRC_RANGE)
                    if (a.hot && ((a.hot[v >> 5] >> (v & 31)) & 1u)) continue;
                const uint64_t b = a.idx[v], d = a.idx[v + 1] - b;
                bpush_vertex((int)v, b - a.ebase, d, a, B, pieces, st);
            }
        }
    } else if constexpr (SRC == SRC_LIST) {
        for (long long i = gwave; i < a.k; i += (long long)gridDim.x * NWAVES) {
            const uint64_t b = a.off[i], d = a.off[i + 1] - b;
            bpush_vertex(a.vids[i], b, d, a, B, pieces, st);
        }
    } else {
        const long long k = *a.hcount;
        for (long long e = gwave; e < k; e += (long long)gridDim.x * NWAVES) {
            const int i = a.hlist[e];
            const uint64_t b = a.hoff[i], d = a.hoff[i + 1] - b;
            bpush_vertex(a.hids[i], b, d, a, B, pieces, st);
        }
    }
    if (lane == 0) {
        if (st[0]) { atomicAdd(&a.counters[C_VERT], (unsigned long long)st[0]); atomicAdd(&a.counters[C_VERT_NM], (unsigned long long)st[1]); }
        if (st[2]) { atomicAdd(&a.counters[C_EDGES], (unsigned long long)st[2]); atomicAdd(&a.counters[C_EDGES_NM], (unsigned long long)st[3]); }
        if (st[4]) { atomicAdd(&a.counters[C_PEDGES], (unsigned long long)st[4]); atomicAdd(&a.counters[C_PEDGES_NM], (unsigned long long)st[5]); }
    }
}

__global__ __launch_bounds__(BLOCK) void k_bpieces(const BPiece* __restrict__ pcs, const uint32_t* count,
                                                    uint32_t cap, const int32_t* __restrict__ edges, BatchArgs B) {
    uint32_t cnt = *count;
    if (cnt > cap) cnt = cap;
    const int lane = lane_id();
    const long long nw = (long long)gridDim.x * NWAVES;
    for (long long w = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE; w < cnt; w += nw) {
        const BPiece pc = pcs[w];
        const float p = lane < B.K ? B.prows[(long long)pc.prow * B.K + lane] : 0.f;
        for (uint32_t e0 = 0; e0 < pc.len; e0 += WAVE) {
            const int id = e0 + lane < pc.len ? edges[pc.start + e0 + lane] : 0;
            const int c = (int)(pc.len - e0 < (uint32_t)WAVE ? pc.len - e0 : WAVE);
            if (B.touch && e0 + lane < pc.len) B.touch[id] = 1;
            for (int t = 0; t < c; t++) {
                const int u = __shfl(id, t, WAVE);
                if (p != 0.f) atomic_add_f32(&B.r[(long long)u * B.K + lane], p);
            }
        }
    }
}

// Hot compaction for a [n][K] state: hot entries in [lo, hi) pending in an active column.
__global__ __launch_bounds__(BLOCK) void k_bcompact_hot(int lo, int hi, const int32_t* __restrict__ hids, BatchArgs B,
                                                         int32_t* out, uint32_t* count) {
    const int lane = lane_id();
    const long long w0 = lo + ((blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE) * WAVE;
    const bool col_on = lane < B.K && ((B.active >> lane) & 1ull);
    for (int i = 0; i < WAVE; i++) {
        const long long e = w0 + i;
        if (e >= hi) break;
        const int v = hids[e];
        const bool pj = col_on && fabsf(B.r[(long long)v * B.K + lane]) > B.eps[lane];
        if (__ballot(pj) && lane == 0) out[atomicAdd(count, 1u)] = (int32_t)e;
    }
}

// Sets r[seeds[i]][col] += value (teleport mass of one column).
__global__ void k_bseed(int cnt, const int32_t* __restrict__ seeds, float* r, int K, int col, float value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) atomic_add_f32(&r[(long long)seeds[i] * K + col], value);
}

// ---- batched dense passes: fold + pull over the transpose, [n][K] rows, no atomics
// (except the hub pieces).  The snapshot bitmap of the pass marks the sources whose p row
// is written this pass; every other row is stale and skipped.

// Batched dense-pass weight rows are stored as PT: float, or bf16 (uint16_t: the high half
// of the float's bits, rounded to nearest even), which makes a 64-column row one 128-byte
// line.  A bf16 weight w is exact for the source's effective delta d' = w deg / damp; the fold
// moves d' into x and leaves the rounding remainder d - d' (at most 2^-9 |d|) in r, where it
// is pushed again once it exceeds eps (error feedback), so x + r = b + W x holds as with
// float rows.
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// bf16 rows have an even stride (K rounded up) so that a row is whole 4-byte words: the
// gather of a bf16 row is one word (two columns) per lane over 32 lanes, two rows per wave
// instruction (BAcc<uint16_t>).
__host__ __device__ constexpr int prow_stride(int K, bool bf16) { return bf16 ? K + (K & 1) : K; }
// Folds delta d of a vertex with deg out-edges: stores its weight at p[i] (deg > 0) and
// returns the part of d it accounts for (d itself for float rows and for deg == 0).
__device__ __forceinline__ float prow_fold(float* p, long long i, float d, uint64_t deg, float damp) {
    if (deg) p[i] = damp * d / (float)deg;
    return d;
}
__device__ __forceinline__ float prow_fold(uint16_t* p, long long i, float d, uint64_t deg, float damp) {
    if (!deg) return d;
    const uint16_t h = f32_to_bf16(damp * d / (float)deg);
    p[i] = h;
    return bf16_to_f32(h) * (float)deg / damp;
}
__device__ __forceinline__ void prow_zero(float* p, long long i) { p[i] = 0.f; }
__device__ __forceinline__ void prow_zero(uint16_t* p, long long i) { p[i] = 0; }

template <typename PT>
struct BDenseArgs {
    int n;
    const uint64_t* idx;          // out-offsets (fold: degrees)
    const uint32_t* pending;      // the pass's snapshot bitmap
    PT* p;                        // [n][pk] per-edge weights of the pass (columns 0..K-1), the
                                  // row of vertex v at rank[v] (decreasing out-degree order)
    int pk;                       // row stride of p: prow_stride(K, PT is bf16)
    const int32_t* rank;          // p row of each vertex
    const int32_t* inv;           // vertex of each p row (rows below nsrc)
    uint32_t* live;               // bit q: row q of p is live this pass (k_blive)
    int nsrc;                     // rows of vertices with out-edges
    const uint64_t* tidx;         // transpose (in-lists of at most PULL_HUB entries)
    const int32_t* tadj;          // in-neighbours (vertex ids)
    unsigned long long* counters; // C_FIELDS row of the dense pass
    unsigned long long* live_count;   // k_bpull_hub: rows read
    // k_bpull: the next pass's snapshot and fold, fused into the row updates
    // (p_next == nullptr: not fused)
    PT* p_next;                   // [n][pk] rows of the next pass
    uint32_t* bitmap_next;        // its snapshot bitmap
    const uint32_t* hubbits;      // destinations whose in-lists went to k_bpull_hub
    const int* cv0;
    int nc;
    unsigned long long* cnt;      // [2 nc] per-chunk pending vertices and degree sums
    unsigned long long* spread;   // C_SPREAD area of the counters
    unsigned long long* tile_next;   // k_bpull: next unclaimed tile (zeroed before the launch)
};

__device__ __forceinline__ bool bit_of(const uint32_t* __restrict__ bm, int v) {
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Fold: each wave takes 64-vertex segments and folds the snapshot's pending vertices one
// at a time, lanes = columns: x += r, r = 0, p = d r / deg for every active column over
// eps, the whole p row written (zeros for the other columns).
template <typename PT>
__global__ __launch_bounds__(BLOCK) void k_bfold(BDenseArgs<PT> a, BatchArgs B) {
    const int lane = lane_id();
    const long long nw = (long long)gridDim.x * NWAVES;
    const long long words = ((long long)a.n + 31) / 32;
    const bool col_on = lane < B.K && ((B.active >> lane) & 1ull);
    uint64_t st[4] = {0, 0, 0, 0};
    for (long long seg = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE; seg * WAVE < a.n; seg += nw) {
        uint64_t mask = a.pending[2 * seg];
        if (2 * seg + 1 < words) mask |= (uint64_t)a.pending[2 * seg + 1] << 32;
        while (mask) {
            const int t = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            const long long v = seg * WAVE + t;
            const uint64_t deg = a.idx[v + 1] - a.idx[v];
            float d = 0.f;
            bool pw = false;   // a weight row entry of this column is live
            if (col_on) {      // the pull reads active columns only
                float* rp = B.r + v * B.K + lane;
                const float rv = *rp;
                if (fabsf(rv) > B.eps[lane]) {
                    d = rv;
                    const float moved = prow_fold(a.p, (long long)a.rank[v] * a.pk + lane, rv, deg, B.damp);
                    *rp = rv - moved;
                    B.x[v * B.K + lane] += moved;
                    pw = deg != 0;
                } else if (deg) {
                    prow_zero(a.p, (long long)a.rank[v] * a.pk + lane);
                }
            }
            const uint64_t fm = __ballot(d != 0.f), pm = __ballot(pw);
            if (lane == 0) {
                if (fm) { st[0] += 1; st[1] += (uint64_t)__popcll(fm); }
                if (pm) { st[2] += deg; st[3] += deg * (uint64_t)__popcll(pm); }
            }
        }
    }
    if (lane == 0) {
        if (st[0]) { atomicAdd(&a.counters[C_VERT], (unsigned long long)st[0]); atomicAdd(&a.counters[C_VERT_NM], (unsigned long long)st[1]); }
        if (st[2]) { atomicAdd(&a.counters[C_EDGES], (unsigned long long)st[2]); atomicAdd(&a.counters[C_EDGES_NM], (unsigned long long)st[3]); }
    }
}

// The pass's live rows in p-row order (the pull tests in-list entries, which are row
// numbers): bit q = the snapshot bit of inv[q].  One ballot per 64 rows.
template <typename PT>
__global__ __launch_bounds__(BLOCK) void k_blive(BDenseArgs<PT> a) {
    const int lane = lane_id();
    const long long waves = (long long)gridDim.x * NWAVES;
    for (long long w = blockIdx.x * (long long)NWAVES + threadIdx.x / WAVE; w * WAVE < a.nsrc; w += waves) {
        const long long q = w * WAVE + lane;
        const uint64_t b = __ballot(q < a.nsrc && bit_of(a.pending, a.inv[q]));
        if (lane < 2) a.live[2 * w + lane] = (uint32_t)(b >> (32 * lane));
    }
}

// Per-owner sums of gathered rows.  BAcc<float>: lane = column.  BAcc<uint16_t>: the two
// halves of the wave hold different rows' partial sums, lane j of a half columns 2j and
// 2j + 1 (lo, hi); cols() adds the halves and hands column c to lane c (four shuffles, once
// per owner).
template <typename PT> struct BAcc;
template <> struct BAcc<float> {
    float v = 0.f;
    __device__ void clear() { v = 0.f; }
    __device__ float cols() const { return v; }
};
template <> struct BAcc<uint16_t> {
    float lo = 0.f, hi = 0.f;
    __device__ void clear() { lo = 0.f; hi = 0.f; }
    __device__ float cols() const {
        const int lane = lane_id();
        const float l = lo + __shfl_xor(lo, 32, WAVE), h = hi + __shfl_xor(hi, 32, WAVE);
        const float a = __shfl(l, lane >> 1, WAVE), b = __shfl(h, lane >> 1, WAVE);
        return (lane & 1) ? b : a;
    }
};

// Sum of the live p rows of in-list entries [b, e) into acc, in slot order; at every owner
// change `flush(owner, sums)` is called with lane = column.  owners(e0) gives, lane-parallel,
// the owner of slot e0 + lane; per row the wave reads that and the row's id with readlane
// (scalar), so the row's address is scalar and a row costs a few instructions besides its
// load.  Float rows: one 4K-byte row per wave instruction (lane = column), Q rows loaded
// before they are added, so a wave keeps Q gathers in flight.  Every load lands in its own
// register and is used only after the batch is issued (a use inside the load's branch makes
// the compiler wait for each load in turn).
template <int Q, typename Owners, typename Start, typename Flush>
__device__ __forceinline__ uint32_t bsum_rows(const BDenseArgs<float>& a, const BatchArgs& B, uint64_t b, uint64_t e,
                                              Owners owners, Start start, Flush flush, int& cur, BAcc<float>& acc) {
    const int lane = lane_id();
    const bool col = lane < B.K && ((B.active >> lane) & 1ull);
    uint32_t taken = 0;
    // windows of 64 entries; the next window's ids and the live-bitmap words of the window
    // after the current one are loaded while the current window's rows are summed
    int v = b + lane < e ? a.tadj[b + lane] : -1;
    uint32_t w = v >= 0 ? a.live[v >> 5] : 0u;
    int vn = b + WAVE + lane < e ? a.tadj[b + WAVE + lane] : -1;
    for (uint64_t e0 = b; e0 < e; e0 += WAVE) {
        uint64_t lm = __ballot(v >= 0 && ((w >> (v & 31)) & 1u));
        const uint32_t wn = vn >= 0 ? a.live[vn >> 5] : 0u;
        const uint64_t e2 = e0 + 2 * WAVE;
        const int vnn = e2 + lane < e ? a.tadj[e2 + lane] : -1;
        const int ownv = lm ? owners(e0) : 0;
        while (lm) {
            int slot[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                slot[q] = lm ? __ffsll((long long)lm) - 1 : -1;
                if (lm) lm &= lm - 1;
            }
            float val[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                val[q] = 0.f;
                if (slot[q] >= 0) {
                    const int vv = __builtin_amdgcn_readlane(v, slot[q]);
                    const float* row = a.p + (long long)vv * a.pk;
                    if (col) val[q] = row[lane];
                }
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (slot[q] < 0) break;
                const int own = __builtin_amdgcn_readlane(ownv, slot[q]);
                if (own != cur) {
                    if (cur >= 0) flush(cur, acc.cols());
                    acc.clear();
                    cur = own;
                    start(own);
                }
                acc.v += val[q];
                taken++;
            }
        }
        v = vn;
        w = wn;
        vn = vnn;
    }
    return taken;
}

// bf16 rows: a wave instruction gathers two rows, lanes 0-31 the first (one 4-byte word =
// two columns per lane) and lanes 32-63 the second, so it has the shape of a float-row
// gather (64 words, two lines) and carries twice the rows.  Q rows in flight = Q / 2
// instructions; each half adds its rows to its own partial sums.
template <int Q, typename Owners, typename Start, typename Flush>
__device__ __forceinline__ uint32_t bsum_rows(const BDenseArgs<uint16_t>& a, const BatchArgs& B, uint64_t b,
                                              uint64_t e, Owners owners, Start start, Flush flush, int& cur,
                                              BAcc<uint16_t>& acc) {
    const int lane = lane_id(), half = lane >> 5, j = lane & 31;
    // halves of inactive columns (and of the padding column of an odd K) hold stale weights
    const bool clo = 2 * j < B.K && ((B.active >> (2 * j)) & 1ull);
    const bool chi = 2 * j + 1 < B.K && ((B.active >> (2 * j + 1)) & 1ull);
    const uint32_t wmask = (clo ? 0x0000ffffu : 0u) | (chi ? 0xffff0000u : 0u);
    const bool col = wmask != 0;
    uint32_t taken = 0;
    int v = b + lane < e ? a.tadj[b + lane] : -1;
    uint32_t w = v >= 0 ? a.live[v >> 5] : 0u;
    int vn = b + WAVE + lane < e ? a.tadj[b + WAVE + lane] : -1;
    for (uint64_t e0 = b; e0 < e; e0 += WAVE) {
        uint64_t lm = __ballot(v >= 0 && ((w >> (v & 31)) & 1u));
        const uint32_t wn = vn >= 0 ? a.live[vn >> 5] : 0u;
        const uint64_t e2 = e0 + 2 * WAVE;
        const int vnn = e2 + lane < e ? a.tadj[e2 + lane] : -1;
        const int ownv = lm ? owners(e0) : 0;
        while (lm) {
            int slot[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                slot[q] = lm ? __ffsll((long long)lm) - 1 : -1;
                if (lm) lm &= lm - 1;
            }
            uint32_t raw[Q / 2];
#pragma unroll
            for (int q = 0; q < Q / 2; q++) {
                raw[q] = 0;
                const int s0 = slot[2 * q], s1 = slot[2 * q + 1];
                const int v0 = s0 >= 0 ? __builtin_amdgcn_readlane(v, s0) : 0;
                const int v1 = s1 >= 0 ? __builtin_amdgcn_readlane(v, s1) : 0;
                const int sl = half ? s1 : s0;
                const int vv = half ? v1 : v0;
                if (sl >= 0 && col)
                    raw[q] = *reinterpret_cast<const uint32_t*>(a.p + (long long)vv * a.pk + 2 * j);
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (slot[q] < 0) break;
                const int own = __builtin_amdgcn_readlane(ownv, slot[q]);
                if (own != cur) {
                    if (cur >= 0) flush(cur, acc.cols());
                    acc.clear();
                    cur = own;
                    start(own);
                }
                if (half == (q & 1)) {
                    const uint32_t wv = raw[q >> 1] & wmask;
                    acc.lo += __uint_as_float(wv << 16);
                    acc.hi += __uint_as_float(wv & 0xffff0000u);
                }
                taken++;
            }
        }
        v = vn;
        w = wn;
        vn = vnn;
    }
    return taken;
}

// Pull: a wave owns 64 consecutive destinations, whose in-lists are one contiguous run of
// the transpose; it walks the run in slot order, reads the live rows (256 bytes each,
// lanes = columns) and adds each destination's sum to its r row once, with a plain
// read-modify-write (no other wave touches these rows).  The owner of a slot is the last
// of the 64 list starts at or before it (one ballot).
constexpr long long BPULL_CLAIM = 4;   // tiles of 64 destinations a wave claims at a time

template <int Q, typename PT>
__global__ __launch_bounds__(BLOCK, 6) void k_bpull(BDenseArgs<PT> a, BatchArgs B) {
    __shared__ uint32_t s_col[NWAVES][WAVE];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const long long gw = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE;
    const long long ntile = ((long long)a.n + WAVE - 1) / WAVE;
    const long long words = ((long long)a.n + 31) / 32;
    const bool fused = a.p_next != nullptr;
    const bool col_on = lane < B.K && ((B.active >> lane) & 1ull);
    const float eps = lane < B.K ? B.eps[lane] : 0.f;
    // waves claim runs of BPULL_CLAIM consecutive tiles from a counter (in-degree sums vary
    // widely between tiles); within a run the per-chunk counts of the next snapshot
    // accumulate in registers

    uint32_t colcnt = 0;                  // pending entries of this lane's column (next pass)
    uint64_t fs[4] = {0, 0, 0, 0};        // fused fold statistics (wave-uniform)
    int cc = -1;
    uint64_t acc_n = 0, acc_d = 0;
    auto flush_counts = [&]() {
        if (cc >= 0 && lane == 0 && acc_n) {
            atomicAdd(&a.cnt[2 * cc], (unsigned long long)acc_n);
            atomicAdd(&a.cnt[2 * cc + 1], (unsigned long long)acc_d);
        }
        acc_n = 0;
        acc_d = 0;
    };
    for (;;) {
    unsigned long long claim = 0;
    if (lane == 0) claim = atomicAdd(a.tile_next, (unsigned long long)BPULL_CLAIM);
    const long long t0 = (long long)__shfl(claim, 0, WAVE);
    if (t0 >= ntile) break;
    const long long t1 = t0 + BPULL_CLAIM < ntile ? t0 + BPULL_CLAIM : ntile;
    for (long long t = t0; t < t1; t++) {
        const long long u0 = t * WAVE;
        const int nt = (int)((long long)a.n - u0 < WAVE ? (long long)a.n - u0 : WAVE);
        const uint64_t lo = a.tidx[u0], hi = a.tidx[u0 + nt];
        const uint64_t my = lane < nt ? a.tidx[u0 + lane] : hi;   // lanes past the tile never own a slot
        uint64_t tmask = 0;   // destinations of the tile pending in the next pass
        uint64_t vis = 0;     // destinations of the tile whose row fold_row has tested
        // lane d: destination d's out-degree and p row, loaded with the tile's list starts so
        // that a fold waits only for its r and x rows (issued together)
        const uint64_t degv = fused && lane < nt ? a.idx[u0 + lane + 1] - a.idx[u0 + lane] : 0;
        const int rankv = fused && lane < nt ? a.rank[u0 + lane] : 0;
        // The next pass's fold of destination u (k_bfold's rule on the row just completed):
        // s = this pass's sum for u (0 for a hub, whose row k_bpull_hub completed).
        // r0, x0: u's r and x rows (lane = column; read by the caller)
        auto fold_row = [&](int own, float s, float r0, float x0) {
            const long long u = u0 + own;
            vis |= 1ull << own;
            float rv = r0 + s;
            const float xv = x0;
            const bool pend = col_on && fabsf(rv) > eps;
            const uint64_t pm = __ballot(pend);
            if (!pm) {
                if (col_on && s != 0.f) B.r[u * B.K + lane] = rv;
                return;
            }
            const uint64_t deg = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)degv, own) |
                                 (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(degv >> 32), own) << 32;
            const long long prow = (long long)__builtin_amdgcn_readlane(rankv, own) * a.pk + lane;
            bool pw = false;
            if (pend) {
                const float moved = prow_fold(a.p_next, prow, rv, deg, B.damp);
                B.x[u * B.K + lane] = xv + moved;
                rv -= moved;
                pw = deg != 0;
            } else if (col_on && deg) {
                prow_zero(a.p_next, prow);
            }
            if (col_on) B.r[u * B.K + lane] = rv;
            tmask |= 1ull << own;
            colcnt += pend ? 1u : 0u;
            const uint64_t vm = __ballot(pw);
            fs[0] += 1;
            fs[1] += (uint64_t)__popcll(pm);
            if (vm) { fs[2] += deg; fs[3] += deg * (uint64_t)__popcll(vm); }
        };
        int cur = -1;
        BAcc<PT> acc;
        // owner of slot e0 + lane: the destinations whose lists start at or before e0, plus
        // one per list that starts inside the window at or before the lane's slot
        auto owners = [&](uint64_t e0) {
            int ow = __popcll(__ballot(my <= e0)) - 1;
            uint64_t sm = __ballot(my > e0 && my < e0 + WAVE);
            while (sm) {
                const int d = __ffsll((long long)sm) - 1;
                sm &= sm - 1;
                const int off = __builtin_amdgcn_readlane((int)(my - e0), d);
                ow += lane >= off ? 1 : 0;
            }
            return ow;
        };
        // fused: the r and x rows of a destination are loaded when its first row is added
        // (start), so they have arrived when its sum is complete (flush)
        float rpre = 0.f, xpre = 0.f;
        auto start = [&](int own) {
            if (fused && col_on) {
                rpre = B.r[(u0 + own) * B.K + lane];
                xpre = B.x[(u0 + own) * B.K + lane];
            }
        };
        auto flush = [&](int own, float sum) {
            if (fused) {
                if (__ballot(sum != 0.f)) fold_row(own, sum, rpre, xpre);
            } else if (lane < B.K && sum != 0.f) {
                B.r[(u0 + own) * B.K + lane] += sum;
            }
        };
        bsum_rows<Q>(a, B, lo, hi, owners, start, flush, cur, acc);
        if (cur >= 0) flush(cur, acc.cols());
        if (!fused) continue;
        // hub destinations of the tile, then the tile's two bitmap words and counts; a
        // destination that received nothing keeps its row: below eps if it was not folded
        // before this pass; with bf16 weights a vertex folded before this pass may hold a
        // rounding remainder above eps, so those are tested too
        uint64_t hm = __ballot(lane < nt && bit_of(a.hubbits, (int)(u0 + lane)));
        if (sizeof(PT) == 2) hm |= __ballot(lane < nt && bit_of(a.pending, (int)(u0 + lane))) & ~vis;
        while (hm) {
            const int i = __ffsll((long long)hm) - 1;
            hm &= hm - 1;
            const long long u = u0 + i;
            fold_row(i, 0.f, col_on ? B.r[u * B.K + lane] : 0.f, col_on ? B.x[u * B.K + lane] : 0.f);
        }
        if (lane == 0) {
            a.bitmap_next[2 * t] = (uint32_t)tmask;
            if (2 * t + 1 < words) a.bitmap_next[2 * t + 1] = (uint32_t)(tmask >> 32);
        }
        if (!tmask) continue;
        const bool pv = (tmask >> lane) & 1ull;
        const uint64_t dv = pv ? degv : 0;
        const int c0 = chunk_of((int)u0, a.cv0, a.nc), c1 = chunk_of((int)(u0 + nt - 1), a.cv0, a.nc);
        if (c0 == c1) {
            if (c0 != cc) { flush_counts(); cc = c0; }
            acc_n += (uint64_t)__popcll(tmask);
            acc_d += wave_sum_u64(dv);
        } else if (pv) {
            const int c = chunk_of((int)(u0 + lane), a.cv0, a.nc);
            atomicAdd(&a.cnt[2 * c], 1ull);
            atomicAdd(&a.cnt[2 * c + 1], (unsigned long long)dv);
        }
    }
    }
    if (!fused) return;
    flush_counts();
    s_col[wid][lane] = colcnt;
    __syncthreads();
    if (wid == 0 && lane < B.K) {
        uint32_t tt = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; w++) tt += s_col[w][lane];
        if (tt) atomicAdd(&B.colpend[lane], tt);
    }
    if (lane == 0 && fs[0]) {
        unsigned long long* slot = a.spread + (gw % SPREAD_SLOTS) * SPREAD_STRIDE;
#pragma unroll
        for (int i = 0; i < 4; i++) if (fs[i]) atomicAdd(&slot[S_VERT + i], (unsigned long long)fs[i]);
    }
}
template <int Q, typename PT>
__global__ __launch_bounds__(BLOCK) void k_bpull_hub(BDenseArgs<PT> a, BatchArgs B, const PullPiece* __restrict__ pieces,
                                                      int npieces) {
    const int lane = lane_id();
    const long long nw = (long long)gridDim.x * NWAVES;
    uint32_t taken = 0;
    for (long long h = (blockIdx.x * (long long)BLOCK + threadIdx.x) / WAVE; h < npieces; h += nw) {
        const PullPiece pc = pieces[h];
        int cur = -1;
        BAcc<PT> acc;
        auto owners = [](uint64_t) { return 0; };
        auto start = [](int) {};
        auto flush = [](int, float) {};
        taken += bsum_rows<Q>(a, B, pc.start, pc.start + pc.len, owners, start, flush, cur, acc);
        const float sum = acc.cols();
        if (lane < B.K && sum != 0.f) atomic_add_f32(&B.r[(long long)pc.u * B.K + lane], sum);
    }
    if (lane == 0 && taken) atomicAdd(a.live_count, (unsigned long long)taken);
}

// ------------------------------------------------------------------ host-side context

struct PushTune {
    int best = -1;               // candidate kept, -1 while undecided
    int next = 0;                // next candidate to time
    int misses = 0;              // trials that pushed too few edges to time
    double ns_edge[3] = {0, 0, 0};
};
struct PushTrial {
    bool on = false;
    int cls = 0, cand = 0;
    unsigned long long before = 0;
};

struct BurstGraph {
    int passes = 0, phase = 0, slot = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double snap_bytes = 0;   // the snapshot bytes its launches count, added per replay
};

struct PieceBuf {
    Piece* d = nullptr;          // pieces (Piece, or BPiece for batched contexts)
    uint32_t* cnt = nullptr;     // phase p: piece count at cnt[2p*CNT_STRIDE], p-row count one line later
    uint32_t cap = 0;
    int phase = 0;
};

// A staging slot: pinned host buffers, their device copies, and two events on the slot's
// stream -- `copied` after the host-to-device copies of its last chunk (the host buffers are
// free again: the next chunk can be read into them while this one is still being pushed; the
// next copy into the device buffers is ordered behind that push on the same stream), `done`
// after its push.  Zero-copy slots are read in place by the push, so they free on `done`.
struct Slot {
    int32_t* h_edges = nullptr; int32_t* d_edges = nullptr;
    int32_t* h_vids = nullptr;  int32_t* d_vids = nullptr;
    uint64_t* h_off = nullptr;  uint64_t* d_off = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t copied = nullptr;
    hipEvent_t frees = nullptr;   // the event exc_slot_acquire waits for
    bool busy = false;
};

struct Timer {
    hipEvent_t a, b;
    int cls;
};

// T_PULL: the dense pull (k_pull / k_pull_sorted / k_bpull) over every measure of the context;
// T_PULL_A1, T_PULL_A2: dense pulls over its first one / two measures (exc_set_active_measures)
enum { T_SNAP = 0, T_PUSH, T_PUSH_LAST = T_PUSH + 3, T_PIECE, T_OTHER, T_FOLD, T_PULL, T_PULL_HUB, T_PULL_A1, T_PULL_A2,
       T_NCLS };

struct Ctx {
    int dev = 0, n = 0, nc = 0, nm = 1, S = 1, push_kernel = EXC_KERNEL_LB;
    uint64_t m = 0;
    bool zc = false, timing = false, timing_pull_only = false;
    bool counts_enqueued = false;   // a dense pull queued its snapshot counts' copy to h_small
    bool counts_copied = false;     // ... and a pass sync has waited for it
    exc_measures M{};
    DevMeasures DM{};
    uint64_t* d_idx = nullptr;
    float* d_x = nullptr;
    float* d_r = nullptr;
    float* d_x0 = nullptr;   // checkpoint
    float* d_r0 = nullptr;
    uint32_t* d_bitmap = nullptr;
    unsigned long long* d_cnt = nullptr;       // [2nc] + hotpend
    int* d_cv0 = nullptr;
    uint16_t* d_trace = nullptr;
    unsigned long long* d_counters = nullptr;  // C_NCOUNTERS
    unsigned long long* h_small = nullptr;     // pinned: cnt + hotpend + residual slots
    // bursts: one count region (counts + residual slots, cnt_words) per pass, on the device
    // and pinned on the host, in two slots of EXC_BURST_MAX regions so that a second burst can
    // be submitted while the first is collected (slot = submission number & 1)
    unsigned long long* d_burst = nullptr;
    unsigned long long* h_burst = nullptr;
    int burst_subs = 0, burst_cols = 0;        // bursts submitted / collected
    int burst_len[2] = {0, 0};
    hipEvent_t burst_ev[2] = {nullptr, nullptr};   // after a slot's copy to the host
    bool in_burst = false;                     // its passes are being enqueued: the regions are
                                               // cleared and copied once for the whole burst
    std::vector<BurstGraph> burst_graphs;      // recorded push bursts (launch_push_burst_graph)
    PushTune ptune[7];                         // EXC_KERNEL_TUNE, by degree class
    hipEvent_t ptune_ev[2] = {nullptr, nullptr};
    uint64_t ptune_bound = 0;                  // edges the launch being made may push
    bool use_burst_graphs = true;              // EXC_BURST_GRAPH=0: push bursts launched directly
    uint64_t burst_graph_launches = 0;
    double l1_last[EXC_MAXM] = {};             // FP64 residual sums of the last snapshot
    float linf_last[EXC_MAXM] = {};            // and the largest |r| per measure
    int snap_nm = 0;                           // measures the last snapshot computed (the others
                                               // were dropped: their sums above stay as they were)
    // Dense passes update measures [0, nm_act) only: the others had nothing pending when they
    // were dropped (exc_set_active_measures), and a measure's residual changes only by its own
    // pushes, so they stay converged.  p_cols: columns of the rows in d_p (the width of the
    // fold that wrote them).
    int nm_act = 1;
    int p_cols = 0;
    // per-launch algorithmic bytes of the plain dense passes, by active measures - 1: the
    // parts known at launch (in-list entries, destinations, live bitmap, hub rows of passes
    // without the bitmap); the rows the kernels count are added in exc_get_stats
    double pull_fixed_bytes[3] = {}, hub_fixed_bytes[3] = {}, fold_fixed_bytes = 0;
    uint64_t pull_launches_w[3] = {};
    int pull_tile_w[3] = {};                   // the sorted tile each width ran with (0: k_pull)
    int sorted_w = 1;                          // widths up to this run the sorted pull (where its LDS fits)
    int pull_small = 2;                        // k_pull tiles of 64 destinations: 0 never, 1 always, 2 small graphs
    excent_dev::SortedTileBuffers sorted;      // source-sorted dense pull (k_pull_sorted)
    int sorted_T = 0;                          // its tile (destinations per block); 0: k_pull
    int32_t* d_hub_keys = nullptr;             // destinations privatised by k_pieces (HubTable)
    uint32_t* d_hub_filter = nullptr;
    int hub_probes = 0, hub_count = 0;
    bool l1_valid = false;
    bool fold_ahead_ok = true;                 // exc_fold_ahead: dense passes may fold the next pass
    uint64_t max_chunk_edges = 0;
    int max_chunk_verts = 0;
    int cus = 256;
    // caches
    int32_t* d_arena = nullptr;
    uint64_t arena_edges = 0;
    int nh = 0;
    int32_t* d_hids = nullptr; uint64_t* d_hoff = nullptr; int32_t* d_hedges = nullptr;
    uint32_t* d_hot = nullptr; int32_t* d_hlist = nullptr; uint32_t* d_hcount = nullptr;
    uint64_t hot_edges = 0;
    hipStream_t st[3] = {nullptr, nullptr, nullptr};   // 0 resident, 1-2 slots
    PieceBuf pb[3];
    Slot slot[2];
    std::vector<Timer> pending_timers;
    std::vector<hipEvent_t> free_events;
    double t_ms[T_NCLS] = {};
    uint64_t t_launch[T_NCLS] = {};
    double snap_bytes = 0;
    float* d_wts = nullptr;           // per-edge coefficients aligned with the arena (weighted graphs)
    float* d_twts = nullptr;          // the same coefficients in transpose order (dense passes)
    // dense passes: fold + pull over the transpose (whole graph resident)
    uint64_t* d_tidx = nullptr;
    int32_t* d_tadj = nullptr;
    float* d_p = nullptr;             // [n][S]
    int32_t* d_rank = nullptr;          // p row of each vertex (out-degree order)
    int32_t* d_inv = nullptr;           // batched: vertex of each p row
    unsigned long long* d_tile_next = nullptr;   // batched: k_bpull's tile counter
    uint32_t* d_pbits = nullptr;        // live-row bitmap of p, 2 * ceil(nsrc / 64) words
    int nsrc = 0;                       // vertices with out-edges (rows of p)
    int pull_live = 2;                  // live-row bitmap: 0 never, 1 always, 2 when demand <= frac * m
    double pull_live_frac = 0.8;          // k_pull: the live-row bitmap below this demand fraction
    double pull_live_frac_sorted = 0.99;  // k_pull_sorted
    bool snap_ready = false;            // the last dense pass computed the next snapshot
    bool folded = false;                // ... and the next pass's fold (p holds its rows): the
                                        // next pass must be a dense pass over all vertices
    int fuse_fold = 1;                  // fold the next pass inside k_pull when demand is high
    float* d_p2 = nullptr;              // the second p buffer (rows of the next pass)
    uint32_t* d_bitmap2 = nullptr;      // batched: the next pass's snapshot bitmap
    uint32_t* d_hubbits = nullptr;      // batched: destinations summed by k_bpull_hub
    int pull_seq = 8;                   // in-list entries per lane in k_pull (4, 8, 16 or 32)
    int bpull_q = 8;                    // batched pull: rows in flight per wave (8 or 16)
    int bhub_q = 0;                     // EXC_BHUB_Q: the hub pieces' rows in flight (0: by bpull_q)
    // batched dense passes: p rows in bf16 (2 bytes per column; EXC_BATCH_ROWS=f32: float)
    bool brows16 = true;
    // Batched pull tuner: the first two dense passes whose demand is at least half the edges
    // run with 8 and then 16 rows in flight and are timed with events; later passes use the
    // faster one.  The rows are added in slot order either way, so the choice changes the
    // time and not the sums.  EXC_BPULL_Q fixes the value and turns the tuner off.
    bool bq_tune = false;
    int bq_trial = 0;                   // next candidate (BQ_CAND index)
    bool bq_recorded = false;           // a trial pass's events are recorded and not yet read
    double bq_demand = 0;               // demand of that pass
    double bq_ns_edge[2] = {0, 0};      // each candidate's time per demand edge
    hipEvent_t bq_ev[2] = {nullptr, nullptr};
    uint64_t pull_passes_live = 0;
    PullPiece* d_pull_hubs = nullptr;   // pieces of in-lists longer than PULL_HUB
    int n_pull_hubs = 0;
    uint64_t pull_hub_edges = 0, pull_hub_vertices = 0;
    double pull_threshold = -1;       // < 0: off; else pull when demand >= threshold * m
    uint64_t last_demand = 0;
    uint64_t pull_passes = 0;
    // batched personalised PageRank (K > 0): x, r are [n][K]
    int K = 0;
    float damp = 0.85f;
    float* d_beps = nullptr;        // [K]
    uint32_t* d_colpend = nullptr;  // [K]
    uint64_t active = ~0ull;
    float* prows[3] = {nullptr, nullptr, nullptr};
    uint32_t prow_cap[3] = {0, 0, 0};
    // frontier snapshots (k_bsnapshot<true>): after a snapshot the touch bytes are clean and
    // front_ok holds until anything but a resident push changes the state; the next snapshot
    // may read the touched vertices only if the pushes since covered every chunk that had
    // pending vertices (front_pend, from the last counts read)
    uint8_t* d_touch = nullptr;     // [n]
    bool use_front = true;          // EXC_BATCH_FRONT=0: every snapshot reads every row
    bool front_ok = false;
    bool front_pend_known = false;
    std::vector<int> h_cv0;
    std::vector<uint64_t> front_pend;
    std::vector<char> front_cov;
    uint64_t front_snaps = 0;
};

BatchArgs batch_args(Ctx* c, int s) {
    BatchArgs B{};
    B.x = c->d_x;
    B.r = c->d_r;
    B.eps = c->d_beps;
    B.damp = c->damp;
    B.K = c->K;
    B.active = c->active;
    B.colpend = c->d_colpend;
    B.prows = s >= 0 ? c->prows[s] : nullptr;
    B.prow_count = s >= 0 ? c->pb[s].cnt + (2 * c->pb[s].phase + 1) * CNT_STRIDE : nullptr;
    B.prow_cap = s >= 0 ? c->prow_cap[s] : 0;
    B.touch = c->d_touch;
    return B;
}

// The counts buffer: per-chunk counts [2 nc], hot-pending, then (from a 128-byte line on)
// the FP64 residual slots; one clear covers all of it.
size_t l1_offset(const Ctx* c) { return ((2 * (size_t)c->nc + 1 + L1_STRIDE - 1) / L1_STRIDE) * L1_STRIDE; }
size_t cnt_words(const Ctx* c) { return l1_offset(c) + (size_t)L1_SLOTS * L1_STRIDE; }
double* l1_slots(const Ctx* c) { return reinterpret_cast<double*>(c->d_cnt + l1_offset(c)); }
void clear_counts(const Ctx* c, hipStream_t st) {
    if (c->in_burst) return;   // exc_run_passes cleared every pass's region at its start
    const int w = (int)cnt_words(c);
    hipLaunchKernelGGL(k_zero_u64, dim3((w + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, c->d_cnt, w);
}

int ensure_bpieces(Ctx* c, int s, uint64_t edges_bound) {
    uint64_t need64 = 2 * (edges_bound / BPIECE) + 64;
    uint64_t rows64 = edges_bound / BHUB + 64;
    if (need64 > 0xffffffffull) need64 = 0xffffffffull;
    if (rows64 > 0xffffffffull) rows64 = 0xffffffffull;
    const uint32_t need = (uint32_t)need64, rows = (uint32_t)rows64;
    PieceBuf& p = c->pb[s];
    if (p.d && p.cap >= need && c->prows[s] && c->prow_cap[s] >= rows) return EXC_OK;
    HIPCHK(hipStreamSynchronize(c->st[s]));
    if (p.d) { (void)hipFree(p.d); p.d = nullptr; p.cap = 0; }
    if (c->prows[s]) { (void)hipFree(c->prows[s]); c->prows[s] = nullptr; c->prow_cap[s] = 0; }
    if (hipMalloc(&p.d, (size_t)need * sizeof(BPiece)) != hipSuccess ||
        hipMalloc(&c->prows[s], (size_t)rows * c->K * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(EXC_ENOMEM, "cannot allocate %u batched hub pieces", need);
    }
    p.cap = need;
    c->prow_cap[s] = rows;
    if (!p.cnt) {   // cleared on the launch stream: the streams are non-blocking, so a clear on
                    // the null stream would not be ordered before the first push
        HIPCHK(hipMalloc(&p.cnt, 4 * CNT_STRIDE * sizeof(uint32_t)));
        HIPCHK(hipMemsetAsync(p.cnt, 0, 4 * CNT_STRIDE * sizeof(uint32_t), c->st[s]));
    }
    return EXC_OK;
}

hipEvent_t get_event(Ctx* c) {
    if (!c->free_events.empty()) { hipEvent_t e = c->free_events.back(); c->free_events.pop_back(); return e; }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct ScopedTimer {
    Ctx* c; hipStream_t s; int cls; hipEvent_t a = nullptr;
    ScopedTimer(Ctx* c_, hipStream_t s_, int cls_) : c(c_), s(s_), cls(cls_) {
        c->t_launch[cls]++;
        if (c->timing && (!c->timing_pull_only || cls == T_PULL || cls == T_PULL_A1 || cls == T_PULL_A2)) { a = get_event(c); if (a) hipEventRecord(a, s); }
    }
    ~ScopedTimer() {
        if (!a) return;
        hipEvent_t b = get_event(c);
        if (!b) return;
        hipEventRecord(b, s);
        c->pending_timers.push_back({a, b, cls});
    }
};

int collect_timers(Ctx* c) {
    for (auto& t : c->pending_timers) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        c->t_ms[t.cls] += ms;
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending_timers.clear();
    return EXC_OK;
}

void burst_graphs_clear(Ctx* c);

int ensure_pieces(Ctx* c, int s, uint64_t edges_bound) {
    if (c->K > 0) return ensure_bpieces(c, s, edges_bound);
    uint64_t need64 = edges_bound / HUB + edges_bound / PIECE + 64;
    if (need64 > 0xffffffffull) need64 = 0xffffffffull;
    uint32_t need = (uint32_t)need64;
    PieceBuf& p = c->pb[s];
    if (p.d && p.cap >= need) return EXC_OK;
    if (s == 0) burst_graphs_clear(c);   // recorded pushes name the old buffer
    if (p.d) { HIPCHK(hipStreamSynchronize(c->st[s])); hipFree(p.d); p.d = nullptr; }
    if (hipMalloc(&p.d, (size_t)need * sizeof(Piece)) != hipSuccess) {
        p.d = nullptr;
        return fail(EXC_ENOMEM, "cannot allocate %u hub pieces", need);
    }
    p.cap = need;
    if (!p.cnt) {   // cleared on the launch stream: the streams are non-blocking, so a clear on
                    // the null stream would not be ordered before the first push
        HIPCHK(hipMalloc(&p.cnt, 4 * CNT_STRIDE * sizeof(uint32_t)));
        HIPCHK(hipMemsetAsync(p.cnt, 0, 4 * CNT_STRIDE * sizeof(uint32_t), c->st[s]));
    }
    return EXC_OK;
}

template <int NM, int S, int SRC, int G>
void launch_push_t(Ctx* c, hipStream_t st, unsigned grid, const PushArgs& a) {
    hipLaunchKernelGGL((k_push<NM, S, SRC, G>), dim3(grid), dim3(BLOCK), 0, st, a, c->DM);
}

template <int NM, int S, int SRC>
int launch_push_g(Ctx* c, hipStream_t st, unsigned grid, const PushArgs& a, int G) {
    switch (G) {
        case 0: launch_push_t<NM, S, SRC, 0>(c, st, grid, a); break;
        case 1: launch_push_t<NM, S, SRC, 1>(c, st, grid, a); break;
        case 2: launch_push_t<NM, S, SRC, 2>(c, st, grid, a); break;
        case 4: launch_push_t<NM, S, SRC, 4>(c, st, grid, a); break;
        case 8: launch_push_t<NM, S, SRC, 8>(c, st, grid, a); break;
        case 16: launch_push_t<NM, S, SRC, 16>(c, st, grid, a); break;
        case 32: launch_push_t<NM, S, SRC, 32>(c, st, grid, a); break;
        default: launch_push_t<NM, S, SRC, 64>(c, st, grid, a); break;
    }
    return EXC_OK;
}

template <int NM, int S>
int launch_push_src(Ctx* c, hipStream_t st, unsigned grid, const PushArgs& a, int src, int G) {
    switch (src) {
        case SRC_RANGE: return launch_push_g<NM, S, SRC_RANGE>(c, st, grid, a, G);
        case SRC_BITMAP: return launch_push_g<NM, S, SRC_BITMAP>(c, st, grid, a, G);
        case SRC_LIST: return launch_push_g<NM, S, SRC_LIST>(c, st, grid, a, G);
        default: return launch_push_g<NM, S, SRC_HOT>(c, st, grid, a, G);
    }
}

template <int NM, int S>
void launch_pieces_t(Ctx* c, int s, const int32_t* edges, const float* wts) {
    const unsigned grid = (unsigned)(c->cus * 4);
    const HubTable ht{c->d_hub_keys, c->d_hub_filter, c->hub_probes, c->hub_count};
    hipLaunchKernelGGL((k_pieces<NM, S>), dim3(grid), dim3(BLOCK), 0, c->st[s], c->pb[s].d,
                       c->pb[s].cnt + 2 * c->pb[s].phase * CNT_STRIDE, c->pb[s].cap, edges, wts, c->DM, ht);
}

int norm_group(int g) {
    if (g <= 0) return 0;
    int p = 1;
    while (p < g && p < 64) p <<= 1;
    return p;
}

// Push autotuner (EXC_KERNEL_TUNE).  A launch's degree class is the log2 of its group hint
// (the host's pick_group of the pushed vertices' average degree).  Candidates of class c: the
// load-balanced kernel (G = 0), G = 2^c, G = max(1, 2^c / 4).  Until a class has timed all
// three, its launches that may push at least PTUNE_MIN_EDGES edges run the next candidate
// alone (the stream is drained first, the kernel is bracketed by events, and the non-hub
// edges it pushed are the change of its source's counter); then the class keeps the fastest
// per pushed edge.  The others, and the launches of a burst (which must not wait), use the
// class's choice once made, else the load-balanced kernel.
constexpr uint64_t PTUNE_MIN_EDGES = 1u << 16;
int ptune_candidate(int cls, int i) {
    const int g = 1 << cls;
    return i == 0 ? 0 : i == 1 ? g : std::max(1, g / 4);
}
int ptune_class(int group) {
    int c = 0;
    while (c < 6 && (1 << c) < group) c++;
    return c;
}

int push_tune_pick(Ctx* c, int s, int src, int group, int* G, PushTrial* trial) {
    const int cls = ptune_class(group > 0 ? group : 1);
    PushTune& t = c->ptune[cls];
    if (t.best >= 0) { *G = ptune_candidate(cls, t.best); return EXC_OK; }
    *G = 0;
    if (c->in_burst || group <= 0 || c->ptune_bound < PTUNE_MIN_EDGES) return EXC_OK;
    if (!c->ptune_ev[0]) {
        HIPCHK(hipEventCreate(&c->ptune_ev[0]));
        HIPCHK(hipEventCreate(&c->ptune_ev[1]));
    }
    HIPCHK(hipStreamSynchronize(c->st[s]));   // nothing else on this stream during the trial
    HIPCHK(hipMemcpy(&trial->before, c->d_counters + src * C_FIELDS + C_EDGES, sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    trial->on = true;
    trial->cls = cls;
    trial->cand = t.next;
    *G = ptune_candidate(cls, t.next);
    return EXC_OK;
}

int push_tune_collect(Ctx* c, int src, const PushTrial& trial) {
    HIPCHK(hipEventSynchronize(c->ptune_ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ptune_ev[0], c->ptune_ev[1]));
    unsigned long long after = 0;
    HIPCHK(hipMemcpy(&after, c->d_counters + src * C_FIELDS + C_EDGES, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    const unsigned long long edges = after - trial.before;
    PushTune& t = c->ptune[trial.cls];
    if (edges >= PTUNE_MIN_EDGES) {
        t.ns_edge[trial.cand] = (double)ms * 1e6 / (double)edges;
        t.next++;
    } else if (++t.misses == 8) {
        t.next = 3;   // this class's launches push too little to time: decide on what was timed
    }
    if (t.next == 3) {   // the fastest timed candidate, the load-balanced kernel if none was
        int b = 0;
        for (int i = 1; i < 3; i++)
            if (t.ns_edge[i] > 0 && (t.ns_edge[b] == 0 || t.ns_edge[i] < t.ns_edge[b])) b = i;
        t.best = b;
    }
    return EXC_OK;
}

// One push launch + its hub-piece launch on stream index s.
int push_launch(Ctx* c, int s, int src, unsigned grid, PushArgs a, int group, uint64_t edges_bound) {
    if (grid == 0) return EXC_OK;
    if (c->K == 0) {   // plain kernels are persistent over tiles: grid = tiles, capped
        a.ntiles = grid;
        const unsigned cap = (unsigned)(c->cus * 8);
        if (grid > cap) grid = cap;
    }
    int rc = ensure_pieces(c, s, edges_bound);
    if (rc) return rc;
    PieceBuf& pb = c->pb[s];
    a.pieces = pb.d;
    a.piece_count = pb.cnt + 2 * pb.phase * CNT_STRIDE;
    a.piece_zero = pb.cnt + 2 * (pb.phase ^ 1) * CNT_STRIDE;
    a.piece_cap = pb.cap;
    a.counters = c->d_counters + src * C_FIELDS;
    a.overflow = c->d_counters + C_OVERFLOW;
    if (c->K > 0) {
        BatchArgs B = batch_args(c, s);
        {
            ScopedTimer t(c, c->st[s], T_PUSH + src);
            switch (src) {
                case SRC_RANGE: hipLaunchKernelGGL(k_bpush<SRC_RANGE>, dim3(grid), dim3(BLOCK), 0, c->st[s], a, B); break;
                case SRC_BITMAP: hipLaunchKernelGGL(k_bpush<SRC_BITMAP>, dim3(grid), dim3(BLOCK), 0, c->st[s], a, B); break;
                case SRC_LIST: hipLaunchKernelGGL(k_bpush<SRC_LIST>, dim3(grid), dim3(BLOCK), 0, c->st[s], a, B); break;
                default: hipLaunchKernelGGL(k_bpush<SRC_HOT>, dim3(grid), dim3(BLOCK), 0, c->st[s], a, B); break;
            }
        }
        {
            ScopedTimer t(c, c->st[s], T_PIECE);
            hipLaunchKernelGGL(k_bpieces, dim3(c->cus * 4), dim3(BLOCK), 0, c->st[s],
                               reinterpret_cast<const BPiece*>(pb.d), pb.cnt + 2 * pb.phase * CNT_STRIDE, pb.cap, a.edges, B);
        }
        pb.phase ^= 1;
        HIPCHK(hipGetLastError());
        return EXC_OK;
    }
    int G = c->push_kernel == EXC_KERNEL_GROUP ? norm_group(group > 0 ? group : 1) : 0;
    PushTrial trial;
    if (c->push_kernel == EXC_KERNEL_TUNE) {
        c->ptune_bound = edges_bound;
        int rc2 = push_tune_pick(c, s, src, group, &G, &trial);
        if (rc2) return rc2;
    }
    {
        ScopedTimer t(c, c->st[s], T_PUSH + src);
        if (trial.on) HIPCHK(hipEventRecord(c->ptune_ev[0], c->st[s]));
        if (c->nm == 1) launch_push_src<1, 1>(c, c->st[s], grid, a, src, G);
        else if (c->nm == 2) launch_push_src<2, 2>(c, c->st[s], grid, a, src, G);
        else launch_push_src<3, 4>(c, c->st[s], grid, a, src, G);
        if (trial.on) HIPCHK(hipEventRecord(c->ptune_ev[1], c->st[s]));
    }
    if (trial.on) {
        int rc2 = push_tune_collect(c, src, trial);
        if (rc2) return rc2;
    }
    {
        ScopedTimer t(c, c->st[s], T_PIECE);
        if (c->nm == 1) launch_pieces_t<1, 1>(c, s, a.edges, a.wts);
        else if (c->nm == 2) launch_pieces_t<2, 2>(c, s, a.edges, a.wts);
        else launch_pieces_t<3, 4>(c, s, a.edges, a.wts);
    }
    pb.phase ^= 1;
    HIPCHK(hipGetLastError());
    return EXC_OK;
}

// Sorted-pull launch for tile T (one of 512 / 1024 / 2048 / 4096).
template <int NM, int S, bool W>
void launch_sorted(Ctx* c, hipStream_t st, const PullArgs& pa) {
    const SortedTiles tiles{c->sorted.start, c->sorted.row, c->sorted.dst, c->sorted.w};
    const unsigned grid = (unsigned)(((long long)c->n + c->sorted_T - 1) / c->sorted_T);
    if (c->sorted_T == 4096) hipLaunchKernelGGL((k_pull_sorted<NM, S, 4096, W>), dim3(grid), dim3(BLOCK), 0, st, pa, tiles, c->DM);
    else if (c->sorted_T == 512) hipLaunchKernelGGL((k_pull_sorted<NM, S, 512, W>), dim3(grid), dim3(BLOCK), 0, st, pa, tiles, c->DM);
    else if (c->sorted_T == 1024) hipLaunchKernelGGL((k_pull_sorted<NM, S, 1024, W>), dim3(grid), dim3(BLOCK), 0, st, pa, tiles, c->DM);
    else hipLaunchKernelGGL((k_pull_sorted<NM, S, 2048, W>), dim3(grid), dim3(BLOCK), 0, st, pa, tiles, c->DM);
}

// A plain dense pass over measures [0, NM) of a context with state stride S (NM = nm_act:
// every measure of the context, or the prefix still pending).
template <int NM, int S>
void launch_fold_pull(Ctx* c, hipStream_t st) {
    static_assert(NM <= S, "measures fit the state row");
    const bool sorted = c->sorted_T && NM <= c->sorted_w && sorted_fits(NM, c->sorted_T);
    // the bitmap costs a probe per in-edge and saves the row gathers of the sources that did
    // not push: it pays once the pass's demand (out-edges of pending sources) is below 0.8 m
    // for k_pull, and below 0.99 m for the sorted pull, whose probes of neighbouring rows
    // share bitmap words (R-MAT scale 24, --measure all: 108.8 GTEPS with both at 0.6, 111.1
    // with both at 0.8; at 1.0 the sorted pull gains another 0.5 ms per run, k_pull loses 8)
    const double frac = sorted ? c->pull_live_frac_sorted : c->pull_live_frac;
    const bool live = c->nsrc && (c->pull_live == 1 ||
                                  (c->pull_live == 2 && (double)c->last_demand <= frac * (double)c->m));
    unsigned long long* row = c->d_counters + (live ? C_DENSE_LIVE : C_DENSE) * C_FIELDS;
    // every measure of the context (state stride S holds 1, 2 or 3 measures)
    constexpr bool full = S == 1 || (S == 2 && NM == 2) || (S == 4 && NM == 3);
    const int hub_field = full ? C_PULL_HUB_LIVE : NM == 1 ? C_PULL_HUB_LIVE_A1 : C_PULL_HUB_LIVE_A2;
    FoldArgs f{c->n, c->d_idx, c->d_p, c->d_rank, row};
    const int ntiles = (int)(((long long)c->n + BLOCK - 1) / BLOCK);
    // small graphs: tiles of 64 destinations (see k_pull)
    const bool small = c->pull_small == 1 || (c->pull_small == 2 && ntiles < 4 * c->cus);
    const int ntiles64 = (int)(((long long)c->n + 63) / 64);
    PullArgs pa{c->n, c->d_tidx, c->d_tadj, c->d_p, live ? c->d_pbits : nullptr, row + hub_field, c->d_twts,
                c->d_idx, c->d_cv0, c->nc, nullptr, nullptr, nullptr, c->d_rank, c->nsrc, c->d_counters + C_SPREAD};
    pa.rows_field = full ? S_ROWS : NM == 1 ? S_ROWS_A1 : S_ROWS_A2;
    c->pull_passes_live += live;
    const bool fuse_snap = !c->d_trace;   // the trace counter stays with k_snapshot
    // fold the next pass here too while the demand is high (the next pass is then a dense pass
    // whatever its demand; below the threshold the fold stays with k_fold, so the switch to
    // push passes needs no unfolding)
    const bool fuse_fold = fuse_snap && c->fuse_fold && c->fold_ahead_ok && c->d_p2 &&
                           (double)c->last_demand >= std::max(0.1, 2.0 * c->pull_threshold) * (double)c->m;
    if (c->folded && c->p_cols != NM) {   // the last pull folded this pass over more measures
        ScopedTimer t(c, st, T_OTHER);
        const unsigned g = (unsigned)std::max(1, std::min((c->nsrc + BLOCK - 1) / BLOCK, c->cus * 8));
        if (c->p_cols == 3 && NM == 2) hipLaunchKernelGGL((k_prepack<3, 2>), dim3(g), dim3(BLOCK), 0, st, c->d_p, c->d_p2, c->nsrc);
        else if (c->p_cols == 3) hipLaunchKernelGGL((k_prepack<3, 1>), dim3(g), dim3(BLOCK), 0, st, c->d_p, c->d_p2, c->nsrc);
        else hipLaunchKernelGGL((k_prepack<2, 1>), dim3(g), dim3(BLOCK), 0, st, c->d_p, c->d_p2, c->nsrc);
        std::swap(c->d_p, c->d_p2);
        pa.p = c->d_p;
        f.p = c->d_p;
        c->p_cols = NM;
    }
    if (fuse_snap) {
        clear_counts(c, st);
        pa.bitmap = c->d_bitmap;
        pa.cnt = c->d_cnt;
        pa.l1 = l1_slots(c);
        if (fuse_fold) pa.p_next = c->d_p2;
        c->snap_nm = NM;
    }
    if (!c->folded) {   // else the last pull already folded this pass into p
        ScopedTimer t(c, st, T_FOLD);
        hipLaunchKernelGGL((k_fold<NM, S>), dim3(std::min(ntiles, c->cus * 8)), dim3(BLOCK), 0, st, f, c->DM);
        c->fold_fixed_bytes += 4.0 * NM * ((double)c->n + (double)c->nsrc);
    }
    if (live) {
        ScopedTimer t(c, st, T_OTHER);
        const int waves = (c->nsrc + WAVE - 1) / WAVE;
        hipLaunchKernelGGL((k_pbits<NM, S>), dim3(std::min((waves + NWAVES - 1) / NWAVES, c->cus * 8)), dim3(BLOCK), 0, st,
                           c->d_p, c->nsrc, c->d_pbits);
    }
    if (c->n_pull_hubs) {
        ScopedTimer t(c, st, T_PULL_HUB);
        hipLaunchKernelGGL((k_pull_hub<NM, S>), dim3((unsigned)std::min(c->n_pull_hubs, c->cus * 8)), dim3(BLOCK), 0, st,
                           pa, c->d_pull_hubs, c->n_pull_hubs, c->DM);
    }
    {
        ScopedTimer t(c, st, full ? T_PULL : NM == 1 ? T_PULL_A1 : T_PULL_A2);
        bool weighted = false;
        if constexpr (NM == 1) {
            if (c->d_twts) {   // weighted graphs (one measure)
                if (sorted) launch_sorted<NM, S, true>(c, st, pa);
                else hipLaunchKernelGGL((k_pull<NM, S, 8, true>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
                weighted = true;
            }
        }
        if (weighted) {
        } else if (sorted) launch_sorted<NM, S, false>(c, st, pa);
        else if (small) hipLaunchKernelGGL((k_pull<NM, S, 8, false, 64>), dim3(ntiles64), dim3(BLOCK), 0, st, pa, c->DM);
        else if constexpr (!full) hipLaunchKernelGGL((k_pull<NM, S, 8>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
        else if (c->pull_seq == 4) hipLaunchKernelGGL((k_pull<NM, S, 4>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
        else if (c->pull_seq == 32) hipLaunchKernelGGL((k_pull<NM, S, 32>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
        else if (c->pull_seq == 16) hipLaunchKernelGGL((k_pull<NM, S, 16>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
        else hipLaunchKernelGGL((k_pull<NM, S, 8>), dim3(ntiles), dim3(BLOCK), 0, st, pa, c->DM);
    }
    {   // algorithmic bytes known at launch: every non-hub in-edge's entry (the sorted pull: row
        // id and 16-bit destination offset) and coefficient, every destination's offsets and r
        // row (read + write), the live bitmap; k_pull_hub: its entries, destinations, bitmap,
        // and without the bitmap every entry's row
        const double me = (double)(c->m - c->pull_hub_edges), nv = (double)(c->n - c->pull_hub_vertices);
        const double bm = live ? 8.0 * ((c->nsrc + 63) / 64) : 0.0;
        const double entry = (sorted ? 6.0 : 4.0) + (c->d_twts ? 4.0 : 0.0);
        c->pull_fixed_bytes[NM - 1] += me * entry + nv * (8.0 + 8.0 * NM) + bm;
        c->hub_fixed_bytes[NM - 1] += (double)c->pull_hub_edges * (c->d_twts ? 8.0 : 4.0) +
                                      (double)c->pull_hub_vertices * (8.0 + 8.0 * NM) + bm +
                                      (live ? 0.0 : 4.0 * NM * (double)c->pull_hub_edges);
        c->pull_launches_w[NM - 1]++;
        c->pull_tile_w[NM - 1] = sorted ? c->sorted_T : 0;
    }
    c->snap_ready = fuse_snap;
    if (fuse_snap) {   // the next snapshot's counts go to the host behind the pull
        c->snap_bytes += 4.0 * (((double)c->n + 31) / 32);
        if (!c->in_burst)
            c->counts_enqueued = hipMemcpyAsync(c->h_small, c->d_cnt, cnt_words(c) * sizeof(unsigned long long),
                                                hipMemcpyDeviceToHost, st) == hipSuccess;
    }
    c->folded = fuse_fold;
    if (fuse_fold) {   // the rows just written serve the next pass
        std::swap(c->d_p, c->d_p2);
        c->p_cols = NM;
    }
}

constexpr int BQ_CAND[2] = {8, 16};

// Reads the timed trial pass (every pass ends with a device synchronisation, so its
// events are complete) and, after the last candidate, keeps the fastest.
void bq_collect(Ctx* c) {
    if (!c->bq_recorded) return;
    c->bq_recorded = false;
    float ms = 0.f;
    if (hipEventSynchronize(c->bq_ev[1]) != hipSuccess || hipEventElapsedTime(&ms, c->bq_ev[0], c->bq_ev[1]) != hipSuccess) {
        c->bq_tune = false;   // no timing: keep the current value
        return;
    }
    c->bq_ns_edge[c->bq_trial] = (double)ms * 1e6 / std::max(1.0, c->bq_demand);
    if (++c->bq_trial == 2) {
        c->bpull_q = c->bq_ns_edge[1] < c->bq_ns_edge[0] ? BQ_CAND[1] : BQ_CAND[0];
        c->bq_tune = false;
    }
}

// Batched dense pass: fold the snapshot's pending vertices, hub pieces, then the pull.
template <typename PT>
void launch_bfold_pull_rows(Ctx* c, hipStream_t st, bool fuse, bool trial) {
    unsigned long long* row = c->d_counters + C_DENSE_LIVE * C_FIELDS;
    BDenseArgs<PT> a{c->n, c->d_idx, c->d_bitmap, reinterpret_cast<PT*>(c->d_p), prow_stride(c->K, sizeof(PT) == 2),
                     c->d_rank, c->d_inv, c->d_pbits, c->nsrc, c->d_tidx, c->d_tadj, row,
                     row + C_PULL_HUB_LIVE, fuse ? reinterpret_cast<PT*>(c->d_p2) : nullptr, c->d_bitmap2,
                     c->d_hubbits, c->d_cv0, c->nc, c->d_cnt, c->d_counters + C_SPREAD, c->d_tile_next};
    const BatchArgs B = batch_args(c, -1);
    const long long segs = ((long long)c->n + WAVE - 1) / WAVE;
    const unsigned cap = (unsigned)c->cus * 8;
    if (!c->folded) {   // else the last pull already folded this pass into p
        ScopedTimer t(c, st, T_FOLD);
        hipLaunchKernelGGL(k_bfold<PT>, dim3((unsigned)std::min<long long>((segs + NWAVES - 1) / NWAVES, cap)),
                           dim3(BLOCK), 0, st, a, B);
    }
    {
        ScopedTimer t(c, st, T_FOLD);
        const long long rw = ((long long)c->nsrc + WAVE - 1) / WAVE;
        hipLaunchKernelGGL(k_blive<PT>, dim3((unsigned)std::max<long long>(1, std::min<long long>((rw + NWAVES - 1) / NWAVES, cap))),
                           dim3(BLOCK), 0, st, a);
    }
    if (fuse) {   // the next snapshot's counts are accumulated by k_bpull
        clear_counts(c, st);
        hipLaunchKernelGGL(k_zero_u32, dim3((c->K + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, c->d_colpend, c->K);
    }
    if (trial) hipEventRecord(c->bq_ev[0], st);
    if (c->n_pull_hubs) {
        ScopedTimer t(c, st, T_PULL_HUB);
        const dim3 g((unsigned)std::min<long long>((c->n_pull_hubs + NWAVES - 1) / NWAVES, cap));
        // bf16 hub pieces: 8 rows in flight (16 measured 2.5x slower there, 1.69 against
        // 4.21 TB/s at R-MAT scale 26, K = 64, while the pull itself gains from 16);
        // EXC_BHUB_Q=16 for measurements
        if ((c->bpull_q == 16 && sizeof(PT) == 4) || c->bhub_q == 16)
            hipLaunchKernelGGL((k_bpull_hub<16, PT>), g, dim3(BLOCK), 0, st, a, B, c->d_pull_hubs, c->n_pull_hubs);
        else
            hipLaunchKernelGGL((k_bpull_hub<8, PT>), g, dim3(BLOCK), 0, st, a, B, c->d_pull_hubs, c->n_pull_hubs);
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(BLOCK), 0, st, c->d_tile_next, 1);
    {
        ScopedTimer t(c, st, T_PULL);
        const dim3 g((unsigned)std::min<long long>((segs + NWAVES - 1) / NWAVES, cap));
        if (c->bpull_q == 16) hipLaunchKernelGGL((k_bpull<16, PT>), g, dim3(BLOCK), 0, st, a, B);
        else hipLaunchKernelGGL((k_bpull<8, PT>), g, dim3(BLOCK), 0, st, a, B);
    }
}

void launch_bfold_pull(Ctx* c, hipStream_t st) {
    bq_collect(c);
    const bool trial = c->bq_tune && (double)c->last_demand >= 0.5 * (double)c->m;
    if (trial) c->bpull_q = BQ_CAND[c->bq_trial];
    const bool fuse = c->fuse_fold && c->fold_ahead_ok && c->d_p2 && c->d_bitmap2 && !c->d_trace &&
                      (double)c->last_demand >= std::max(0.1, 2.0 * c->pull_threshold) * (double)c->m;
    if (c->brows16) launch_bfold_pull_rows<uint16_t>(c, st, fuse, trial);
    else launch_bfold_pull_rows<float>(c, st, fuse, trial);
    if (trial) {
        hipEventRecord(c->bq_ev[1], st);
        c->bq_recorded = true;
        c->bq_demand = (double)c->last_demand;
    }
    c->pull_passes_live++;
    c->folded = fuse;
    c->snap_ready = fuse;
    if (fuse)   // the next snapshot's counts go to the host behind the pull
        c->counts_enqueued = hipMemcpyAsync(c->h_small, c->d_cnt, cnt_words(c) * sizeof(unsigned long long),
                                            hipMemcpyDeviceToHost, st) == hipSuccess;
    if (fuse) {   // the rows and bitmap just written serve the next pass
        std::swap(c->d_p, c->d_p2);
        std::swap(c->d_bitmap, c->d_bitmap2);
        c->snap_bytes += 4.0 * (((double)c->n + 31) / 32);
    }
}

int fold_pull(Ctx* c) {
    hipStream_t st = c->st[0];
    c->front_ok = false;   // a dense pass changes every row
    if (c->K) launch_bfold_pull(c, st);
    else if (c->nm == 1) launch_fold_pull<1, 1>(c, st);
    else if (c->nm == 2 && c->nm_act == 2) launch_fold_pull<2, 2>(c, st);
    else if (c->nm == 2) launch_fold_pull<1, 2>(c, st);
    else if (c->nm_act == 3) launch_fold_pull<3, 4>(c, st);
    else if (c->nm_act == 2) launch_fold_pull<2, 4>(c, st);
    else launch_fold_pull<1, 4>(c, st);
    c->pull_passes++;
    HIPCHK(hipGetLastError());
    return EXC_OK;
}

// Range or bitmap push of vertices [v0, v1) whose edges sit at edges[idx[v] - ebase].
int push_chunk(Ctx* c, int s, int v0, int v1, const int32_t* edges, uint64_t ebase, int use_list,
               int group, int hot_skip, uint64_t edges_bound, const float* wts = nullptr) {
    if (v1 <= v0) return EXC_OK;
    PushArgs a{};
    a.v0 = v0; a.v1 = v1;
    a.idx = c->d_idx;
    a.edges = edges;
    a.wts = wts;
    a.ebase = ebase;
    if (c->K > 0) {   // batched: one wave per 64 vertices
        const long long waves = ((long long)v1 - ((long long)v0 & ~63ll) + WAVE - 1) / WAVE;
        const unsigned grid = (unsigned)((waves + NWAVES - 1) / NWAVES);
        if (use_list) a.bitmap = c->d_bitmap;
        else a.hot = hot_skip ? c->d_hot : nullptr;
        return push_launch(c, s, use_list ? SRC_BITMAP : SRC_RANGE, grid, a, group, edges_bound);
    }
    if (use_list) {
        a.bitmap = c->d_bitmap;
        const long long w0 = v0 >> 5, w1 = ((long long)v1 + 31) >> 5;
        const unsigned grid = (unsigned)((w1 - w0 + BLOCK - 1) / BLOCK);
        return push_launch(c, s, SRC_BITMAP, grid, a, group, edges_bound);
    }
    a.hot = hot_skip ? c->d_hot : nullptr;
    const unsigned grid = (unsigned)(((long long)(v1 - v0) + BLOCK - 1) / BLOCK);
    return push_launch(c, s, SRC_RANGE, grid, a, group, edges_bound);
}

// every column of a batched context (a state change can make any of them pending again)
uint64_t all_columns(int K) { return K >= 64 ? ~0ull : ((1ull << K) - 1); }

Ctx* as_ctx(void* p) { return static_cast<Ctx*>(p); }

// Entry points that read or change the state refuse while a submitted burst is not collected.
int burst_busy(const char* fn) {
    return fail(EXC_ESTATE, "%s: a submitted burst is not collected (exc_burst_collect)", fn);
}

unsigned grid_for(long long items, int cap_blocks) {
    long long g = (items + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    if (g > cap_blocks) g = cap_blocks;
    return (unsigned)g;
}

// Plain (non-batched) snapshot launch: waves take consecutive 64-vertex segments, at most 16
// blocks per CU in flight.  Shared by exc_snapshot and exc_snapshot_noalloc.
void launch_snapshot_plain(SnapArgs a, int nm, int cus, hipStream_t st) {
    const long long nseg = ((long long)a.n + WAVE - 1) / WAVE;
    const long long max_waves = (long long)cus * 16 * NWAVES;
    long long spw = (nseg + max_waves - 1) / max_waves;
    if (spw < 1) spw = 1;
    a.segs_per_wave = (int)spw;
    const long long waves = (nseg + spw - 1) / spw;
    const unsigned grid = (unsigned)((waves + NWAVES - 1) / NWAVES);
    if (nm == 1) hipLaunchKernelGGL((k_snapshot<1, 1>), dim3(grid), dim3(BLOCK), 0, st, a);
    else if (nm == 2) hipLaunchKernelGGL((k_snapshot<2, 2>), dim3(grid), dim3(BLOCK), 0, st, a);
    else hipLaunchKernelGGL((k_snapshot<3, 4>), dim3(grid), dim3(BLOCK), 0, st, a);
}

// ---------------------------------------------------------- allocation-free snapshot
//
// exc_snapshot_noalloc's scratch: the per-chunk count accumulators and hot-pending counter
// ([2 nc + 1], rounded up to a 128-byte line) followed by the FP64 residual slots
// (L1_SLOTS x L1_STRIDE doubles), laid out as the context's counts buffer.  One array per
// GPU for the life of the process; every call clears the part it uses first.
constexpr int NOALLOC_MAX_CHUNKS = 4096;
constexpr size_t NOALLOC_L1_OFF = ((2 * (size_t)NOALLOC_MAX_CHUNKS + 1 + L1_STRIDE - 1) / L1_STRIDE) * L1_STRIDE;
constexpr size_t NOALLOC_WORDS = NOALLOC_L1_OFF + (size_t)L1_SLOTS * L1_STRIDE;   // 9232 words, 72 KiB

// Whether a batched snapshot may read only the touched vertices (k_bsnapshot<true>): the
// state changed since the last snapshot by resident pushes alone, and they covered every
// chunk that snapshot found pending vertices in.
bool front_usable(const Ctx* c, int hot_skip) {
    if (!c->use_front || !c->front_ok || !c->front_pend_known || hot_skip || c->d_hot || c->d_trace) return false;
    for (int i = 0; i < c->nc; i++)
        if (c->front_pend[i] && !c->front_cov[i]) return false;
    return true;
}

// Enqueues a snapshot of the current state on `st`: counts and residual slots cleared, then
// k_snapshot (k_bsnapshot for batched contexts) into c->d_cnt and the bitmap.
void enqueue_snapshot(Ctx* c, hipStream_t st, int hot_skip) {
    clear_counts(c, st);
    SnapArgs a{};
    a.n = c->n;
    a.r = c->d_r;
    for (int k = 0; k < EXC_MAXM; k++) a.eps[k] = c->DM.eps[k];
    a.idx = c->d_idx;
    a.cv0 = c->d_cv0;
    a.nc = c->nc;
    a.hot = hot_skip ? c->d_hot : nullptr;
    a.bitmap = c->d_bitmap;
    a.cnt = c->d_cnt;
    a.hotpend = c->d_cnt + 2 * c->nc;
    a.trace = c->d_trace;
    a.l1 = c->K ? nullptr : l1_slots(c);
    const long long nseg = ((long long)c->n + WAVE - 1) / WAVE;
    if (c->K) {
        hipLaunchKernelGGL(k_zero_u32, dim3((c->K + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, c->d_colpend, c->K);
        const long long bw = (long long)c->cus * 64 * NWAVES;
        long long bspw = (nseg + bw - 1) / bw;
        if (bspw < 1) bspw = 1;
        a.segs_per_wave = (int)bspw;
        const long long bwaves = (nseg + bspw - 1) / bspw;
        const unsigned bgrid = (unsigned)((bwaves + NWAVES - 1) / NWAVES);
        const bool front = front_usable(c, hot_skip);
        {
            ScopedTimer t(c, st, T_SNAP);
            if (front) hipLaunchKernelGGL(k_bsnapshot<true>, dim3(bgrid), dim3(BLOCK), 0, st, a, batch_args(c, -1));
            else hipLaunchKernelGGL(k_bsnapshot<false>, dim3(bgrid), dim3(BLOCK), 0, st, a, batch_args(c, -1));
        }
        if (front) {   // the touch bytes and the bitmap (the touched rows are not counted)
            c->front_snaps++;
            c->snap_bytes += 2.0 * c->n + 4.0 * (((double)c->n + 31) / 32);
        } else {
            c->snap_bytes += 4.0 * c->K * c->n + 4.0 * (((double)c->n + 31) / 32);
            if (c->d_touch)   // a full snapshot leaves the touch bytes clean too
                hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)(((long long)(c->n + 3) / 4 + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                                   st, reinterpret_cast<uint32_t*>(c->d_touch), (c->n + 3) / 4);
        }
        c->front_ok = c->d_touch && !c->d_hot && !c->d_trace;
        c->front_pend_known = false;
        c->front_cov.assign(c->nc, 0);
    } else {
        ScopedTimer t(c, st, T_SNAP);
        c->snap_nm = c->nm;
        launch_snapshot_plain(a, c->nm, c->cus, st);
        c->snap_bytes += 4.0 * c->nm * c->n + 4.0 * (((double)c->n + 31) / 32);
    }
}

// The launches of a burst of `passes` passes on stream 0 (see exc_run_passes): every pass's
// count region cleared at once, the passes, then the regions copied to the host and the last
// one into the context's own count buffer.  Nothing here waits or allocates, so a push burst
// can be recorded into a graph.
int enqueue_burst(Ctx* c, int passes, int dense, int slot) {
    hipStream_t st = c->st[0];
    const size_t W = cnt_words(c);
    unsigned long long* const d_cnt = c->d_cnt;
    unsigned long long* const h_small = c->h_small;
    unsigned long long* const d_reg = c->d_burst + (size_t)slot * EXC_BURST_MAX * W;
    unsigned long long* const h_reg = c->h_burst + (size_t)slot * EXC_BURST_MAX * W;
    const long long words = (long long)passes * (long long)W;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((words + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, d_reg, (int)words);
    c->in_burst = true;
    int rc = EXC_OK;
    for (int i = 0; i < passes && rc == EXC_OK; i++) {
        // pass i writes the snapshot of the state it leaves into region i
        c->d_cnt = d_reg + i * W;
        c->h_small = h_reg + i * W;
        c->snap_ready = false;
        c->l1_valid = false;
        if (dense) {
            rc = fold_pull(c);   // the snapshot is fused into the pull
        } else {
            rc = push_chunk(c, 0, 0, c->n, c->d_arena, 0, 1, 0, 0, c->arena_edges, c->d_wts);
            if (rc == EXC_OK) enqueue_snapshot(c, st, 0);
        }
    }
    c->in_burst = false;
    c->d_cnt = d_cnt;
    c->h_small = h_small;
    if (rc != EXC_OK) return rc;
    HIPCHK(hipMemcpyAsync(h_reg, d_reg, (size_t)words * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    // the bitmap is already the last snapshot's
    HIPCHK(hipMemcpyAsync(c->d_cnt, d_reg + (passes - 1) * W, W * sizeof(unsigned long long), hipMemcpyDeviceToDevice, st));
    return EXC_OK;
}

// Measure slots from `count` on are frozen: their eps becomes infinite in the device measure
// table, so no kernel finds a vertex pending in them (they are neither pushed nor folded,
// their x and r stay as they are, and snapshots still sum their |r|); dense passes gather
// rows of the first `count` measures only.  count = nm restores every slot.
void set_active(Ctx* c, int count) {
    c->nm_act = count;
    for (int k = 0; k < c->nm; k++) c->DM.eps[k] = k < count ? c->M.eps[k] : INFINITY;
    burst_graphs_clear(c);   // recorded launches carry the measure table by value
}

// Push bursts replayed from graphs: a push burst's launches are fixed by its length and the
// phase of stream 0's piece counters (it flips at every push), so one recorded graph per
// (length, phase) serves every later burst.  Anything that changes a buffer or a by-value
// argument of those launches drops the graphs (burst_graphs_clear).
void burst_graphs_clear(Ctx* c) {
    for (auto& g : c->burst_graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    c->burst_graphs.clear();
}

int launch_push_burst_graph(Ctx* c, int passes, int slot) {
    hipStream_t st = c->st[0];
    // the launches reserve the piece buffer for the whole cache: done before recording, so
    // recording never allocates or waits
    int rc = ensure_pieces(c, 0, c->arena_edges);
    if (rc) return rc;
    const int phase = c->pb[0].phase;
    BurstGraph* hit = nullptr;
    for (auto& g : c->burst_graphs)
        if (g.passes == passes && g.phase == phase && g.slot == slot) hit = &g;
    if (!hit) {
        BurstGraph g;
        g.passes = passes;
        g.phase = phase;
        g.slot = slot;
        const bool timing = c->timing;
        const double sb = c->snap_bytes;
        c->timing = false;   // no event nodes: the burst is timed as a whole
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = enqueue_burst(c, passes, 0, slot);
        const hipError_t ce = hipStreamEndCapture(st, &g.graph);
        c->timing = timing;
        g.snap_bytes = c->snap_bytes - sb;
        c->snap_bytes = sb;
        c->pb[0].phase = phase;   // recorded, not run
        if (rc) { if (g.graph) (void)hipGraphDestroy(g.graph); return rc; }
        if (ce != hipSuccess) return fail(EXC_EHIP, "exc_run_passes: graph capture: %s", hipGetErrorString(ce));
        HIPCHK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        c->burst_graphs.push_back(g);
        hit = &c->burst_graphs.back();
    }
    {
        ScopedTimer t(c, st, T_PUSH + SRC_BITMAP);
        HIPCHK(hipGraphLaunch(hit->exec, st));
    }
    c->pb[0].phase ^= passes & 1;
    c->snap_bytes += hit->snap_bytes;
    c->snap_nm = c->nm;
    c->burst_graph_launches++;
    return EXC_OK;
}

}  // namespace

static __device__ __align__(256) unsigned long long scratch[NOALLOC_WORDS];

namespace {

// Outputs of an allocation-free snapshot: the counts copied out of the scratch, and the FP64
// residual per measure summed over the slots in slot order (the order exc_snapshot's host
// sum uses).
__global__ void k_snapshot_finish(const unsigned long long* __restrict__ scr, int nc, int nm,
                                  unsigned long long* counts, double* l1) {
    const int ncnt = 2 * nc + 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncnt; i += gridDim.x * blockDim.x) counts[i] = scr[i];
    if (blockIdx.x == 0 && threadIdx.x < nm) {
        const double* slots = reinterpret_cast<const double*>(scr + NOALLOC_L1_OFF);
        double t = 0.0;
        for (int q = 0; q < L1_SLOTS; q++) t += slots[q * L1_STRIDE + threadIdx.x];
        l1[threadIdx.x] = t;
    }
}

// Device address of `scratch` and the CU count, resolved once per device (with the device
// current) by the first call on it; host code uses only these addresses, never the
// array's name.
constexpr int NOALLOC_MAX_DEVICES = 64;
struct NoallocDev {
    unsigned long long* ptr = nullptr;
    int cus = 0;
};
NoallocDev g_noalloc[NOALLOC_MAX_DEVICES];
std::mutex g_noalloc_mu;

int noalloc_setup(int device, NoallocDev* out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoallocDev& d = g_noalloc[device];
    if (!d.ptr) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(EXC_EHIP, "exc_snapshot_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess || cus < 1)
            return fail(EXC_EHIP, "exc_snapshot_noalloc: CU count of device %d: %s", device, hipGetErrorString(e));
        d.cus = cus;
        d.ptr = static_cast<unsigned long long*>(ptr);
    }
    *out = d;
    return EXC_OK;
}

}  // namespace

void exc_internal_set_error(const char* msg) { g_err = msg; }

// ====================================================================== C ABI

extern "C" {

const char* exc_last_error(void) { return g_err.c_str(); }

int exc_device_count(int* count) {
    HIPCHK(hipGetDeviceCount(count));
    return EXC_OK;
}

int exc_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
    return EXC_OK;
}

// Device memory held for bin/gpu_ballast: blocks are taken while more than keep_free bytes
// are free, starting at 1 GiB and quartering the request after a refusal (down to 2 MiB),
// so fragmentation or other processes' allocations do not stop it short of the target.
struct MemoryHold {
    int device;
    std::vector<void*> blocks;
};

int exc_hold_memory(int device, size_t keep_free, void** handle, size_t* held_bytes) {
    if (!handle) return fail(EXC_EINVAL, "exc_hold_memory: null handle");
    *handle = nullptr;
    HIPCHK(hipSetDevice(device));
    auto* h = new MemoryHold{device, {}};
    size_t piece = 1ull << 30, held = 0;
    for (;;) {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess || freeB <= keep_free) break;
        const size_t want = std::min(piece, freeB - keep_free);
        void* b = nullptr;
        if (hipMalloc(&b, want) == hipSuccess) {
            h->blocks.push_back(b);
            held += want;
            continue;
        }
        (void)hipGetLastError();
        if (piece <= (2ull << 20)) break;
        piece /= 4;
    }
    if (held_bytes) *held_bytes = held;
    *handle = h;
    return EXC_OK;
}

int exc_release_memory(void* handle) {
    auto* h = static_cast<MemoryHold*>(handle);
    if (!h) return EXC_OK;
    int rc = EXC_OK;
    if (hipSetDevice(h->device) != hipSuccess) rc = fail(EXC_EHIP, "exc_release_memory: device %d", h->device);
    for (void* b : h->blocks)
        if (hipFree(b) != hipSuccess && rc == EXC_OK) rc = fail(EXC_EHIP, "exc_release_memory: hipFree");
    delete h;
    return rc;
}

int exc_device_info(int device, int* cus, int* integrated, char* name, int name_len) {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    if (cus) *cus = p.multiProcessorCount;
    if (integrated) *integrated = p.integrated;
    if (name && name_len > 0) { strncpy(name, p.gcnArchName, name_len - 1); name[name_len - 1] = 0; }
    return EXC_OK;
}

int exc_device_synchronize(int device) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    return EXC_OK;
}

int exc_create(const exc_config* cfg, const exc_measures* M, void** out) {
    if (!cfg || !M || !out) return fail(EXC_EINVAL, "exc_create: null argument");
    if (cfg->n < 1 || cfg->nc < 1 || !cfg->cv0) return fail(EXC_EINVAL, "exc_create: bad n/nc/cv0");
    if (M->nm < 1 || M->nm > EXC_MAXM) return fail(EXC_EINVAL, "exc_create: nm must be 1..3");
    if (cfg->push_kernel != EXC_KERNEL_LB && cfg->push_kernel != EXC_KERNEL_GROUP && cfg->push_kernel != EXC_KERNEL_TUNE)
        return fail(EXC_EINVAL, "exc_create: unknown push kernel %d", cfg->push_kernel);
    Ctx* c = new Ctx();
    *out = nullptr;
    c->dev = cfg->device; c->n = cfg->n; c->m = cfg->m; c->nc = cfg->nc;
    c->nm = M->nm; c->S = M->nm == 1 ? 1 : M->nm == 2 ? 2 : 4;
    if (cfg->batch_k) {
        if (cfg->batch_k < 1 || cfg->batch_k > 64) { delete c; return fail(EXC_EINVAL, "exc_create: batch_k must be 1..64"); }
        c->K = cfg->batch_k;
        c->S = c->K;
        c->nm = 1;
        c->damp = M->scale[0];
        c->active = all_columns(c->K);
    }
    c->M = *M;
    c->nm_act = c->nm;
    c->zc = cfg->zero_copy != 0;
    c->timing = cfg->timing != 0;
    c->timing_pull_only = cfg->timing == 2;
    c->push_kernel = cfg->push_kernel;
    if (const char* bg = getenv("EXC_BURST_GRAPH")) c->use_burst_graphs = atoi(bg) != 0;   // measurements
    c->max_chunk_edges = cfg->max_chunk_edges;
    c->max_chunk_verts = cfg->max_chunk_verts;
    auto bail = [&](int rc) { exc_destroy(c); return rc; };
#define CK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) \
    return bail(fail(EXC_EHIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__)); } while (0)
    CK(hipSetDevice(c->dev));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, c->dev));
    c->cus = prop.multiProcessorCount;
    const size_t n = (size_t)c->n, words = (n + 31) / 32;
    CK(hipMalloc(&c->d_idx, (n + 1) * sizeof(uint64_t)));
    CK(hipMalloc(&c->d_x, n * c->S * sizeof(float)));
    CK(hipMalloc(&c->d_r, n * c->S * sizeof(float)));
    CK(hipMemset(c->d_x, 0, n * c->S * sizeof(float)));
    CK(hipMemset(c->d_r, 0, n * c->S * sizeof(float)));
    CK(hipMalloc(&c->d_bitmap, words * sizeof(uint32_t)));
    CK(hipMalloc(&c->d_cnt, cnt_words(c) * sizeof(unsigned long long)));
    CK(hipMemset(c->d_cnt, 0, cnt_words(c) * sizeof(unsigned long long)));
    CK(hipMalloc(&c->d_cv0, (size_t)c->nc * sizeof(int)));
    CK(hipMemcpy(c->d_cv0, cfg->cv0, (size_t)c->nc * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMalloc(&c->d_counters, C_NCOUNTERS * sizeof(unsigned long long)));
    CK(hipMemset(c->d_counters, 0, C_NCOUNTERS * sizeof(unsigned long long)));
    CK(hipHostMalloc((void**)&c->h_small, cnt_words(c) * sizeof(unsigned long long), hipHostMallocDefault));
    if (cfg->trace) {
        CK(hipMalloc(&c->d_trace, n * sizeof(uint16_t)));
        CK(hipMemset(c->d_trace, 0, n * sizeof(uint16_t)));
    }
    CK(hipMalloc(&c->d_hcount, sizeof(uint32_t)));
    c->h_cv0.assign(cfg->cv0, cfg->cv0 + c->nc);
    if (c->K) {
        CK(hipMalloc(&c->d_beps, (size_t)c->K * sizeof(float)));
        CK(hipMalloc(&c->d_colpend, (size_t)c->K * sizeof(uint32_t)));
        CK(hipMemset(c->d_colpend, 0, (size_t)c->K * sizeof(uint32_t)));
        if (const char* ev = getenv("EXC_BATCH_FRONT")) c->use_front = atoi(ev) != 0;
        if (!cfg->trace) {   // touch bytes, whole 32-bit words (cleared as such)
            CK(hipMalloc(&c->d_touch, ((size_t)n + 3) / 4 * 4));
            CK(hipMemset(c->d_touch, 0, ((size_t)n + 3) / 4 * 4));
        }
        std::vector<float> e(c->K, M->eps[0]);
        CK(hipMemcpy(c->d_beps, e.data(), e.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    for (int s = 0; s < 3; s++) CK(hipStreamCreateWithFlags(&c->st[s], hipStreamNonBlocking));
    for (auto& sl : c->slot) {
        const size_t eb = (c->max_chunk_edges ? c->max_chunk_edges : 1) * sizeof(int32_t);
        const size_t vb = ((size_t)c->max_chunk_verts + 1) * sizeof(int32_t);
        const size_t ob = ((size_t)c->max_chunk_verts + 1) * sizeof(uint64_t);
        if (c->zc) {
            CK(hipHostMalloc((void**)&sl.h_edges, eb, hipHostMallocMapped));
            CK(hipHostMalloc((void**)&sl.h_vids, vb, hipHostMallocMapped));
            CK(hipHostMalloc((void**)&sl.h_off, ob, hipHostMallocMapped));
            CK(hipHostGetDevicePointer((void**)&sl.d_edges, sl.h_edges, 0));
            CK(hipHostGetDevicePointer((void**)&sl.d_vids, sl.h_vids, 0));
            CK(hipHostGetDevicePointer((void**)&sl.d_off, sl.h_off, 0));
        } else {
            CK(hipHostMalloc((void**)&sl.h_edges, eb, hipHostMallocDefault));
            CK(hipHostMalloc((void**)&sl.h_vids, vb, hipHostMallocDefault));
            CK(hipHostMalloc((void**)&sl.h_off, ob, hipHostMallocDefault));
            CK(hipMalloc(&sl.d_edges, eb));
            CK(hipMalloc(&sl.d_vids, vb));
            CK(hipMalloc(&sl.d_off, ob));
        }
        CK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    }
    // the clears above ran on the null stream; the context's streams are non-blocking
    CK(hipDeviceSynchronize());
#undef CK
    c->DM.x = c->d_x;
    c->DM.r = c->d_r;
    for (int k = 0; k < EXC_MAXM; k++) {
        c->DM.eps[k] = k < c->nm ? M->eps[k] : 0.f;
        c->DM.scale[k] = k < c->nm ? M->scale[k] : 0.f;
        c->DM.divdeg[k] = k < c->nm ? M->divdeg[k] : 0;
    }
    *out = c;
    return EXC_OK;
}

int exc_destroy(void* p) {
    Ctx* c = as_ctx(p);
    if (!c) return EXC_OK;
    hipSetDevice(c->dev);
    hipDeviceSynchronize();
    for (auto& t : c->pending_timers) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    for (auto e : c->free_events) hipEventDestroy(e);
    for (auto e : c->bq_ev) if (e) hipEventDestroy(e);
    for (auto e : c->ptune_ev) if (e) hipEventDestroy(e);
    for (auto e : c->burst_ev) if (e) hipEventDestroy(e);
    burst_graphs_clear(c);
    for (auto& sl : c->slot) {
        if (sl.h_edges) hipHostFree(sl.h_edges);
        if (sl.h_vids) hipHostFree(sl.h_vids);
        if (sl.h_off) hipHostFree(sl.h_off);
        if (!c->zc) { if (sl.d_edges) hipFree(sl.d_edges); if (sl.d_vids) hipFree(sl.d_vids); if (sl.d_off) hipFree(sl.d_off); }
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.copied) hipEventDestroy(sl.copied);
    }
    for (auto& pb : c->pb) { if (pb.d) hipFree(pb.d); if (pb.cnt) hipFree(pb.cnt); }
    for (auto s : c->st) if (s) hipStreamDestroy(s);
    for (auto q : c->prows) if (q) (void)hipFree(q);
    void* pull_ptrs[] = {c->d_tidx, c->d_tadj, c->d_p, c->d_pull_hubs, c->d_wts, c->d_twts, c->d_rank, c->d_pbits, c->d_p2,
                         c->d_bitmap2, c->d_hubbits, c->d_inv, c->d_tile_next};
    for (void* q : pull_ptrs) if (q) (void)hipFree(q);
    if (c->d_beps) (void)hipFree(c->d_beps);
    excent_dev::free_sorted_tiles(&c->sorted);
    if (c->d_hub_keys) (void)hipFree(c->d_hub_keys);
    if (c->d_hub_filter) (void)hipFree(c->d_hub_filter);
    if (c->d_colpend) (void)hipFree(c->d_colpend);
    if (c->d_touch) (void)hipFree(c->d_touch);
    void* dev_ptrs[] = {c->d_idx, c->d_x, c->d_r, c->d_x0, c->d_r0, c->d_bitmap, c->d_cnt, c->d_cv0, c->d_trace, c->d_counters,
                        c->d_arena, c->d_hids, c->d_hoff, c->d_hedges, c->d_hot, c->d_hlist, c->d_hcount};
    for (void* q : dev_ptrs) if (q) hipFree(q);
    if (c->h_small) hipHostFree(c->h_small);
    if (c->d_burst) hipFree(c->d_burst);
    if (c->h_burst) hipHostFree(c->h_burst);
    delete c;
    return EXC_OK;
}

int exc_upload_offsets(void* p, const uint64_t* h_idx, size_t v0, size_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || !h_idx || v0 + cnt > (size_t)c->n + 1) return fail(EXC_EINVAL, "exc_upload_offsets: bad range");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(c->d_idx + v0, h_idx, cnt * sizeof(uint64_t), hipMemcpyHostToDevice));
    return EXC_OK;
}

int exc_init_measure(void* p, int i, float beta, float lambda) {
    Ctx* c = as_ctx(p);
    if (!c || i < 0 || i >= c->nm) return fail(EXC_EINVAL, "exc_init_measure: bad slot");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    c->folded = false;       // the state is replaced: a folded pass no longer applies
    set_active(c, c->nm);    // ... and every measure may be pending again
    if (c->K) c->active = all_columns(c->K);
    if (c->K) {   // batched: every column starts at x = r = 0 (exc_batch_seed adds the teleport mass)
        const size_t bytes = (size_t)c->n * c->S * sizeof(float);
        HIPCHK(hipMemsetAsync(c->d_x, 0, bytes, c->st[0]));
        HIPCHK(hipMemsetAsync(c->d_r, 0, bytes, c->st[0]));
        HIPCHK(hipStreamSynchronize(c->st[0]));
        return EXC_OK;
    }
    ScopedTimer t(c, c->st[0], T_OTHER);
    hipLaunchKernelGGL(k_init, dim3(c->cus * 8), dim3(BLOCK), 0, c->st[0], c->n, c->d_idx, c->d_x, c->d_r,
                       c->S, i, c->M.kind[i], beta, lambda);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st[0]));
    return EXC_OK;
}

int exc_set_state(void* p, int i, const float* h_x, const float* h_r) {
    Ctx* c = as_ctx(p);
    if (!c || i < 0 || i >= (c->K ? c->K : c->nm)) return fail(EXC_EINVAL, "exc_set_state: bad slot");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    c->folded = false;       // the state is replaced: a folded pass no longer applies
    set_active(c, c->nm);    // ... and every measure may be pending again
    if (c->K) c->active = all_columns(c->K);
    float* tmp;
    const size_t bytes = (size_t)c->n * sizeof(float);
    HIPCHK(hipMalloc(&tmp, bytes));
    const float* src[2] = {h_x, h_r};
    float* dst[2] = {c->d_x, c->d_r};
    for (int w = 0; w < 2; w++) {
        if (!src[w]) continue;
        hipError_t e = hipMemcpy(tmp, src[w], bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) { hipFree(tmp); return fail(EXC_EHIP, "exc_set_state: %s", hipGetErrorString(e)); }
        hipLaunchKernelGGL(k_scatter_col, dim3(c->cus * 8), dim3(BLOCK), 0, c->st[0], c->n, tmp, dst[w], c->S, i);
        HIPCHK(hipStreamSynchronize(c->st[0]));
    }
    hipFree(tmp);
    return EXC_OK;
}

int exc_read_state(void* p, int i, int which, float* h_out, size_t v0, size_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || i < 0 || i >= (c->K ? c->K : c->nm) || v0 + cnt > (size_t)c->n || !h_out)
        return fail(EXC_EINVAL, "exc_read_state: bad arguments");
    if (!cnt) return EXC_OK;
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    float* tmp;
    HIPCHK(hipMalloc(&tmp, cnt * sizeof(float)));
    const float* src = (which ? c->d_r : c->d_x) + v0 * c->S;
    hipLaunchKernelGGL(k_gather_col, dim3(c->cus * 8), dim3(BLOCK), 0, c->st[0], (long long)cnt, src, tmp, c->S, i);
    hipError_t e = hipMemcpyAsync(h_out, tmp, cnt * sizeof(float), hipMemcpyDeviceToHost, c->st[0]);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st[0]);
    hipFree(tmp);
    if (e != hipSuccess) return fail(EXC_EHIP, "exc_read_state: %s", hipGetErrorString(e));
    return EXC_OK;
}

int exc_state_checkpoint(void* p, int restore) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_state_checkpoint: null context");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    if (c->K) c->active = all_columns(c->K);
    const size_t bytes = (size_t)c->n * c->S * sizeof(float);
    if (restore) {
        c->folded = false;
        set_active(c, c->nm);
        if (!c->d_x0) return fail(EXC_ESTATE, "exc_state_checkpoint: nothing saved");
        HIPCHK(hipMemcpyAsync(c->d_x, c->d_x0, bytes, hipMemcpyDeviceToDevice, c->st[0]));
        HIPCHK(hipMemcpyAsync(c->d_r, c->d_r0, bytes, hipMemcpyDeviceToDevice, c->st[0]));
    } else {
        if (!c->d_x0) {
            if (hipMalloc(&c->d_x0, bytes) != hipSuccess || hipMalloc(&c->d_r0, bytes) != hipSuccess) {
                (void)hipGetLastError();
                if (c->d_x0) (void)hipFree(c->d_x0);
                c->d_x0 = nullptr;
                c->d_r0 = nullptr;
                return fail(EXC_ENOMEM, "exc_state_checkpoint: cannot allocate %.2f GB", 2.0 * bytes / 1e9);
            }
        }
        HIPCHK(hipMemcpyAsync(c->d_x0, c->d_x, bytes, hipMemcpyDeviceToDevice, c->st[0]));
        HIPCHK(hipMemcpyAsync(c->d_r0, c->d_r, bytes, hipMemcpyDeviceToDevice, c->st[0]));
    }
    HIPCHK(hipStreamSynchronize(c->st[0]));
    return EXC_OK;
}

// Per-chunk counts (and hot-pending) of the snapshot on the device to the host.
// Parses the counts buffer's host copy (h_small): per-chunk counts, hot-pending, residual sums
// and maxima.
void parse_counts(Ctx* c, uint64_t* h_cnt, uint64_t* h_hotpend);

int snapshot_counts(Ctx* c, uint64_t* h_cnt, uint64_t* h_hotpend) {
    hipStream_t st = c->st[0];
    HIPCHK(hipMemcpyAsync(c->h_small, c->d_cnt, cnt_words(c) * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    parse_counts(c, h_cnt, h_hotpend);
    return EXC_OK;
}

void parse_counts(Ctx* c, uint64_t* h_cnt, uint64_t* h_hotpend) {
    memcpy(h_cnt, c->h_small, 2 * (size_t)c->nc * sizeof(uint64_t));
    if (h_hotpend) *h_hotpend = c->h_small[2 * c->nc];
    const double* l1 = reinterpret_cast<const double*>(c->h_small + l1_offset(c));
    for (int k = 0; k < EXC_MAXM; k++) {
        double t = 0.0;
        uint32_t m = 0;
        for (int q = 0; q < L1_SLOTS; q++) {
            t += l1[q * L1_STRIDE + k];
            uint32_t b;
            memcpy(&b, reinterpret_cast<const char*>(l1 + q * L1_STRIDE + LINF_AT) + 4 * k, 4);
            m = std::max(m, b);
        }
        float f;
        memcpy(&f, &m, 4);
        if (k >= c->nm) { c->l1_last[k] = 0.0; c->linf_last[k] = 0.f; }
        else if (k < c->snap_nm) { c->l1_last[k] = t; c->linf_last[k] = f; }
        // else: a dropped measure, not summed by the pull that took this snapshot; its state
        // has not changed since the snapshot that found it converged

    }
    c->l1_valid = c->K == 0;
    c->last_demand = 0;
    for (int i = 0; i < c->nc; i++) c->last_demand += c->h_small[2 * i + 1];
    if (c->K) {
        c->front_pend.resize(c->nc);
        for (int i = 0; i < c->nc; i++) c->front_pend[i] = c->h_small[2 * i];
        c->front_pend_known = true;
    }
}

int exc_snapshot(void* p, int hot_skip, uint64_t* h_cnt, uint64_t* h_hotpend) {
    Ctx* c = as_ctx(p);
    if (!c || !h_cnt) return fail(EXC_EINVAL, "exc_snapshot: bad arguments");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    hipStream_t st = c->st[0];
    if ((c->snap_ready || c->folded) && !hot_skip) {   // computed by the dense pass that ended the last pass
        c->snap_ready = false;
        if (c->counts_copied) {   // its copy was queued behind the pull and the pass sync waited for it
            c->counts_copied = false;
            parse_counts(c, h_cnt, h_hotpend);
            return EXC_OK;
        }
        return snapshot_counts(c, h_cnt, h_hotpend);
    }
    c->counts_copied = false;
    c->snap_ready = false;
    enqueue_snapshot(c, st, hot_skip);
    HIPCHK(hipGetLastError());
    return snapshot_counts(c, h_cnt, h_hotpend);
}

int exc_bitmap_to_host(void* p, uint32_t* h_bitmap) {
    Ctx* c = as_ctx(p);
    if (!c || !h_bitmap) return fail(EXC_EINVAL, "exc_bitmap_to_host: bad arguments");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpyAsync(h_bitmap, c->d_bitmap, (((size_t)c->n + 31) / 32) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, c->st[0]));
    HIPCHK(hipStreamSynchronize(c->st[0]));
    return EXC_OK;
}

int exc_read_trace(void* p, uint16_t* h_out, size_t v0, size_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || !c->d_trace || v0 + cnt > (size_t)c->n) return fail(EXC_EINVAL, "exc_read_trace: no trace or bad range");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(h_out, c->d_trace + v0, cnt * sizeof(uint16_t), hipMemcpyDeviceToHost));
    return EXC_OK;
}

int exc_cache_chunks(void* p, uint64_t resident_edges) {
    Ctx* c = as_ctx(p);
    if (!c || resident_edges > c->m) return fail(EXC_EINVAL, "exc_cache_chunks: bad size");
    HIPCHK(hipSetDevice(c->dev));
    burst_graphs_clear(c);   // recorded pushes name the old arena
    if (c->d_arena) { (void)hipFree(c->d_arena); c->d_arena = nullptr; c->arena_edges = 0; }
    // a prefix of zero-edge chunks is still a cache: keep one element so the pointer is valid
    if (hipMalloc(&c->d_arena, (resident_edges ? resident_edges : 1) * sizeof(int32_t)) != hipSuccess) {
        (void)hipGetLastError();
        c->d_arena = nullptr;
        return fail(EXC_ENOMEM, "cannot allocate %.2f GB edge cache", resident_edges * 4.0 / 1e9);
    }
    c->arena_edges = resident_edges;
    return EXC_OK;
}

int exc_cache_upload(void* p, const int32_t* h_edges, uint64_t e0, uint64_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || !c->d_arena || e0 + cnt > c->arena_edges) return fail(EXC_EINVAL, "exc_cache_upload: bad range");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(c->d_arena + e0, h_edges, cnt * sizeof(int32_t), hipMemcpyHostToDevice));
    return EXC_OK;
}

int exc_cache_hot(void* p, int nh, const int32_t* ids, const uint64_t* off, const uint32_t* bitmask) {
    Ctx* c = as_ctx(p);
    if (!c || nh < 1 || !ids || !off || !bitmask) return fail(EXC_EINVAL, "exc_cache_hot: bad arguments");
    HIPCHK(hipSetDevice(c->dev));
    const uint64_t he = off[nh];
    const size_t words = ((size_t)c->n + 31) / 32;
    bool ok = true;
    ok = ok && hipMalloc(&c->d_hedges, (he ? he : 1) * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_hids, (size_t)nh * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_hoff, ((size_t)nh + 1) * sizeof(uint64_t)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_hot, words * sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_hlist, (size_t)nh * sizeof(int32_t)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        void* q[] = {c->d_hedges, c->d_hids, c->d_hoff, c->d_hot, c->d_hlist};
        for (void* x : q) if (x) hipFree(x);
        c->d_hedges = nullptr; c->d_hids = nullptr; c->d_hoff = nullptr; c->d_hot = nullptr; c->d_hlist = nullptr;
        return fail(EXC_ENOMEM, "cannot allocate %d hot lists (%.2f GB)", nh, he * 4.0 / 1e9);
    }
    HIPCHK(hipMemcpy(c->d_hids, ids, (size_t)nh * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_hoff, off, ((size_t)nh + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_hot, bitmask, words * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->nh = nh;
    c->hot_edges = he;
    return EXC_OK;
}

int exc_cache_hot_upload(void* p, const int32_t* h_edges, uint64_t off0, uint64_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || !c->d_hedges || off0 + cnt > c->hot_edges) return fail(EXC_EINVAL, "exc_cache_hot_upload: bad range");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(c->d_hedges + off0, h_edges, cnt * sizeof(int32_t), hipMemcpyHostToDevice));
    return EXC_OK;
}

int exc_slot_buffers(void* p, int s, int32_t** h_edges, int32_t** h_vids, uint64_t** h_off) {
    Ctx* c = as_ctx(p);
    if (!c || s < 0 || s > 1) return fail(EXC_EINVAL, "exc_slot_buffers: bad slot");
    if (h_edges) *h_edges = c->slot[s].h_edges;
    if (h_vids) *h_vids = c->slot[s].h_vids;
    if (h_off) *h_off = c->slot[s].h_off;
    return EXC_OK;
}

int exc_slot_acquire(void* p, int s, double* waited) {
    Ctx* c = as_ctx(p);
    if (!c || s < 0 || s > 1) return fail(EXC_EINVAL, "exc_slot_acquire: bad slot");
    if (waited) *waited = 0;
    Slot& sl = c->slot[s];
    if (!sl.busy) return EXC_OK;
    auto a = std::chrono::steady_clock::now();
    HIPCHK(hipEventSynchronize(sl.frees));
    if (waited) *waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    sl.busy = false;
    return EXC_OK;
}

int exc_slot_copy_probe(void* p, int s, uint64_t bytes, double* seconds) {
    Ctx* c = as_ctx(p);
    if (!c || s < 0 || s > 1 || bytes > c->max_chunk_edges * sizeof(int32_t) || !seconds)
        return fail(EXC_EINVAL, "exc_slot_copy_probe: bad arguments");
    HIPCHK(hipSetDevice(c->dev));
    Slot& sl = c->slot[s];
    auto a = std::chrono::steady_clock::now();
    if (!c->zc && bytes) {
        HIPCHK(hipMemcpyAsync(sl.d_edges, sl.h_edges, bytes, hipMemcpyHostToDevice, c->st[1 + s]));
        HIPCHK(hipStreamSynchronize(c->st[1 + s]));
    }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    return EXC_OK;
}

int exc_push_resident(void* p, const exc_chunk* ch, int use_list, int group, int hot_skip) {
    Ctx* c = as_ctx(p);
    if (!c || !ch || ch->v0 < 0 || ch->v1 > c->n || ch->v0 > ch->v1) return fail(EXC_EINVAL, "exc_push_resident: bad chunk");
    if (!c->d_arena || ch->e1 > c->arena_edges) return fail(EXC_ESTATE, "exc_push_resident: chunk not in the edge cache");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->l1_valid = false;
    const bool whole = ch->v0 == 0 && ch->v1 == c->n && !hot_skip;
    if (c->folded && !whole)
        return fail(EXC_ESTATE, "exc_push_resident: the last dense pass folded this pass; it must be a dense pass over all vertices");
    if (c->d_tadj && (!c->d_wts || c->d_twts) && whole &&
        (c->folded || (double)c->last_demand >= c->pull_threshold * (double)c->m))
        return fold_pull(c);
    if (c->K && c->front_ok) {   // the chunks whose every vertex this push covers
        if ((int)c->front_cov.size() != c->nc) c->front_cov.assign(c->nc, 0);
        for (int i = 0; i < c->nc; i++) {
            const int lo = c->h_cv0[i], hi = i + 1 < c->nc ? c->h_cv0[i + 1] : c->n;
            if (ch->v0 <= lo && hi <= ch->v1) c->front_cov[i] = 1;
        }
    }
    return push_chunk(c, 0, ch->v0, ch->v1, c->d_arena, 0, use_list, group, hot_skip, c->arena_edges, c->d_wts);
}

int exc_stream_chunk(void* p, int s, const exc_chunk* ch, int to_resident, int hot_lo, int hot_hi,
                     int use_list, int group, int hot_skip) {
    Ctx* c = as_ctx(p);
    if (!c || s < 0 || s > 1 || !ch || ch->v0 < 0 || ch->v1 > c->n || ch->e1 < ch->e0)
        return fail(EXC_EINVAL, "exc_stream_chunk: bad arguments");
    const uint64_t ne = ch->e1 - ch->e0;
    if (c->d_wts) return fail(EXC_ESTATE, "exc_stream_chunk: weighted graphs must be fully cached");
    if (ne > c->max_chunk_edges) return fail(EXC_EINVAL, "exc_stream_chunk: chunk larger than the staging slot");
    if (to_resident && (!c->d_arena || ch->e1 > c->arena_edges)) return fail(EXC_ESTATE, "exc_stream_chunk: chunk not in the edge cache");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    if (c->folded)
        return fail(EXC_ESTATE, "exc_stream_chunk: the last dense pass folded this pass; it must be a dense pass over all vertices");
    Slot& sl = c->slot[s];
    hipStream_t st = c->st[1 + s];
    int32_t* dst;
    uint64_t ebase;
    if (to_resident) { dst = c->d_arena; ebase = 0; }
    else { dst = sl.d_edges; ebase = ch->e0; }
    if (ne) {
        if (!c->zc) HIPCHK(hipMemcpyAsync(to_resident ? dst + ch->e0 : dst, sl.h_edges, ne * sizeof(int32_t), hipMemcpyHostToDevice, st));
        else if (to_resident) HIPCHK(hipMemcpyAsync(dst + ch->e0, sl.d_edges, ne * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    }
    // the host buffer is free once this copy has run (a zero-copy chunk copied to the cache is
    // read from the host buffer, so it frees after that copy too)
    const bool free_after_copy = !c->zc || to_resident;
    if (free_after_copy) HIPCHK(hipEventRecord(sl.copied, st));
    if (hot_hi > hot_lo && c->d_hedges) {
        ScopedTimer t(c, st, T_OTHER);
        const int32_t* src = to_resident ? dst + ch->e0 : dst;
        const long long items = (long long)(hot_hi - hot_lo) * WAVE;   // one wave per hot entry, strided
        hipLaunchKernelGGL(k_fill_hot, dim3(grid_for(items, c->cus * 16)), dim3(BLOCK), 0, st,
                           hot_lo, hot_hi, c->d_hids, c->d_hoff, c->d_idx, ch->e0, src, c->d_hedges);
    }
    int rc = push_chunk(c, 1 + s, ch->v0, ch->v1, dst, ebase, use_list, group, hot_skip,
                        to_resident ? c->arena_edges : c->max_chunk_edges);
    if (rc) return rc;
    HIPCHK(hipEventRecord(sl.done, st));
    sl.frees = free_after_copy ? sl.copied : sl.done;
    sl.busy = true;
    return EXC_OK;
}

int exc_gather_chunk(void* p, int s, int k, int group) {
    Ctx* c = as_ctx(p);
    if (!c || s < 0 || s > 1 || k < 0 || k > c->max_chunk_verts) return fail(EXC_EINVAL, "exc_gather_chunk: bad arguments");
    Slot& sl = c->slot[s];
    const uint64_t ne = k ? sl.h_off[k] : 0;
    if (ne > c->max_chunk_edges) return fail(EXC_EINVAL, "exc_gather_chunk: gathered lists exceed the slot");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    if (c->folded)
        return fail(EXC_ESTATE, "exc_gather_chunk: the last dense pass folded this pass; it must be a dense pass over all vertices");
    hipStream_t st = c->st[1 + s];
    if (!c->zc) {
        HIPCHK(hipMemcpyAsync(sl.d_vids, sl.h_vids, (size_t)k * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(sl.d_off, sl.h_off, ((size_t)k + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        if (ne) HIPCHK(hipMemcpyAsync(sl.d_edges, sl.h_edges, ne * sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(sl.copied, st));   // the host buffers are free again
    }
    if (k) {
        PushArgs a{};
        a.vids = sl.d_vids;
        a.off = sl.d_off;
        a.k = k;
        a.edges = sl.d_edges;
        const unsigned grid = c->K ? grid_for((long long)k * WAVE, c->cus * 16) : (unsigned)((k + BLOCK - 1) / BLOCK);
        int rc = push_launch(c, 1 + s, SRC_LIST, grid, a, group, c->max_chunk_edges);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(sl.done, st));
    sl.frees = c->zc ? sl.done : sl.copied;
    sl.busy = true;
    return EXC_OK;
}

int exc_push_hot(void* p, int lo, int hi, int group, int after_slots) {
    Ctx* c = as_ctx(p);
    if (!c || lo < 0 || hi > c->nh || lo > hi) return fail(EXC_EINVAL, "exc_push_hot: bad range");
    if (hi == lo) return EXC_OK;
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    if (c->folded)
        return fail(EXC_ESTATE, "exc_push_hot: the last dense pass folded this pass; it must be a dense pass over all vertices");
    hipStream_t st = c->st[0];
    if (after_slots)
        for (auto& sl : c->slot) if (sl.busy) HIPCHK(hipStreamWaitEvent(st, sl.done, 0));
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, c->d_hcount, 1);
    if (c->K) {
        {
            ScopedTimer t(c, st, T_OTHER);
            const long long waves = ((long long)(hi - lo) + WAVE - 1) / WAVE;
            hipLaunchKernelGGL(k_bcompact_hot, dim3((unsigned)((waves + NWAVES - 1) / NWAVES)), dim3(BLOCK), 0, st, lo, hi,
                               c->d_hids, batch_args(c, -1), c->d_hlist, c->d_hcount);
        }
        PushArgs a{};
        a.hlist = c->d_hlist;
        a.hcount = c->d_hcount;
        a.hids = c->d_hids;
        a.hoff = c->d_hoff;
        a.edges = c->d_hedges;
        return push_launch(c, 0, SRC_HOT, grid_for((long long)(hi - lo) * WAVE, c->cus * 16), a, group, c->hot_edges);
    }
    {
        ScopedTimer t(c, st, T_OTHER);
        const unsigned grid = (unsigned)(((long long)(hi - lo) + BLOCK - 1) / BLOCK);
        if (c->nm == 1) hipLaunchKernelGGL((k_compact_hot<1, 1>), dim3(grid), dim3(BLOCK), 0, st, lo, hi, c->d_hids, c->DM, c->d_hlist, c->d_hcount);
        else if (c->nm == 2) hipLaunchKernelGGL((k_compact_hot<2, 2>), dim3(grid), dim3(BLOCK), 0, st, lo, hi, c->d_hids, c->DM, c->d_hlist, c->d_hcount);
        else hipLaunchKernelGGL((k_compact_hot<3, 4>), dim3(grid), dim3(BLOCK), 0, st, lo, hi, c->d_hids, c->DM, c->d_hlist, c->d_hcount);
    }
    PushArgs a{};
    a.hlist = c->d_hlist;
    a.hcount = c->d_hcount;
    a.hids = c->d_hids;
    a.hoff = c->d_hoff;
    a.edges = c->d_hedges;
    return push_launch(c, 0, SRC_HOT, (unsigned)(((long long)(hi - lo) + BLOCK - 1) / BLOCK), a, group, c->hot_edges);
}

int exc_batch_eps(void* p, const float* eps) {
    Ctx* c = as_ctx(p);
    if (!c || !c->K || !eps) return fail(EXC_EINVAL, "exc_batch_eps: not a batched context");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(c->d_beps, eps, (size_t)c->K * sizeof(float), hipMemcpyHostToDevice));
    c->front_ok = false;   // the pending test changes for every vertex
    return EXC_OK;
}

int exc_batch_seed(void* p, int col, const int32_t* seeds, int count, float value) {
    Ctx* c = as_ctx(p);
    if (!c || !c->K || col < 0 || col >= c->K || count < 0 || (count && !seeds))
        return fail(EXC_EINVAL, "exc_batch_seed: bad arguments");
    for (int i = 0; i < count; i++)
        if (seeds[i] < 0 || seeds[i] >= c->n) return fail(EXC_EINVAL, "exc_batch_seed: seed %d out of range", seeds[i]);
    if (!count) return EXC_OK;
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    c->folded = false;       // the state is replaced: a folded pass no longer applies
    set_active(c, c->nm);    // ... and every measure may be pending again
    if (c->K) c->active = all_columns(c->K);
    int32_t* d;
    HIPCHK(hipMalloc(&d, (size_t)count * sizeof(int32_t)));
    hipError_t e = hipMemcpy(d, seeds, (size_t)count * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bseed, dim3((count + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->st[0], count, d, c->d_r, c->K, col, value);
        e = hipStreamSynchronize(c->st[0]);
    }
    (void)hipFree(d);
    if (e != hipSuccess) return fail(EXC_EHIP, "exc_batch_seed: %s", hipGetErrorString(e));
    return EXC_OK;
}

int exc_batch_colpend(void* p, uint32_t* out) {
    Ctx* c = as_ctx(p);
    if (!c || !c->K || !out) return fail(EXC_EINVAL, "exc_batch_colpend: not a batched context");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipMemcpy(out, c->d_colpend, (size_t)c->K * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return EXC_OK;
}

int exc_batch_set_active(void* p, uint64_t mask) {
    Ctx* c = as_ctx(p);
    if (!c || !c->K) return fail(EXC_EINVAL, "exc_batch_set_active: not a batched context");
    const uint64_t m = c->K == 64 ? mask : (mask & ((1ull << c->K) - 1));
    if (m & ~c->active) c->front_ok = false;   // a column back on may hold pending vertices nobody pushed
    c->active = m;
    return EXC_OK;
}

int exc_set_edge_weights(void* p, const float* h_w, uint64_t cnt) {
    Ctx* c = as_ctx(p);
    if (!c || c->K || !h_w) return fail(EXC_EINVAL, "exc_set_edge_weights: bad arguments");
    // the weighted dense pull exists for one measure only; with more, the pull would sum the
    // unweighted rows while the push scales them
    if (c->nm != 1) return fail(EXC_EINVAL, "exc_set_edge_weights: weighted graphs take exactly one measure");
    if (!c->d_arena || c->arena_edges != c->m || cnt != c->m)
        return fail(EXC_ESTATE, "exc_set_edge_weights: the whole graph must be in the edge cache and cnt == m");
    if (c->d_tadj) return fail(EXC_ESTATE, "exc_set_edge_weights: set the weights before exc_enable_pull");
    HIPCHK(hipSetDevice(c->dev));
    c->snap_ready = false;   // the state changes: a fused snapshot no longer holds
    c->front_ok = false;     // ... nor may the next batched snapshot read the touched rows only
    c->l1_valid = false;
    c->folded = false;       // the state is replaced: a folded pass no longer applies
    set_active(c, c->nm);    // ... and every measure may be pending again
    burst_graphs_clear(c);   // recorded pushes ran unweighted
    if (!c->d_wts) HIPCHK(hipMalloc(&c->d_wts, (c->m ? c->m : 1) * sizeof(float)));
    if (c->m) HIPCHK(hipMemcpy(c->d_wts, h_w, c->m * sizeof(float), hipMemcpyHostToDevice));
    return EXC_OK;
}

int exc_burst_submit(void* p, int passes, int dense) {
    Ctx* c = as_ctx(p);
    if (!c || passes < 1 || passes > EXC_BURST_MAX)
        return fail(EXC_EINVAL, "exc_burst_submit: 1 to %d passes", EXC_BURST_MAX);
    if (c->K) return fail(EXC_ESTATE, "exc_burst_submit: plain contexts only");
    if (!c->d_arena || c->arena_edges != c->m) return fail(EXC_ESTATE, "exc_burst_submit: the whole graph must be in the edge cache");
    if (c->d_trace) return fail(EXC_ESTATE, "exc_burst_submit: not with the pending-pass trace");
    if (dense && !(c->d_tadj && (!c->d_wts || c->d_twts))) return fail(EXC_ESTATE, "exc_burst_submit: dense passes need exc_enable_pull");
    if (!dense && c->folded)
        return fail(EXC_ESTATE, "exc_burst_submit: the last dense pass folded the next one; it must be a dense pass");
    if (c->burst_subs - c->burst_cols >= 2) return fail(EXC_ESTATE, "exc_burst_submit: two bursts are already in flight");
    HIPCHK(hipSetDevice(c->dev));
    const size_t W = cnt_words(c);
    if (!c->d_burst) {
        HIPCHK(hipMalloc(&c->d_burst, 2 * EXC_BURST_MAX * W * sizeof(unsigned long long)));
        HIPCHK(hipHostMalloc((void**)&c->h_burst, 2 * EXC_BURST_MAX * W * sizeof(unsigned long long), hipHostMallocDefault));
        for (auto& e : c->burst_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const int slot = c->burst_subs & 1;
    int rc = !dense && c->use_burst_graphs ? launch_push_burst_graph(c, passes, slot) : enqueue_burst(c, passes, dense, slot);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->burst_ev[slot], c->st[0]));
    c->burst_len[slot] = passes;
    c->burst_subs++;
    return EXC_OK;
}

int exc_burst_collect(void* p, uint64_t* h_cnt, double* h_l1, float* h_linf, int* passes) {
    Ctx* c = as_ctx(p);
    if (!c || !h_cnt) return fail(EXC_EINVAL, "exc_burst_collect: bad arguments");
    if (c->burst_cols == c->burst_subs) return fail(EXC_ESTATE, "exc_burst_collect: no burst in flight");
    HIPCHK(hipSetDevice(c->dev));
    const int slot = c->burst_cols & 1;
    HIPCHK(hipEventSynchronize(c->burst_ev[slot]));
    c->burst_cols++;
    const bool last = c->burst_cols == c->burst_subs;
    if (last) {   // nothing else in flight: the context is idle again
        for (auto s2 : c->st) HIPCHK(hipStreamSynchronize(s2));
        c->counts_enqueued = false;
        for (auto& sl : c->slot) sl.busy = false;
        HIPCHK(hipGetLastError());
        int rc = collect_timers(c);   // (their events may belong to a burst still running otherwise)
        if (rc) return rc;
    }
    const size_t W = cnt_words(c);
    const int np = c->burst_len[slot];
    const unsigned long long* h_reg = c->h_burst + (size_t)slot * EXC_BURST_MAX * W;
    for (int i = 0; i < np; i++) {
        memcpy(c->h_small, h_reg + i * W, W * sizeof(unsigned long long));
        parse_counts(c, h_cnt + (size_t)i * 2 * c->nc, nullptr);
        for (int k = 0; k < c->nm; k++) {
            if (h_l1) h_l1[(size_t)i * c->nm + k] = c->l1_last[k];
            if (h_linf) h_linf[(size_t)i * c->nm + k] = c->linf_last[k];
        }
    }
    if (passes) *passes = np;
    // with nothing left in flight, the next exc_snapshot returns the last pass's snapshot
    // without a launch
    c->snap_ready = last;
    c->counts_copied = last;
    return EXC_OK;
}

int exc_run_passes(void* p, int passes, int dense, uint64_t* h_cnt, double* h_l1, float* h_linf) {
    Ctx* c = as_ctx(p);
    if (!c || passes < 1 || passes > EXC_BURST_MAX || !h_cnt)
        return fail(EXC_EINVAL, "exc_run_passes: 1 to %d passes and a count array", EXC_BURST_MAX);
    if (c->burst_subs != c->burst_cols) return fail(EXC_ESTATE, "exc_run_passes: a submitted burst is not collected");
    int rc = exc_burst_submit(p, passes, dense);
    if (rc) {   // the same checks, reported under this name
        std::string m = exc_last_error();
        const std::string from = "exc_burst_submit";
        if (m.compare(0, from.size(), from) == 0) m = "exc_run_passes" + m.substr(from.size());
        return fail(rc, "%s", m.c_str());
    }
    return exc_burst_collect(p, h_cnt, h_l1, h_linf, nullptr);
}

int exc_pass_sync(void* p, double* waited) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_pass_sync: null context");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    auto a = std::chrono::steady_clock::now();
    for (auto s : c->st) HIPCHK(hipStreamSynchronize(s));
    if (waited) *waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    if (c->counts_enqueued) {   // the copy queued behind the last dense pull is complete
        c->counts_copied = c->snap_ready;
        c->counts_enqueued = false;
    }
    for (auto& sl : c->slot) sl.busy = false;
    HIPCHK(hipGetLastError());
    return collect_timers(c);
}

int exc_pushed(void* p, uint64_t* edges) {
    Ctx* c = as_ctx(p);
    if (!c || !edges) return fail(EXC_EINVAL, "exc_pushed: bad arguments");
    HIPCHK(hipSetDevice(c->dev));
    unsigned long long h[C_NCOUNTERS];
    HIPCHK(hipMemcpy(h, c->d_counters, sizeof h, hipMemcpyDeviceToHost));
    uint64_t e = 0;
    for (int s = 0; s < 6; s++) e += h[s * C_FIELDS + C_EDGES] + (s < 4 ? h[s * C_FIELDS + C_PEDGES] : 0ull);
    for (int q = 0; q < SPREAD_SLOTS; q++) e += h[C_SPREAD + q * SPREAD_STRIDE + S_EDGES];
    *edges = e;
    if (h[C_OVERFLOW]) return fail(EXC_ESTATE, "hub piece buffer overflowed %llu times", h[C_OVERFLOW]);
    return EXC_OK;
}

int exc_get_stats(void* p, exc_stats* o) {
    Ctx* c = as_ctx(p);
    if (!c || !o) return fail(EXC_EINVAL, "exc_get_stats: bad arguments");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    unsigned long long h[C_NCOUNTERS];
    HIPCHK(hipMemcpy(h, c->d_counters, sizeof h, hipMemcpyDeviceToHost));
    memset(o, 0, sizeof *o);
    o->snapshot_ms = c->t_ms[T_SNAP]; o->snapshot_launches = c->t_launch[T_SNAP]; o->snapshot_bytes = c->snap_bytes;
    o->piece_ms = c->t_ms[T_PIECE]; o->piece_launches = c->t_launch[T_PIECE];
    o->other_ms = c->t_ms[T_OTHER]; o->other_launches = c->t_launch[T_OTHER];
    for (int s = 0; s < 4; s++) {
        const unsigned long long* r = h + s * C_FIELDS;
        // algorithmic bytes: per pushed vertex 16 (offset pair) + 16 per measure (r and x
        // read + write); per pushed edge 4 (column id) + 8 per measure (r read + write)
        const double vb = 16.0 * r[C_VERT] + 16.0 * r[C_VERT_NM] + 4.0 * r[C_EDGES] + 8.0 * r[C_EDGES_NM];
        const double pb = 4.0 * r[C_PEDGES] + 8.0 * r[C_PEDGES_NM];
        o->src_ms[s] = c->t_ms[T_PUSH + s];
        o->src_launches[s] = c->t_launch[T_PUSH + s];
        o->src_bytes[s] = vb;
        o->push_ms += c->t_ms[T_PUSH + s];
        o->push_launches += c->t_launch[T_PUSH + s];
        o->push_bytes += vb;
        o->piece_bytes += pb;
        o->pushed_vertices += r[C_VERT];
        o->pushed_vertex_measures += r[C_VERT_NM];
        o->pushed_edges += r[C_EDGES] + r[C_PEDGES];
        o->pushed_edge_measures += r[C_EDGES_NM] + r[C_PEDGES_NM];
    }
    {   // dense passes: fold (rows 4 and 5) and pull
        const unsigned long long* r4 = h + C_DENSE * C_FIELDS;
        const unsigned long long* r5 = h + C_DENSE_LIVE * C_FIELDS;
        const int nm = c->K ? c->K : c->nm;
        const double pb = c->K && c->brows16 ? 2.0 : 4.0;   // bytes per column of a p row
        const double P = (double)c->pull_passes, PL = (double)c->pull_passes_live;
        o->fold_ms = c->t_ms[T_FOLD];
        o->pull_ms = c->t_ms[T_PULL] + c->t_ms[T_PULL_A1] + c->t_ms[T_PULL_A2];
        o->pull_hub_ms = c->t_ms[T_PULL_HUB];
        o->pull_passes = c->pull_passes;
        o->pull_tile = c->pull_tile_w[c->nm - 1];
        // fold: read r for every vertex, write p for every source; offsets, x read + write for
        // folded ones
        o->fold_bytes = c->fold_fixed_bytes + 16.0 * (r4[C_VERT] + r5[C_VERT]) + 8.0 * (r4[C_VERT_NM] + r5[C_VERT_NM]);
        if (c->K)   // batched: the bitmap, then per pending vertex its offsets, p-row number, r row
                    // and p row, per folded column r written and x read + written; per pass the
                    // live-row bitmap from the row->vertex map
            o->fold_bytes = P * 4.0 * (((double)c->n + 31) / 32) + (double)r5[C_VERT] * (20.0 + (4.0 + pb) * nm) +
                            12.0 * (double)r5[C_VERT_NM] + P * (4.0 * c->nsrc + c->nsrc / 8.0);
        // pull: every in-edge's id (the sorted pull: row id and 16-bit destination offset) and
        // weight on weighted graphs,
        // every vertex's offsets and r (read + write), and the p rows
        // read: every in-edge's in a plain pass; in a pass with the live-row bitmap, the
        // bitmap plus one row per out-edge of a pushing source (the fold's edge count -- a
        // pushed weight is zero only on underflow), split between the kernels by
        // k_pull_hub's own count.  In-lists above the hub threshold go to the hub kernel.
        const double me = (double)(c->m - c->pull_hub_edges), nv = (double)(c->n - c->pull_hub_vertices);
        const double bm = 8.0 * ((c->nsrc + 63) / 64);
        unsigned long long sp[S_ROWS_A2 + 1] = {};
        for (int q = 0; q < SPREAD_SLOTS; q++)
            for (int f = 0; f <= S_ROWS_A2; f++) sp[f] += h[C_SPREAD + q * SPREAD_STRIDE + f];
        const double hub_rows = (P - PL) * c->pull_hub_edges + (double)r5[C_PULL_HUB_LIVE];
        // k_pull counts the rows it reads (batched passes count them through the fold)
        const double rows = c->K ? std::max(0.0, (double)(r5[C_EDGES] + sp[S_EDGES]) - (double)r5[C_PULL_HUB_LIVE])
                                 : (double)sp[S_ROWS];
        if (c->K) {
            o->pull_bytes = P * (me * 4.0 + nv * (8.0 + 8.0 * nm)) + PL * bm + pb * nm * rows;
            o->pull_hub_bytes = P * ((double)c->pull_hub_edges * 4.0 + (double)c->pull_hub_vertices * (8.0 + 8.0 * nm)) +
                                PL * bm + pb * nm * hub_rows;
            o->pull_w_ms[0] = o->pull_ms;
            o->pull_w_launches[0] = c->pull_passes;
            o->pull_w_bytes[0] = o->pull_bytes;
        } else {   // plain: per width (measures the pass updated), the rows counted by the kernels
            const int srow[3] = {S_ROWS_A1, S_ROWS_A2, S_ROWS};
            const int hrow[3] = {C_PULL_HUB_LIVE_A1, C_PULL_HUB_LIVE_A2, C_PULL_HUB_LIVE};
            const int tcls[3] = {T_PULL_A1, T_PULL_A2, T_PULL};
            for (int w = 1; w <= c->nm; w++) {
                const int i = w == c->nm ? 2 : w - 1;   // the context's full width counts in S_ROWS
                o->pull_w_ms[w - 1] = c->t_ms[tcls[i]];
                o->pull_w_launches[w - 1] = c->pull_launches_w[w - 1];
                o->pull_w_tile[w - 1] = c->pull_tile_w[w - 1];
                o->pull_w_bytes[w - 1] = c->pull_fixed_bytes[w - 1] + 4.0 * w * (double)sp[srow[i]];
                o->pull_bytes += o->pull_w_bytes[w - 1];
                o->pull_hub_bytes += c->hub_fixed_bytes[w - 1] + 4.0 * w * (double)r5[hrow[i]];
            }
        }
        for (const unsigned long long* r : {r4, r5}) {
            o->pushed_vertices += r[C_VERT];
            o->pushed_vertex_measures += r[C_VERT_NM];
            o->pushed_edges += r[C_EDGES];
            o->pushed_edge_measures += r[C_EDGES_NM];
        }
        // folds done inside k_pull for the following pass
        o->pushed_vertices += sp[S_VERT];
        o->pushed_vertex_measures += sp[S_VERT_NM];
        o->pushed_edges += sp[S_EDGES];
        o->pushed_edge_measures += sp[S_EDGES_NM];
    }
    o->pull_rows_in_flight = c->K ? c->bpull_q : 0;
    o->burst_graph_launches = c->burst_graph_launches;
    o->front_snapshots = c->front_snaps;
    for (int k = 0; k < 7; k++) {
        o->push_tune_best[k] = c->ptune[k].best;
        for (int i = 0; i < 3; i++) o->push_tune_ns_edge[k][i] = c->ptune[k].ns_edge[i];
    }
    o->pull_tune_ns_per_edge[0] = c->bq_ns_edge[0];
    o->pull_tune_ns_per_edge[1] = c->bq_ns_edge[1];
    if (h[C_OVERFLOW]) return fail(EXC_ESTATE, "hub piece buffer overflowed %llu times", h[C_OVERFLOW]);
    return EXC_OK;
}

int exc_enable_pull(void* p, double threshold) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_enable_pull: null context");
    if (!c->d_arena || c->arena_edges != c->m) return fail(EXC_ESTATE, "exc_enable_pull: the whole graph must be in the edge cache");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipDeviceSynchronize());
    c->pull_threshold = threshold;
    {   // EXC_PULL_LIVE=0 / 1: never / always use the live-row bitmap (measurements)
        const char* ev = getenv("EXC_PULL_LIVE");
        c->pull_live = ev && (ev[0] == '0' || ev[0] == '1') ? ev[0] - '0' : 2;
        if (const char* lf = getenv("EXC_PULL_LIVE_FRAC")) c->pull_live_frac = atof(lf);
        if (const char* lf = getenv("EXC_PULL_LIVE_FRAC_SORTED")) c->pull_live_frac_sorted = atof(lf);
        if (const char* ff = getenv("EXC_FUSE_FOLD")) c->fuse_fold = atoi(ff);   // 0: fold with k_fold always
        if (const char* hq = getenv("EXC_BHUB_Q")) c->bhub_q = atoi(hq) == 16 ? 16 : 8;
        if (const char* ps = getenv("EXC_PULL_SMALL")) c->pull_small = std::min(2, std::max(0, atoi(ps)));
        if (const char* bq = getenv("EXC_BPULL_Q")) {   // 8 / 16 rows per wave, no tuning
            c->bpull_q = atoi(bq) == 16 ? 16 : 8;
        } else if (c->K && !c->bq_ev[0] && c->bq_trial == 0) {
            HIPCHK(hipEventCreate(&c->bq_ev[0]));
            HIPCHK(hipEventCreate(&c->bq_ev[1]));
            c->bq_tune = true;
        }
        const char* sq = getenv("EXC_PULL_SEQ");   // 4 / 8 / 16 / 32 entries per lane (measurements)
        c->pull_seq = sq ? atoi(sq) : 8;
    }
    if (c->d_tadj) return EXC_OK;
    const size_t n = (size_t)c->n;
    // rows of p in decreasing out-degree order (ties by id), vertices without out-edges
    // last: the most-read rows share cache lines (plain) and stay in the caches (batched)
    const bool by_id = false;
    std::vector<uint64_t> hidx(n + 1);
    HIPCHK(hipMemcpy(hidx.data(), c->d_idx, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> rank(by_id ? 0 : n);
    if (!by_id) {
        uint64_t maxd = 0;
        for (size_t v = 0; v < n; v++) maxd = std::max(maxd, hidx[v + 1] - hidx[v]);
        std::vector<uint64_t> start(maxd + 2, 0);
        for (size_t v = 0; v < n; v++) start[maxd - (hidx[v + 1] - hidx[v]) + 1]++;
        for (uint64_t d = 0; d <= maxd; d++) start[d + 1] += start[d];
        for (size_t v = 0; v < n; v++) rank[v] = (int32_t)start[maxd - (hidx[v + 1] - hidx[v])]++;
    }
    unsigned long long* cnt = nullptr;
    bool ok = hipMalloc(&c->d_tidx, (n + 1) * sizeof(uint64_t)) == hipSuccess;
    ok = ok && (by_id || hipMalloc(&c->d_rank, n * sizeof(int32_t)) == hipSuccess);
    c->nsrc = 0;
    for (size_t v = 0; v < n; v++) c->nsrc += hidx[v + 1] > hidx[v];
    ok = ok && hipMalloc(&c->d_pbits, (2 * (((size_t)c->nsrc + 63) / 64) + 2) * sizeof(uint32_t)) == hipSuccess;
    ok = ok && hipMalloc(&c->d_tadj, (c->m ? c->m : 1) * sizeof(int32_t)) == hipSuccess;
    if (ok && c->d_wts) ok = hipMalloc(&c->d_twts, (c->m ? c->m : 1) * sizeof(float)) == hipSuccess;
    if (c->K) {   // EXC_BATCH_ROWS=f32: float weight rows in batched dense passes (measurements)
        const char* br = getenv("EXC_BATCH_ROWS");
        c->brows16 = !(br && !strcmp(br, "f32"));
    }
    const bool b16 = c->K && c->brows16;
    const size_t prow_bytes = (size_t)(c->K ? prow_stride(c->K, b16) : c->S) * (b16 ? sizeof(uint16_t) : sizeof(float));
    ok = ok && hipMalloc(&c->d_p, n * prow_bytes) == hipSuccess;
    ok = ok && hipMemset(c->d_p, 0, n * prow_bytes) == hipSuccess;
    if (ok && c->fuse_fold) ok = hipMalloc(&c->d_p2, n * prow_bytes) == hipSuccess;
    if (ok && c->fuse_fold) ok = hipMemset(c->d_p2, 0, n * prow_bytes) == hipSuccess;
    if (ok && c->K) ok = hipMalloc(&c->d_inv, std::max<size_t>(1, (size_t)c->nsrc) * sizeof(int32_t)) == hipSuccess;
    if (ok && c->K) ok = hipMalloc(&c->d_tile_next, sizeof(unsigned long long)) == hipSuccess;
    if (ok && c->K && c->fuse_fold) {
        ok = hipMalloc(&c->d_bitmap2, ((n + 63) / 64) * 2 * sizeof(uint32_t)) == hipSuccess;
        ok = ok && hipMalloc(&c->d_hubbits, ((n + 63) / 64) * 2 * sizeof(uint32_t)) == hipSuccess;
    }
    ok = ok && hipMalloc(&cnt, (n + 1) * sizeof(unsigned long long)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        void* q[] = {c->d_tidx, c->d_tadj, c->d_p, cnt, c->d_rank, c->d_pbits, c->d_p2, c->d_bitmap2, c->d_hubbits, c->d_twts,
                     c->d_inv, c->d_tile_next};
        for (void* x : q) if (x) (void)hipFree(x);
        c->d_tidx = nullptr; c->d_tadj = nullptr; c->d_p = nullptr; c->d_rank = nullptr; c->d_pbits = nullptr;
        c->d_inv = nullptr;
        c->d_tile_next = nullptr;
        c->d_p2 = nullptr; c->d_bitmap2 = nullptr; c->d_hubbits = nullptr; c->d_twts = nullptr;
        c->nsrc = 0;
        c->pull_threshold = -1;
        return fail(EXC_ENOMEM, "exc_enable_pull: cannot allocate the transpose (%.2f GB)", c->m * 4.0 / 1e9);
    }
    auto fail_free = [&](hipError_t e) {
        (void)hipFree(cnt);
        return fail(EXC_EHIP, "exc_enable_pull: %s", hipGetErrorString(e));
    };
    hipError_t e = by_id ? hipSuccess : hipMemcpy(c->d_rank, rank.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && c->d_inv) {
        std::vector<int32_t> inv(std::max<size_t>(1, (size_t)c->nsrc), 0);
        for (size_t v = 0; v < n; v++)
            if (rank[v] < c->nsrc) inv[rank[v]] = (int32_t)v;
        e = hipMemcpy(c->d_inv, inv.data(), inv.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(cnt, 0, (n + 1) * sizeof(unsigned long long));
    if (e != hipSuccess) return fail_free(e);
    hipLaunchKernelGGL(k_count_in64, dim3(c->cus * 16), dim3(BLOCK), 0, 0, c->m, c->d_arena, cnt);
    std::vector<uint64_t> h(n + 1);
    e = hipMemcpy(h.data(), cnt, n * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_free(e);
    // in-lists of at most PULL_HUB entries first, in vertex order (d_tidx; a hub's range there
    // is empty), then the hub lists (cut into pieces for k_pull_hub); cur = fill cursors
    std::vector<PullPiece> hubs;
    std::vector<uint64_t> off(n + 1), cur(n);
    uint64_t hub_thr = PULL_HUB;
    if (const char* ev = getenv("EXC_PULL_HUB")) hub_thr = std::max<uint64_t>(64, strtoull(ev, nullptr, 10));   // measurements
    uint64_t acc = 0;
    for (size_t v = 0; v < n; v++) {
        off[v] = cur[v] = acc;
        if (h[v] <= hub_thr) acc += h[v];
    }
    off[n] = acc;
    for (size_t v = 0; v < n; v++) {
        const uint64_t d = h[v];
        if (d <= hub_thr) continue;
        c->pull_hub_edges += d;
        c->pull_hub_vertices++;
        const uint64_t piece = c->K ? BPULL_PIECE : PULL_PIECE;
        for (uint64_t o = 0; o < d; o += piece)
            hubs.push_back({(int32_t)v, (uint32_t)std::min<uint64_t>(piece, d - o), acc + o});
        cur[v] = acc;
        acc += d;
    }
    e = hipMemcpy(c->d_tidx, off.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(cnt, cur.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_free(e);
    {
        const uint64_t ranges = (c->m + FILL_EDGES - 1) / FILL_EDGES;
        const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(ranges, (uint64_t)c->cus * 8));
        hipLaunchKernelGGL(k_fill_in64, dim3(grid), dim3(BLOCK), 0, 0, c->n, c->m, c->d_idx, c->d_arena, cnt, c->d_rank,
                           c->d_tadj, c->d_wts, c->d_twts);
    }
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_free(e);
    (void)hipFree(cnt);
    // each hub's in-list sorted by source row: k_pull_hub's waves read 64 consecutive
    // entries, and sorted they share row lines (at R-MAT scale 24, 0.61 lines per entry at one
    // measure, 0.77 at three, from about one)
    if (!c->K && c->pull_hub_vertices && c->m - off[n] < (1ull << 31) && c->nsrc > 0) {
        std::vector<int> starts;
        starts.reserve(c->pull_hub_vertices + 1);
        for (size_t v = 0; v < n; v++)
            if (h[v] > hub_thr) starts.push_back((int)(cur[v] - off[n]));
        starts.push_back((int)(c->m - off[n]));
        int bits = 1;
        while (bits < 31 && (1ll << bits) < (long long)c->nsrc) bits++;
        e = excent_dev::sort_runs(c->d_tadj + off[n], c->d_twts ? c->d_twts + off[n] : nullptr, c->m - off[n], starts, bits, 0);
        if (e == hipErrorOutOfMemory) (void)hipGetLastError();   // keep the fill order
        else if (e != hipSuccess) return fail(EXC_EHIP, "exc_enable_pull: hub lists: %s", hipGetErrorString(e));
    }
    // the non-hub in-edges again, per tile of destinations sorted by source row
    // (k_pull_sorted), for the dense passes over one measure (a one-measure context, or the
    // first measure once the others are dropped).  With one measure it cuts the launch by a
    // sixth at R-MAT scale 24 (its L1-to-L2 requests drop by a third); over two or three
    // measures its LDS atomics per entry cost more than that saves, so k_pull runs those.
    // EXC_PULL_TILE = T (512 / 1024 / 2048 / 4096) picks the tile and the sorted pull at
    // every width, "T/w" at widths up to w only; 0 keeps k_pull.  Larger tiles share more
    // row lines per wave but fit fewer blocks per CU; 8192 and 16384 ran slower at R-MAT
    // scale 24 (PageRank alone 1.39 and 1.77 ms against 1.21 ms per launch).
    if (!c->K) {
        // the largest tile that still gives every CU four blocks; small graphs keep k_pull
        int T = 0;
        for (int t : {2048, 1024, 512})
            if ((long long)c->n / t >= 4ll * c->cus) { T = t; break; }
        if (const char* ev = getenv("EXC_PULL_TILE")) {
            T = atoi(ev);
            const char* sl = strchr(ev, '/');
            c->sorted_w = sl ? std::max(1, atoi(sl + 1)) : 3;
        }
        if ((T == 512 || T == 1024 || T == 2048 || T == 4096) && off[n] < (1ull << 31) && c->nsrc > 0) {
            int bits = 1;
            while (bits < 31 && (1ll << bits) < (long long)c->nsrc) bits++;
            const hipError_t se = excent_dev::build_sorted_tiles(c->n, T, c->d_tidx, c->d_tadj, c->d_twts, off[n], bits, 0,
                                                                 &c->sorted);
            if (se == hipSuccess) c->sorted_T = T;
            else if (se == hipErrorOutOfMemory) (void)hipGetLastError();   // keep k_pull
            else return fail(EXC_EHIP, "exc_enable_pull: sorted tiles: %s", hipGetErrorString(se));
        }
    }
    if (c->d_hubbits) {   // batched: which destinations k_bpull_hub completes
        std::vector<uint32_t> hb(((n + 63) / 64) * 2, 0);
        for (const PullPiece& pc : hubs) hb[pc.u >> 5] |= 1u << (pc.u & 31);
        HIPCHK(hipMemcpy(c->d_hubbits, hb.data(), hb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    c->n_pull_hubs = (int)hubs.size();
    if (c->n_pull_hubs) {
        HIPCHK(hipMalloc(&c->d_pull_hubs, hubs.size() * sizeof(PullPiece)));
        HIPCHK(hipMemcpy(c->d_pull_hubs, hubs.data(), hubs.size() * sizeof(PullPiece), hipMemcpyHostToDevice));
    }
    return EXC_OK;
}

int exc_set_hub_destinations(void* p, const int32_t* ids, int count) {
    Ctx* c = as_ctx(p);
    if (!c || count < 0 || count > HUB_MAX || (count && !ids))
        return fail(EXC_EINVAL, "exc_set_hub_destinations: 0 to %d vertex ids", HUB_MAX);
    std::vector<int32_t> keys(HUB_SLOTS, -1);
    std::vector<uint32_t> filter(HUB_FILTER_WORDS, 0u);
    int probes = 0, kept = 0;
    for (int i = 0; i < count; i++) {
        const int32_t u = ids[i];
        if (u < 0 || u >= c->n) return fail(EXC_EINVAL, "exc_set_hub_destinations: vertex %d out of range", u);
        uint32_t h = hub_slot((uint32_t)u);
        int q = 1;
        while (keys[h] >= 0 && keys[h] != u) { h = (h + 1) & (HUB_SLOTS - 1); q++; }
        if (keys[h] == u) continue;   // listed twice
        keys[h] = u;
        kept++;
        probes = std::max(probes, q);
        const uint32_t fb = hub_fbit((uint32_t)u);
        filter[fb >> 5] |= 1u << (fb & 31);
    }
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipDeviceSynchronize());   // no push in flight reads the table being replaced
    burst_graphs_clear(c);            // recorded pushes carry the table's size and probe length
    if (!c->d_hub_keys) HIPCHK(hipMalloc(&c->d_hub_keys, HUB_SLOTS * sizeof(int32_t)));
    if (!c->d_hub_filter) HIPCHK(hipMalloc(&c->d_hub_filter, HUB_FILTER_WORDS * sizeof(uint32_t)));
    HIPCHK(hipMemcpy(c->d_hub_keys, keys.data(), HUB_SLOTS * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_hub_filter, filter.data(), HUB_FILTER_WORDS * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->hub_probes = probes;
    c->hub_count = kept;
    return EXC_OK;
}

int exc_residual_l1(void* p, double* out) {
    Ctx* c = as_ctx(p);
    if (!c || !out) return fail(EXC_EINVAL, "exc_residual_l1: bad arguments");
    if (c->K) return fail(EXC_ESTATE, "exc_residual_l1: not kept for batched contexts");
    if (!c->l1_valid) return fail(EXC_ESTATE, "exc_residual_l1: no snapshot yet");
    for (int k = 0; k < c->nm; k++) out[k] = c->l1_last[k];
    return EXC_OK;
}

int exc_residual_max(void* p, float* out) {
    Ctx* c = as_ctx(p);
    if (!c || !out) return fail(EXC_EINVAL, "exc_residual_max: bad arguments");
    if (c->K) return fail(EXC_ESTATE, "exc_residual_max: not kept for batched contexts");
    if (!c->l1_valid) return fail(EXC_ESTATE, "exc_residual_max: no snapshot yet");
    for (int k = 0; k < c->nm; k++) out[k] = c->linf_last[k];
    return EXC_OK;
}

int exc_set_active_measures(void* p, int count) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_set_active_measures: null context");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    if (c->K) return fail(EXC_ESTATE, "exc_set_active_measures: batched contexts drop columns (exc_batch_set_active)");
    if (count < 1 || count > c->nm) return fail(EXC_EINVAL, "exc_set_active_measures: count must be 1..%d", c->nm);
    if (count > c->nm_act && c->folded)
        return fail(EXC_ESTATE, "exc_set_active_measures: the last dense pass folded the next one over %d measures",
                    c->nm_act);
    if (count < c->nm_act) {
        if (!c->l1_valid) return fail(EXC_ESTATE, "exc_set_active_measures: no snapshot yet");
        // a folded pass holds the deltas of every slot that had a vertex above eps: narrowing
        // then would lose them
        if (c->folded)
            for (int k = count; k < c->nm_act; k++)
                if (c->linf_last[k] > c->M.eps[k])
                    return fail(EXC_ESTATE, "exc_set_active_measures: the last dense pass folded slot %d (|r| %g > eps %g)",
                                k, (double)c->linf_last[k], (double)c->M.eps[k]);
    }
    set_active(c, count);
    return EXC_OK;
}

int exc_fold_ahead(void* p, int allow) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_fold_ahead: null context");
    c->fold_ahead_ok = allow != 0;
    return EXC_OK;
}

int exc_dense_required(void* p) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_dense_required: null context");
    return c->folded ? 1 : 0;
}

int exc_reset_stats(void* p) {
    Ctx* c = as_ctx(p);
    if (!c) return fail(EXC_EINVAL, "exc_reset_stats: null context");
    if (c->burst_subs != c->burst_cols) return burst_busy(__func__);
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipDeviceSynchronize());
    int rc = collect_timers(c);
    if (rc) return rc;
    HIPCHK(hipMemset(c->d_counters, 0, C_NCOUNTERS * sizeof(unsigned long long)));
    HIPCHK(hipDeviceSynchronize());   // order the clear before work on the non-blocking streams
    for (int i = 0; i < T_NCLS; i++) { c->t_ms[i] = 0; c->t_launch[i] = 0; }
    c->snap_bytes = 0;
    c->pull_passes = 0;
    c->pull_passes_live = 0;
    c->fold_fixed_bytes = 0;
    for (int w = 0; w < 3; w++) {
        c->pull_fixed_bytes[w] = c->hub_fixed_bytes[w] = 0;
        c->pull_launches_w[w] = 0;
    }
    c->front_snaps = 0;
    return EXC_OK;
}

int exc_snapshot_noalloc(int device, void* stream, int n, int nm, const float* d_r, const float* eps,
                         const uint64_t* d_idx, const int* d_cv0, int nc, const uint32_t* d_hot, uint32_t* d_bitmap,
                         unsigned long long* d_counts, double* d_l1) {
    if (n < 1 || nm < 1 || nm > EXC_MAXM || !d_r || !eps || !d_idx || !d_cv0 || !d_bitmap || !d_counts || !d_l1)
        return fail(EXC_EINVAL, "exc_snapshot_noalloc: bad arguments (n >= 1, 1 <= nm <= %d, non-null buffers)", EXC_MAXM);
    if (nc < 1 || nc > NOALLOC_MAX_CHUNKS)
        return fail(EXC_EINVAL, "exc_snapshot_noalloc: %d chunks; its scratch holds the counts of 1 to %d chunks", nc,
                    NOALLOC_MAX_CHUNKS);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= NOALLOC_MAX_DEVICES)
        return fail(EXC_EINVAL, "exc_snapshot_noalloc: no device %d", device);
    HIPCHK(hipSetDevice(device));
    NoallocDev d;
    int rc = noalloc_setup(device, &d);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // clear the counts and the residual slots (a fill kernel, so a captured call replays it)
    const int words = (int)NOALLOC_WORDS;
    hipLaunchKernelGGL(k_zero_u64, dim3((words + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, d.ptr, words);
    SnapArgs a{};
    a.n = n;
    a.r = d_r;
    for (int k = 0; k < EXC_MAXM; k++) a.eps[k] = k < nm ? eps[k] : 0.f;
    a.idx = d_idx;
    a.cv0 = d_cv0;
    a.nc = nc;
    a.hot = d_hot;
    a.bitmap = d_bitmap;
    a.cnt = d.ptr;
    a.hotpend = d.ptr + 2 * nc;
    a.trace = nullptr;
    a.l1 = reinterpret_cast<double*>(d.ptr + NOALLOC_L1_OFF);
    launch_snapshot_plain(a, nm, d.cus, st);
    hipLaunchKernelGGL(k_snapshot_finish, dim3((2 * nc + 1 + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, d.ptr, nc, nm,
                       d_counts, d_l1);
    HIPCHK(hipGetLastError());
    return EXC_OK;
}

int exc_indegree(void* p, uint32_t* h_out) {
    Ctx* c = as_ctx(p);
    if (!c || !h_out) return fail(EXC_EINVAL, "exc_indegree: bad arguments");
    if (!c->d_arena || c->arena_edges != c->m) return fail(EXC_ESTATE, "exc_indegree: the whole graph must be in the edge cache");
    HIPCHK(hipSetDevice(c->dev));
    uint32_t* d;
    HIPCHK(hipMalloc(&d, (size_t)c->n * sizeof(uint32_t)));
    hipMemsetAsync(d, 0, (size_t)c->n * sizeof(uint32_t), c->st[0]);
    hipLaunchKernelGGL(k_indegree, dim3(c->cus * 16), dim3(BLOCK), 0, c->st[0], c->m, c->d_arena, d);
    hipError_t e = hipMemcpyAsync(h_out, d, (size_t)c->n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st[0]);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st[0]);
    hipFree(d);
    if (e != hipSuccess) return fail(EXC_EHIP, "exc_indegree: %s", hipGetErrorString(e));
    return EXC_OK;
}

}  // extern "C"
