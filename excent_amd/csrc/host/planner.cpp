// planner.cpp -- see planner.hpp.
#include "planner.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <utility>

namespace excent {

uint64_t chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges) {
    if (chunk_medges > 0) return (uint64_t)(chunk_medges * 1e6);
    const uint64_t lo = (uint64_t)(min_chunk_medges * 1e6), hi = 64000000ull;
    return std::min(hi, std::max(lo, (uint64_t)std::ceil((double)m / target_chunks)));
}

int pick_group(double avg, int cap) {
    int g = 1;
    while (g < cap && g < avg) g <<= 1;
    return g;
}

// Per chunk: nothing pending -> skip; cached and filled -> resident; cached but not yet
// filled, or demand/edges >= tau -> stream; pass 0 of hot mode with hot lists to fill ->
// stream; otherwise gather (which needs the bitmap on the host).
PlanOut fetch_plan(int nc, const uint64_t* cnt, const std::vector<char>& resident, const std::vector<char>& loaded,
                   double tau, const std::vector<uint64_t>& chunk_edges, bool hot_mode, int pass,
                   const std::vector<int>& hlo, const std::vector<int>& hhi, int8_t* mode) {
    PlanOut o;
    for (int c = 0; c < nc; c++) {
        const uint64_t cv = cnt[2 * c], ce = cnt[2 * c + 1];
        o.pend_v += cv;
        o.demand += ce;
        mode[c] = F_SKIP;
        if (cv == 0) continue;
        if (resident[c] && loaded[c]) mode[c] = F_RESIDENT;
        else if (resident[c] || (double)ce >= tau * (double)chunk_edges[c]) mode[c] = F_STREAM;
        else if (hot_mode && pass == 0 && hhi[c] > hlo[c]) mode[c] = F_STREAM;
        else { mode[c] = F_GATHER; o.need_bitmap = true; }
    }
    return o;
}

// ---------------------------------------------------------------- hot-set selection
//
// The rule (src/excent.cu:432-491), with cost(d) = 4 d + 16 bytes per list:
//  * degree rule: the vertices with d >= 1 form classes of equal degree, taken whole in
//    decreasing degree while the class fits in what is left of the budget.  At the first
//    class that does not fit (degree T), a class of degree < 2^20 gives its first
//    floor(left / cost(T)) vertices in vertex order; a class of degree >= 2^20 gives none.
//    When every class fits, T is the smallest degree below 2^20 (2^20 - 1 if there is none)
//    and all of its vertices are taken.  deg_min reports T.
//  * score rule (--hot-rank): classes of equal score s >= 1 (lists with d >= 1) are taken
//    whole from s = 65535 down while they fit; the first that does not fit (score S) is
//    then offered in vertex order, each list taken if it fits in what is left.
//
// Finding T: a weighted radix select.  Pass 1 sums count and cost per bit-length bucket of
// the degree (64 buckets); the buckets above the crossing one fit whole.  Pass 2 resolves
// exact degree classes only inside the crossing bucket.  The selection is a parallel
// compaction over blocks of 2^16 vertices (counts, block offsets, then the writes), with
// the quota of class T handed to blocks in vertex order.
namespace {

inline uint64_t list_cost(uint64_t d) { return d * sizeof(int32_t) + 2 * sizeof(int32_t) + sizeof(uint64_t); }
inline int bit_length(uint64_t d) { return 64 - __builtin_clzll(d); }   // d >= 1
constexpr uint64_t kBigDegree = 1ull << 20;   // classes at or above it are never split
constexpr int kSelBlock = 1 << 16;            // vertices per compaction block (a multiple of 32)

struct DegreeCut {
    uint64_t T = kBigDegree - 1;   // threshold degree
    uint64_t quota = 0;            // vertices of degree T taken, in vertex order
};

DegreeCut degree_cut(int n, const uint64_t* off, uint64_t budget, int threads) {
    // pass 1: count and cost per bit-length bucket
    uint64_t bcnt[65] = {0}, bcost[65] = {0};
#pragma omp parallel num_threads(threads)
    {
        uint64_t lc[65] = {0}, lw[65] = {0};
#pragma omp for schedule(static) nowait
        for (int v = 0; v < n; v++) {
            const uint64_t d = off[v + 1] - off[v];
            if (!d) continue;
            const int b = bit_length(d);
            lc[b]++;
            lw[b] += list_cost(d);
        }
#pragma omp critical
        for (int b = 0; b <= 64; b++) { bcnt[b] += lc[b]; bcost[b] += lw[b]; }
    }
    uint64_t used = 0;
    int cross = 0;
    for (int b = 64; b >= 1; b--) {
        if (!bcnt[b]) continue;
        if (bcost[b] <= budget - used) { used += bcost[b]; continue; }
        cross = b;
        break;
    }
    DegreeCut cut;
    if (!cross) {   // everything fits: T is the smallest degree below 2^20
        uint64_t dmin = UINT64_MAX;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(min : dmin)
        for (int v = 0; v < n; v++) {
            const uint64_t d = off[v + 1] - off[v];
            if (d && d < dmin) dmin = d;
        }
        if (dmin >= kBigDegree) return cut;   // T = 2^20 - 1, quota 0: every list is above it
        uint64_t cnt = 0;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : cnt)
        for (int v = 0; v < n; v++) cnt += (off[v + 1] - off[v]) == dmin;
        cut.T = dmin;
        cut.quota = cnt;
        return cut;
    }
    // pass 2: exact classes of the crossing bucket [lo, 2 lo), in decreasing degree
    const uint64_t lo = 1ull << (cross - 1);
    std::vector<std::pair<uint64_t, uint64_t>> classes;   // (degree, count), decreasing
    if (lo >= kBigDegree) {   // few lists this long: collect and sort their degrees
        std::vector<uint64_t> degs;
#pragma omp parallel num_threads(threads)
        {
            std::vector<uint64_t> mine;
#pragma omp for schedule(static) nowait
            for (int v = 0; v < n; v++) {
                const uint64_t d = off[v + 1] - off[v];
                if (d && bit_length(d) == cross) mine.push_back(d);
            }
#pragma omp critical
            degs.insert(degs.end(), mine.begin(), mine.end());
        }
        std::sort(degs.begin(), degs.end(), std::greater<uint64_t>());
        for (uint64_t d : degs) {
            if (classes.empty() || classes.back().first != d) classes.push_back({d, 0});
            classes.back().second++;
        }
    } else {   // a dense count over the bucket's lo degrees
        const size_t width = (size_t)lo;
        std::vector<uint64_t> cnt(width, 0);
        if (width <= 65536) {   // narrow buckets hold many lists: counts per thread, then merged
#pragma omp parallel num_threads(threads)
            {
                std::vector<uint64_t> mine(width, 0);
#pragma omp for schedule(static) nowait
                for (int v = 0; v < n; v++) {
                    const uint64_t d = off[v + 1] - off[v];
                    if (d >= lo && d < 2 * lo) mine[d - lo]++;
                }
#pragma omp critical
                for (size_t i = 0; i < width; i++) cnt[i] += mine[i];
            }
        } else {   // wide buckets hold few lists per degree: shared counters
            uint64_t* c = cnt.data();
#pragma omp parallel for schedule(static) num_threads(threads)
            for (int v = 0; v < n; v++) {
                const uint64_t d = off[v + 1] - off[v];
                if (d >= lo && d < 2 * lo) __atomic_fetch_add(&c[d - lo], 1ull, __ATOMIC_RELAXED);
            }
        }
        for (size_t i = width; i-- > 0;)
            if (cnt[i]) classes.push_back({lo + i, cnt[i]});
    }
    for (const auto& [d, c] : classes) {
        const uint64_t need = c * list_cost(d);
        if (need <= budget - used) { used += need; continue; }
        cut.T = d;
        cut.quota = d >= kBigDegree ? 0 : std::min<uint64_t>(c, (budget - used) / list_cost(d));
        return cut;
    }
    return cut;   // not reached: the crossing bucket holds a class that does not fit
}

}  // namespace

HotSet hot_select(int n, const uint64_t* off, uint64_t budget, const uint16_t* score, int threads) {
    threads = std::max(1, threads);
    HotSet h;
    h.bitmask.assign(((size_t)n + 31) / 32, 0u);
    const int nblk = (int)(((long long)n + kSelBlock - 1) / kSelBlock);
    // per block: lists taken outright, their edges, and the candidates of the split class
    std::vector<uint64_t> take(nblk, 0), edges(nblk, 0), split_cnt(nblk, 0);
    DegreeCut cut;
    int S = 0;
    std::vector<std::vector<int>> split_taken(nblk);   // score rule: class-S lists taken, per block
    if (!score) {
        cut = degree_cut(n, off, budget, threads);
        h.deg_min = cut.T;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
        for (int b = 0; b < nblk; b++) {
            const int v0 = b * kSelBlock, v1 = std::min(n, v0 + kSelBlock);
            uint64_t t = 0, e = 0, eq = 0;
            for (int v = v0; v < v1; v++) {
                const uint64_t d = off[v + 1] - off[v];
                if (d > cut.T) { t++; e += d; }
                else if (d == cut.T && d) eq++;
            }
            take[b] = t; edges[b] = e; split_cnt[b] = eq;
        }
        uint64_t left = cut.quota;   // class T's quota, first come first served in vertex order
        for (int b = 0; b < nblk; b++) {
            const uint64_t q = std::min(left, split_cnt[b]);
            left -= q;
            split_cnt[b] = q;
            take[b] += q;
            edges[b] += q * cut.T;
        }
    } else {
        std::vector<uint64_t> cls(65536, 0);   // cost per score class
#pragma omp parallel num_threads(threads)
        {
            std::vector<uint64_t> mine(65536, 0);
#pragma omp for schedule(static) nowait
            for (int v = 0; v < n; v++) {
                const uint64_t d = off[v + 1] - off[v];
                if (d && score[v]) mine[score[v]] += list_cost(d);
            }
#pragma omp critical
            for (int s = 0; s < 65536; s++) cls[s] += mine[s];
        }
        // whole classes while they fit (the reference sums in double; integer sums are the
        // same numbers while they stay below 2^53 bytes)
        uint64_t used = 0, left = 0;
        for (int s = 65535; s >= 1; s--) {
            if (cls[s] <= budget - used) { used += cls[s]; continue; }
            S = s;
            left = budget - used;
            break;
        }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
        for (int b = 0; b < nblk; b++) {
            const int v0 = b * kSelBlock, v1 = std::min(n, v0 + kSelBlock);
            uint64_t t = 0, e = 0;
            for (int v = v0; v < v1; v++) {
                const uint64_t d = off[v + 1] - off[v];
                if (d && score[v] && score[v] > S) { t++; e += d; }
            }
            take[b] = t; edges[b] = e;
        }
        if (S) {   // class S, first fit in vertex order (sequential: each take shrinks `left`)
            for (int b = 0; b < nblk; b++) {
                const int v0 = b * kSelBlock, v1 = std::min(n, v0 + kSelBlock);
                for (int v = v0; v < v1; v++) {
                    const uint64_t d = off[v + 1] - off[v];
                    if (!d || score[v] != S || list_cost(d) > left) continue;
                    left -= list_cost(d);
                    split_taken[b].push_back(v);
                    take[b]++;
                    edges[b] += d;
                }
            }
        }
    }
    // block offsets, then every block writes its ids, list offsets and mask words
    std::vector<uint64_t> pos(nblk + 1, 0), ebase(nblk + 1, 0);
    for (int b = 0; b < nblk; b++) { pos[b + 1] = pos[b] + take[b]; ebase[b + 1] = ebase[b] + edges[b]; }
    h.ids.resize(pos[nblk]);
    h.off.resize(pos[nblk] + 1);
    h.off[0] = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int b = 0; b < nblk; b++) {
        const int v0 = b * kSelBlock, v1 = std::min(n, v0 + kSelBlock);
        uint64_t p = pos[b], e = ebase[b], q = score ? 0 : split_cnt[b];
        size_t next_split = 0;
        for (int v = v0; v < v1; v++) {
            const uint64_t d = off[v + 1] - off[v];
            if (!d) continue;
            bool in;
            if (score) {
                in = score[v] && score[v] > S;
                if (!in && score[v] == S && next_split < split_taken[b].size() && split_taken[b][next_split] == v) {
                    in = true;
                    next_split++;
                }
            } else {
                in = d > cut.T || (d == cut.T && q > 0 && (q--, true));
            }
            if (!in) continue;
            h.ids[p] = v;
            e += d;
            h.off[++p] = e;
            h.bitmask[v >> 5] |= 1u << (v & 31);
        }
    }
    if (score) h.deg_min = 0;
    return h;
}

// Break-even demand ratio: extrapolate from the end points when gathering is always
// slower (or always faster) than streaming, else interpolate linearly between the two
// calibration points around the crossing; clamp to [0.005, 1].
double tau_from_timings(double t_stream, const double* t_g, const double* f_got) {
    double s;
    if (t_g[0] >= t_stream) s = f_got[0] * t_stream / t_g[0];
    else if (t_g[2] < t_stream) s = f_got[2] * t_stream / t_g[2];
    else {
        const int j = t_g[1] >= t_stream ? 1 : 2;
        s = f_got[j - 1] + (f_got[j] - f_got[j - 1]) * (t_stream - t_g[j - 1]) / (t_g[j] - t_g[j - 1]);
    }
    return std::min(1.0, std::max(0.005, s));
}

int gather_list(const uint32_t* bitmap, int v0, int v1, int32_t* out) {
    int k = 0;
    for (int v = v0; v < v1;) {
        const uint32_t w = bitmap[v >> 5] >> (v & 31);
        if (!w) { v = (v | 31) + 1; continue; }
        v += __builtin_ctz(w);
        if (v >= v1) break;
        out[k++] = v;
        v++;
    }
    return k;
}

}  // namespace excent
