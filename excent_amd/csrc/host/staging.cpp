// staging.cpp -- see staging.hpp.
#include "staging.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "planner.hpp"

namespace excent {
namespace {

typedef std::chrono::steady_clock Clock;
double since(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

// SplitMix64 output function (Steele, Lea, Flood 2014), keyed by a counter.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

GatherStats stage_lists(const EgrGraph& g, const int32_t* vids, int k, uint64_t* off_out, int32_t* edges_out,
                        bool dedup, int threads) {
    GatherStats st;
    off_out[0] = 0;
    if (k <= 0) return st;
    threads = std::max(1, threads);
    // offsets: per-block degree sums, their prefix, then each block fills its own offsets
    const int nb = std::max(1, std::min(threads, k / 4096));
    std::vector<uint64_t> part(nb + 1, 0);
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; b++) {
        const int i0 = (int)((long long)k * b / nb), i1 = (int)((long long)k * (b + 1) / nb);
        uint64_t t = 0;
        for (int i = i0; i < i1; i++) t += g.degree(vids[i]);
        part[b + 1] = t;
    }
    for (int b = 0; b < nb; b++) part[b + 1] += part[b];
#pragma omp parallel for schedule(static, 1) num_threads(nb)
    for (int b = 0; b < nb; b++) {
        const int i0 = (int)((long long)k * b / nb), i1 = (int)((long long)k * (b + 1) / nb);
        uint64_t t = part[b];
        for (int i = i0; i < i1; i++) {
            t += g.degree(vids[i]);
            off_out[i + 1] = t;
        }
    }
    st.edges = off_out[k];
    if (dedup) g.prefetch_lists(vids, k, &st.pages, &st.runs);
    // columns: worker w copies columns [E w / W, E (w+1) / W) of the packed output
    const uint64_t E = st.edges;
    const int W = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)threads, E / 65536));
#pragma omp parallel for schedule(static, 1) num_threads(W)
    for (int w = 0; w < W; w++) {
        const uint64_t c0 = E * w / W, c1 = E * (w + 1) / W;
        if (c0 >= c1) continue;
        // the list holding column c0: the last i with off_out[i] <= c0
        int i = (int)(std::upper_bound(off_out, off_out + k + 1, c0) - off_out) - 1;
        for (uint64_t c = c0; c < c1; i++) {
            const uint64_t lo = off_out[i], hi = off_out[i + 1];
            if (hi <= c) continue;
            const uint64_t take = std::min(hi, c1) - c;
            memcpy(edges_out + c, g.edges(vids[i]) + (c - lo), take * sizeof(int32_t));
            c += take;
        }
    }
    return st;
}

TauProbe calibrate_tau(const EgrGraph& g, const EdgeChunk& ch, int32_t* h_edges, int32_t* h_vids, uint64_t* h_off,
                       bool dedup, int threads, const std::function<double(uint64_t)>& copy_probe) {
    TauProbe p;
    const auto t0 = Clock::now();
    const uint64_t E = ch.edges();
    const double chunk_bytes = (double)E * sizeof(int32_t);

    // stream: a cold read of the chunk, then its copy over the link
    g.drop_cache(ch.e0, ch.e1, ch.v0, ch.v1);
    auto a = Clock::now();
    g.read_edges(ch.e0, ch.e1, h_edges);
    const double t_read = since(a);
    const double t_link = copy_probe(E * sizeof(int32_t));
    p.storage_seq_gbs = t_read > 0 ? chunk_bytes / t_read / 1e9 : 0;
    p.link_gbs = t_link > 0 ? chunk_bytes / t_link / 1e9 : 0;
    p.t_stream = std::max(t_read, t_link);

    // gathers: lists drawn without replacement with probability proportional to their
    // length (exponential clocks: key = -ln(u) / d, smallest keys first), nested by demand
    const int nv = ch.v1 - ch.v0;
    std::vector<std::pair<double, int32_t>> order;
    order.reserve(nv);
    for (int v = ch.v0; v < ch.v1; v++) {
        const uint64_t d = g.degree(v);
        if (!d) continue;
        const double u = ((double)(mix64(0x7A0C1B5Eull + (uint64_t)v) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        order.push_back({-std::log(u) / (double)d, v});
    }
    std::sort(order.begin(), order.end());
    const double frac[3] = {0.05, 0.15, 0.4};
    size_t taken = 0;
    uint64_t acc = 0;
    std::vector<int32_t> pick;
    for (int j = 0; j < 3; j++) {
        while (taken < order.size() && (double)acc < frac[j] * (double)E) {
            acc += g.degree(order[taken].second);
            pick.push_back(order[taken].second);
            taken++;
        }
        std::vector<int32_t> sorted(pick);
        std::sort(sorted.begin(), sorted.end());
        const int k = (int)sorted.size();
        std::copy(sorted.begin(), sorted.end(), h_vids);
        g.drop_cache(ch.e0, ch.e1, ch.v0, ch.v1);
        a = Clock::now();
        const GatherStats gs = stage_lists(g, h_vids, k, h_off, h_edges, dedup, threads);
        const double t_host = since(a);
        // the slot ships the packed columns, the ids and the offsets
        const uint64_t bytes = gs.edges * sizeof(int32_t) + (uint64_t)k * (sizeof(int32_t) + sizeof(uint64_t));
        const double t_copy = copy_probe(std::min<uint64_t>(bytes, E * sizeof(int32_t)));
        p.t_gather[j] = std::max(t_host, t_copy);
        p.f_got[j] = E ? (double)gs.edges / (double)E : 0.0;
    }
    p.tau = tau_from_timings(p.t_stream, p.t_gather, p.f_got);
    p.seconds = since(t0);
    return p;
}

std::vector<int32_t> top_in_degree(const EgrGraph& g, int count, double sample, int threads) {
    const uint64_t m = g.m, block = (1u << 20) / sizeof(int32_t);   // columns per 1 MiB block
    const uint64_t blocks = (m + block - 1) / block;
    const uint64_t want = std::min<uint64_t>(m, std::max<uint64_t>(16ull << 20, (uint64_t)(sample * (double)m)));
    const uint64_t stride = std::max<uint64_t>(1, blocks * block / std::max<uint64_t>(want, 1));
    std::vector<uint32_t> cnt((size_t)g.n, 0);
    threads = std::max(1, threads);
    const long long nb = (long long)((blocks + stride - 1) / stride);
#pragma omp parallel num_threads(threads)
    {
        std::vector<int32_t> buf(block);
#pragma omp for schedule(dynamic, 1)
        for (long long i = 0; i < nb; i++) {
            const uint64_t e0 = (uint64_t)i * stride * block, e1 = std::min(m, e0 + block);
            g.read_edges(e0, e1, buf.data());
            for (uint64_t e = 0; e < e1 - e0; e++)
                if ((uint32_t)buf[e] < (uint32_t)g.n) __atomic_fetch_add(&cnt[buf[e]], 1u, __ATOMIC_RELAXED);
        }
    }
    // the largest counts: each thread keeps its best `count` in a min-heap, then they merge
    typedef std::pair<uint32_t, int32_t> CV;
    std::vector<CV> all;
#pragma omp parallel num_threads(threads)
    {
        std::vector<CV> heap;
        auto cmp = [](const CV& a, const CV& b) { return a.first > b.first; };
#pragma omp for schedule(static) nowait
        for (long long v = 0; v < (long long)g.n; v++) {
            if (!cnt[v]) continue;
            if ((int)heap.size() < count) {
                heap.push_back({cnt[v], (int32_t)v});
                std::push_heap(heap.begin(), heap.end(), cmp);
            } else if (cnt[v] > heap.front().first) {
                std::pop_heap(heap.begin(), heap.end(), cmp);
                heap.back() = {cnt[v], (int32_t)v};
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
        }
#pragma omp critical
        all.insert(all.end(), heap.begin(), heap.end());
    }
    std::sort(all.begin(), all.end(), [](const CV& a, const CV& b) { return a.first > b.first; });
    if ((int)all.size() > count) all.resize(count);
    std::vector<int32_t> out;
    out.reserve(all.size());
    for (const CV& c : all) out.push_back(c.second);
    return out;
}

}  // namespace excent
