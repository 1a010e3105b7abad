// planner.cpp -- see planner.hpp.
#include "planner.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>

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

// Hot set: cost(d) = 4d + 16 bytes.  Degrees >= 2^20 are sorted explicitly, smaller ones
// go through a histogram; the largest degrees are taken until the budget runs out,
// giving a threshold T and a quota of degree-T vertices taken in vertex order.  With
// scores, whole score classes from 65535 down are taken while they fit, then the first
// class that does not fit is taken in vertex order while its remainder lasts.
HotSet hot_select(int n, const uint64_t* off, uint64_t budget, const uint16_t* score, int threads) {
    auto cost = [](uint64_t d) { return d * sizeof(int32_t) + 2 * sizeof(int32_t) + sizeof(uint64_t); };
    const uint64_t CAP = 1ull << 20;
    std::vector<uint64_t> hist(CAP, 0), big;
#pragma omp parallel num_threads(threads)
    {
        std::vector<uint64_t> lh(CAP, 0), lb;
#pragma omp for schedule(static) nowait
        for (int v = 0; v < n; v++) {
            const uint64_t d = off[v + 1] - off[v];
            if (d >= CAP) lb.push_back(d); else lh[d]++;
        }
#pragma omp critical
        {
            for (uint64_t d = 0; d < CAP; d++) hist[d] += lh[d];
            big.insert(big.end(), lb.begin(), lb.end());
        }
    }
    std::sort(big.rbegin(), big.rend());
    uint64_t used = 0, T = 0, quota = 0;
    bool done = false;
    for (uint64_t d : big) {
        if (used + cost(d) > budget) { T = d; quota = 0; done = true; break; }
        used += cost(d); T = d; quota++;
    }
    if (!done) { quota = 0; T = CAP - 1; }
    for (uint64_t d = CAP - 1; !done && d >= 1; d--) {
        if (!hist[d]) continue;
        const uint64_t take = std::min<uint64_t>(hist[d], (budget - used) / cost(d));
        used += take * cost(d);
        T = d; quota = take;
        if (take < hist[d]) done = true;
    }
    int S = 0;
    uint64_t s_left = 0;
    if (score) {
        std::vector<double> cs(65536, 0.0);
        for (int v = 0; v < n; v++) {
            const uint64_t d = off[v + 1] - off[v];
            if (d && score[v]) cs[score[v]] += (double)cost(d);
        }
        double u = 0;
        for (int s = 65535; s >= 1; s--) {
            if (u + cs[s] <= (double)budget) { u += cs[s]; continue; }
            S = s; s_left = budget - (uint64_t)u; break;
        }
    }
    HotSet h;
    h.bitmask.assign(((size_t)n + 31) / 32, 0u);
    for (int v = 0; v < n; v++) {
        const uint64_t d = off[v + 1] - off[v];
        if (d == 0) continue;
        if (score) {
            if (score[v] == 0 || score[v] < S) continue;
            if (score[v] == S) { if (cost(d) > s_left) continue; s_left -= cost(d); }
        } else {
            const bool take = d > T || (d == T && quota > 0);
            if (!take) continue;
            if (d == T) quota--;
        }
        h.ids.push_back(v);
        h.off.push_back(h.off.back() + d);
        h.bitmask[v >> 5] |= 1u << (v & 31);
    }
    h.deg_min = score ? 0 : T;
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
