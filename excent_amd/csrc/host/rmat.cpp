// rmat.cpp -- bin/excent_rmat: seeded R-MAT graphs written as EGR (no reference
// counterpart; the reference ships no generator).
//
// Edge i of the 2^scale * edge_factor generated edges picks `scale` quadrants with
// probabilities (a, b, c, 1-a-b-c) from a counter-based SplitMix64 stream keyed by
// (seed, i), so any subset of edges can be regenerated independently and in
// parallel.  Vertex ids go through a seeded bijection of [0, 2^scale) (multiply by an
// odd constant, add, xor-shift right; four rounds), the Graph500-style relabelling
// that spreads the hubs over the id range.  Self-loops and duplicate edges are
// removed and each list is sorted ascending.  Optional sidecars: the transpose as
// <out>.csc.egr and raw edge weights uniform(0.5, 1.5) as <out>.w.f32 (float32[m]).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "excent.h"

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {   // SplitMix64 output function (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rmat {
    int scale;
    uint64_t seed, mask;
    uint64_t ta, tab, tabc;
    uint64_t mul[4], add[4];
    int shift;
    bool permute;

    Rmat(int s, double a, double b, double c, uint64_t sd, bool perm) : scale(s), seed(sd), permute(perm) {
        mask = s >= 64 ? ~0ull : ((1ull << s) - 1);
        auto th = [](double p) { double t = std::floor(p * 4294967296.0); return (uint64_t)std::min(std::max(t, 0.0), 4294967296.0); };
        ta = th(a); tab = th(a + b); tabc = th(a + b + c);
        for (int r = 0; r < 4; r++) {
            mul[r] = mix64(seed * 4 + r + 0x1234567ull) | 1ull;
            add[r] = mix64(seed * 4 + r + 0x7654321ull);
        }
        shift = (s + 1) / 2;
        if (shift < 1) shift = 1;
    }
    uint64_t perm(uint64_t x) const {
        if (!permute) return x;
        for (int r = 0; r < 4; r++) {
            x = (x * mul[r] + add[r]) & mask;
            x ^= x >> shift;
        }
        return x;
    }
    void edge(uint64_t i, uint32_t* src, uint32_t* dst) const {
        const uint64_t h = mix64(seed ^ mix64(i + GOLDEN));
        uint64_t z = 0, s = 0, d = 0;
        for (int l = 0; l < scale; l++) {
            uint64_t u;
            if ((l & 1) == 0) { z = mix64(h + (uint64_t)(l / 2 + 1) * GOLDEN); u = z & 0xffffffffull; }
            else u = z >> 32;
            const uint64_t row = u >= tab ? 1 : 0;
            const uint64_t col = (u >= ta && u < tab) || u >= tabc ? 1 : 0;
            s = (s << 1) | row;
            d = (d << 1) | col;
        }
        *src = (uint32_t)perm(s);
        *dst = (uint32_t)perm(d);
    }
};

bool write_egr(const std::string& path, uint64_t n, uint64_t m, const std::vector<uint64_t>& off,
               const std::vector<int32_t>& adj, int wide_mode) {
    const bool wide = wide_mode == 1 || (wide_mode < 0 && m >= (1ull << 31));
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    int32_t h0 = (int32_t)n;
    fwrite(&h0, 4, 1, f);
    if (wide) {
        int32_t s = -1;
        fwrite(&s, 4, 1, f);
        fwrite(&m, 8, 1, f);
        fwrite(off.data(), 8, n + 1, f);
    } else {
        int32_t e = (int32_t)m;
        fwrite(&e, 4, 1, f);
        std::vector<int32_t> t(1 << 20);
        for (uint64_t v = 0; v <= n; v += t.size()) {
            const uint64_t c = std::min<uint64_t>(t.size(), n + 1 - v);
            for (uint64_t i = 0; i < c; i++) t[i] = (int32_t)(uint32_t)off[v + i];
            fwrite(t.data(), 4, c, f);
        }
    }
    const size_t B = 1ull << 26;
    for (uint64_t e = 0; e < m; e += B) fwrite(adj.data() + e, 4, std::min<uint64_t>(B, m - e), f);
    return fclose(f) == 0;
}

}  // namespace

extern "C" int excent_rmat_main(int argc, const char* const* argv) {
    int scale = 16, ef = 16, wide = -1;
    double a = 0.57, b = 0.19, c = 0.19;
    uint64_t seed = 1;
    bool permute = true, dedup = true, loops = false, csc = false, weights = false, symmetric = false;
    std::string out;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", s.c_str()); exit(1); }
            return argv[++i];
        };
        if (s == "--scale") scale = atoi(next());
        else if (s == "--edge-factor") ef = atoi(next());
        else if (s == "--a") a = atof(next());
        else if (s == "--b") b = atof(next());
        else if (s == "--c") c = atof(next());
        else if (s == "--seed") seed = strtoull(next(), nullptr, 10);
        else if (s == "--no-permute") permute = false;
        else if (s == "--keep-duplicates") dedup = false;
        else if (s == "--keep-self-loops") loops = true;
        else if (s == "--wide") { std::string v = next(); wide = v == "on" ? 1 : v == "off" ? 0 : -1; }
        else if (s == "--csc") csc = true;
        else if (s == "--weights") weights = true;
        else if (s == "--symmetric") symmetric = true;   // every edge in both directions
        else if (s[0] == '-') { fprintf(stderr, "unknown option %s\n", s.c_str()); return 1; }
        else out = s;
    }
    if (out.empty() || scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1) {
        fprintf(stderr, "usage: excent_rmat [--scale 16] [--edge-factor 16] [--a .57 --b .19 --c .19] [--seed 1] "
                        "[--no-permute] [--keep-duplicates] [--keep-self-loops] [--symmetric] [--wide auto|on|off] [--csc] [--weights] out.egr\n");
        return 1;
    }
    const uint64_t n = 1ull << scale, M = (uint64_t)ef * n;
    Rmat R(scale, a, b, c, seed, permute);
    // pass 1: out-degrees
    std::vector<uint64_t> off(n + 1, 0);
    {
        std::vector<uint32_t> deg(n, 0);
#pragma omp parallel for schedule(static, 1 << 16)
        for (long long i = 0; i < (long long)M; i++) {
            uint32_t s, d;
            R.edge((uint64_t)i, &s, &d);
            if (!loops && s == d) continue;
            __atomic_fetch_add(&deg[s], 1u, __ATOMIC_RELAXED);
            if (symmetric && s != d) __atomic_fetch_add(&deg[d], 1u, __ATOMIC_RELAXED);
        }
        for (uint64_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
    }
    fprintf(stderr, "rmat: out-degrees counted\n");
    uint64_t m0 = off[n];
    std::vector<int32_t> adj(m0 ? m0 : 1);
    {   // pass 2: scatter
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static, 1 << 16)
        for (long long i = 0; i < (long long)M; i++) {
            uint32_t s, d;
            R.edge((uint64_t)i, &s, &d);
            if (!loops && s == d) continue;
            const uint64_t p = __atomic_fetch_add(&pos[s], 1ull, __ATOMIC_RELAXED);
            adj[p] = (int32_t)d;
            if (symmetric && s != d) adj[__atomic_fetch_add(&pos[d], 1ull, __ATOMIC_RELAXED)] = (int32_t)s;
        }
    }
    fprintf(stderr, "rmat: edges placed\n");
    std::vector<uint32_t> ndeg(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long v = 0; v < (long long)n; v++) {
        int32_t* b0 = adj.data() + off[v];
        int32_t* b1 = adj.data() + off[v + 1];
        std::sort(b0, b1);
        ndeg[v] = (uint32_t)(dedup ? std::unique(b0, b1) - b0 : b1 - b0);
    }
    uint64_t m = 0;
    {   // compact in place, in vertex order (new offsets never exceed old ones)
        uint64_t w = 0;
        for (uint64_t v = 0; v < n; v++) {
            const uint64_t r0 = off[v], d = ndeg[v];
            if (w != r0 && d) memmove(adj.data() + w, adj.data() + r0, d * sizeof(int32_t));
            off[v] = w;
            w += d;
        }
        off[n] = w;
        m = w;
    }
    adj.resize(m ? m : 1);
    fprintf(stderr, "rmat: lists sorted and deduplicated, writing %s\n", out.c_str());
    if (!write_egr(out, n, m, off, adj, wide)) { fprintf(stderr, "cannot write %s\n", out.c_str()); return 1; }
    if (weights) {
        std::vector<float> w(m);
#pragma omp parallel for schedule(static, 1 << 16)
        for (long long e = 0; e < (long long)m; e++) {
            const uint64_t z = mix64((seed * GOLDEN) ^ mix64((uint64_t)e + 0x5bd1e995ull));
            w[e] = 0.5f + (float)((z >> 40) * (1.0 / 16777216.0));   // 24-bit uniform in [0.5, 1.5)
        }
        std::string p = out + ".w.f32";
        FILE* f = fopen(p.c_str(), "wb");
        if (!f || fwrite(w.data(), 4, m, f) != m) { fprintf(stderr, "cannot write %s\n", p.c_str()); if (f) fclose(f); return 1; }
        fclose(f);
    }
    if (csc) {
        std::vector<uint64_t> toff(n + 1, 0);
        for (uint64_t e = 0; e < m; e++) toff[(uint32_t)adj[e] + 1]++;
        for (uint64_t v = 0; v < n; v++) toff[v + 1] += toff[v];
        std::vector<int32_t> tadj(m ? m : 1);
        std::vector<uint64_t> pos(toff.begin(), toff.end() - 1);
        for (uint64_t v = 0; v < n; v++)
            for (uint64_t e = off[v]; e < off[v + 1]; e++) tadj[pos[(uint32_t)adj[e]]++] = (int32_t)v;
        if (!write_egr(out + ".csc.egr", n, m, toff, tadj, wide)) { fprintf(stderr, "cannot write csc\n"); return 1; }
    }
    printf("rmat: scale %d edge factor %d -> n=%llu m=%llu (%llu generated, %s)\n", scale, ef,
           (unsigned long long)n, (unsigned long long)m, (unsigned long long)M, out.c_str());
    return 0;
}
