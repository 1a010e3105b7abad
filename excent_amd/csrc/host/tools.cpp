// tools.cpp -- bin/egr_convert, bin/centrality_reference, bin/excent_params.
//
// egr_convert follows tools/efficient_converter.py:5-113 (paths relative to the
// reference root): two passes over a 1-indexed .mtx edge list, lists in file order,
// duplicates and self-loops kept, sources >= nodes dropped, always the wide format.
// centrality_reference follows tools/centrality_reference.cu:78-222 (FP64 solve,
// --compare, --compare-dumps, --rank); excent_params follows tools/excent_params.cu:73-174.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "egr_graph.hpp"
#include "excent.h"
#include "excent_hip.h"

namespace {

std::vector<char> slurp(const char* p, bool* ok) {
    std::vector<char> b;
    FILE* f = fopen(p, "rb");
    if (!f) { perror(p); *ok = false; return b; }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    b.resize(sz);
    *ok = fread(b.data(), 1, sz, f) == (size_t)sz;
    fclose(f);
    return b;
}

void unit(std::vector<double>& v) {   // tools/centrality_reference.cu:71-76
    double s = 0, t = 0;
    for (double a : v) { s += a * a; t += a; }
    const double f = (t < 0 ? -1.0 : 1.0) / std::sqrt(s);
    for (double& a : v) a *= f;
}

// Order of the largest values, ties by vertex id.
std::vector<int> top_order(const std::vector<double>& x, size_t k) {
    std::vector<int> o(x.size());
    std::iota(o.begin(), o.end(), 0);
    k = std::min(k, o.size());
    std::partial_sort(o.begin(), o.begin() + k, o.end(), [&](int a, int b) { return x[a] > x[b] || (x[a] == x[b] && a < b); });
    o.resize(k);
    return o;
}
size_t overlap(std::vector<int> a, std::vector<int> b) {
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    std::vector<int> c;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(c));
    return c.size();
}

int load_pair(const std::string& measure, const char* refp, const char* dumpp, bool ref_f32,
              std::vector<double>& r, std::vector<double>& x) {
    bool ok1 = true, ok2 = true;
    auto rb = slurp(refp, &ok1), xb = slurp(dumpp, &ok2);
    if (!ok1 || !ok2) return 1;
    const size_t n = rb.size() / (ref_f32 ? sizeof(float) : sizeof(double));
    if (xb.size() != n * sizeof(float)) { fprintf(stderr, "size mismatch: ref n=%zu, dump %zu floats\n", n, xb.size() / 4); return 1; }
    r.resize(n);
    x.resize(n);
    if (ref_f32) { const float* rf = (const float*)rb.data(); for (size_t i = 0; i < n; i++) r[i] = rf[i]; }
    else memcpy(r.data(), rb.data(), n * sizeof(double));
    const float* xf = (const float*)xb.data();
    for (size_t i = 0; i < n; i++) x[i] = xf[i];
    if (measure == "ev") { unit(r); unit(x); }
    return 0;
}

// tools/centrality_reference.cu:78-106
int compare(const std::string& measure, const char* refp, const char* dumpp, bool ref_f32) {
    std::vector<double> r, x;
    if (load_pair(measure, refp, dumpp, ref_f32, r, x)) return 1;
    const size_t n = r.size();
    double l1 = 0, ref1 = 0, l2 = 0, maxrel = 0;
    for (size_t i = 0; i < n; i++) {
        const double d = std::fabs(x[i] - r[i]);
        l1 += d; ref1 += std::fabs(r[i]); l2 += d * d;
        if (std::fabs(r[i]) > 1e-12) maxrel = std::max(maxrel, d / std::fabs(r[i]));
    }
    const size_t K = std::min<size_t>(100, n);
    const size_t common = overlap(top_order(r, K), top_order(x, K));
    printf("%s: n=%zu  L1 rel err %.3e  L2 err %.3e%s  top-%zu overlap %zu/%zu  max rel err %.3e\n", measure.c_str(), n,
           l1 / ref1, std::sqrt(l2), measure == "ev" ? " (unit vectors)" : "", K, common, K, maxrel);
    return 0;
}

// tools/centrality_reference.cu:109-139
int rank_quality(const std::string& measure, const char* refp, const char* dumpp) {
    std::vector<double> r, x;
    if (load_pair(measure, refp, dumpp, false, r, x)) return 1;
    const size_t n = r.size(), KMAX = std::min<size_t>(100000, n);
    auto ir = top_order(r, KMAX), ix = top_order(x, KMAX);
    printf("%s:", measure.c_str());
    for (size_t k : {10, 100, 1000, 10000, 100000}) {
        if (k > KMAX) break;
        std::vector<int> a(ir.begin(), ir.begin() + k), b(ix.begin(), ix.begin() + k);
        printf("  P@%zu %.4f", k, (double)overlap(a, b) / k);
    }
    const int T = (int)std::min<size_t>(1000, n);
    long long conc = 0, disc = 0;
    for (int i = 0; i < T; i++)
        for (int j = i + 1; j < T; j++) {
            const double dr = r[ir[i]] - r[ir[j]], dx = x[ir[i]] - x[ir[j]];
            if (dr * dx > 0) conc++; else if (dr * dx < 0) disc++;
        }
    printf("  tau@1000 %.4f\n", T > 1 ? (double)(conc - disc) / ((double)T * (T - 1) / 2) : 1.0);
    return 0;
}

bool load_graph(const char* path, int* n, uint64_t* m, std::vector<uint64_t>& idx, std::vector<int32_t>& adj) {
    try {
        excent::EgrGraph g(path, 16, true);
        *n = g.n; *m = g.m;
        idx.resize((size_t)g.n + 1);
        g.read_offsets(0, idx.size(), idx.data());
        adj.resize(g.m ? g.m : 1);
        if (g.m) g.read_edges(0, g.m, adj.data());
        return true;
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        return false;
    }
}

}  // namespace

extern "C" int excent_convert_main(int argc, const char* const* argv) {
    if (argc != 3) { printf("Usage: egr_convert <input.mtx> <output.egr>\n"); return 1; }
    FILE* f = fopen(argv[1], "r");
    if (!f) { perror(argv[1]); return 1; }
    long long num_nodes = 0, num_edges = 0;
    char line[4096];
    long hdr_pos = -1;
    while (fgets(line, sizeof line, f)) {   // efficient_converter.py:10-18
        if (line[0] == '%') continue;
        long long a, b, c;
        int k = sscanf(line, "%lld %lld %lld", &a, &b, &c);
        if (k >= 2) { num_nodes = std::max(a, b); num_edges = k >= 3 ? c : 0; hdr_pos = ftell(f); break; }
    }
    if (hdr_pos < 0) { fprintf(stderr, "%s: no size line\n", argv[1]); fclose(f); return 1; }
    printf("Graph stats: %lld nodes, %lld edges.\n", num_nodes, num_edges);
    std::vector<uint64_t> deg((size_t)num_nodes + 1, 0);
    auto parse_edge = [&](const char* s, long long* u, long long* v) { return sscanf(s, "%lld %lld", u, v) == 2; };
    fseek(f, hdr_pos, SEEK_SET);
    while (fgets(line, sizeof line, f)) {   // pass 1 (py:36-53)
        long long u, v;
        if (!parse_edge(line, &u, &v)) continue;
        if (u < 1) { fprintf(stderr, "vertex ids are 1-indexed: %lld\n", u); fclose(f); return 1; }
        if (u - 1 < num_nodes) deg[u - 1]++;
    }
    std::vector<uint64_t> nindex((size_t)num_nodes + 1, 0);
    for (long long i = 0; i < num_nodes; i++) nindex[i + 1] = nindex[i] + deg[i];
    const uint64_t m = nindex[num_nodes];
    std::vector<int32_t> nlist(m ? m : 1);
    std::vector<uint64_t> pos(nindex.begin(), nindex.end() - 1);
    fseek(f, hdr_pos, SEEK_SET);
    while (fgets(line, sizeof line, f)) {   // pass 2 (py:89-110)
        long long u, v;
        if (!parse_edge(line, &u, &v)) continue;
        if (u - 1 < num_nodes) nlist[pos[u - 1]++] = (int32_t)(v - 1);
    }
    fclose(f);
    FILE* o = fopen(argv[2], "wb");
    if (!o) { perror(argv[2]); return 1; }
    const int32_t nn = (int32_t)num_nodes, s = -1;
    fwrite(&nn, 4, 1, o);
    fwrite(&s, 4, 1, o);
    fwrite(&m, 8, 1, o);
    fwrite(nindex.data(), 8, nindex.size(), o);
    if (m) fwrite(nlist.data(), 4, m, o);
    if (fclose(o) != 0) { perror(argv[2]); return 1; }
    printf("Conversion complete successfully!\n");
    return 0;
}

extern "C" int excent_reference_main(int argc, const char* const* argv) {
    if (argc >= 5 && !strcmp(argv[1], "--compare")) return compare(argv[2], argv[3], argv[4], false);
    if (argc >= 5 && !strcmp(argv[1], "--rank")) return rank_quality(argv[2], argv[3], argv[4]);
    if (argc >= 5 && !strcmp(argv[1], "--compare-dumps")) return compare(argv[2], argv[3], argv[4], true);
    if (argc < 4) {
        fprintf(stderr, "usage: %s graph.egr pr|katz|ev ref.bin [--alpha A] [--beta B] [--max-iter N] [--tol T]\n"
                        "       %s --compare pr|katz|ev ref.bin dump.bin\n", argv[0], argv[0]);
        return 1;
    }
    const std::string measure = argv[2];
    double alpha = -1, beta = 1, tol = 1e-9;
    int max_iter = 20000;
    for (int i = 4; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--alpha")) alpha = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--beta")) beta = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--max-iter")) max_iter = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--tol")) tol = atof(argv[i + 1]);
    }
    if (measure == "katz" && alpha <= 0) { fprintf(stderr, "katz needs --alpha\n"); return 1; }
    if (measure != "pr" && measure != "katz" && measure != "ev") { fprintf(stderr, "measure must be pr, katz or ev\n"); return 1; }
    int n;
    uint64_t m;
    std::vector<uint64_t> idx;
    std::vector<int32_t> adj;
    if (!load_graph(argv[1], &n, &m, idx, adj)) return 1;
    void* S = nullptr;
    if (exc_fp64_create(0, n, m, idx.data(), adj.data(), &S)) { fprintf(stderr, "%s\n", exc_last_error()); return 1; }
    const int kind = measure == "pr" ? EXC_PR : measure == "katz" ? EXC_KATZ : EXC_EV;
    std::vector<double> h(n);
    double rel = 0, lam = 0;
    int it = 0;
    if (exc_fp64_solve(S, kind, alpha, beta, max_iter, tol, nullptr, h.data(), &rel, &lam, &it)) {
        fprintf(stderr, "%s\n", exc_last_error());
        exc_fp64_destroy(S);
        return 1;
    }
    exc_fp64_destroy(S);
    double sum = 0;
    for (double v : h) sum += v;
    printf("reference %s: %d iterations, final rel change %.2e, sum %.6e", measure.c_str(), it, rel, sum);
    if (measure == "ev") printf(" (unit L2; lambda1 ~ %f)", lam);
    printf("\n");
    FILE* o = fopen(argv[3], "wb");
    if (!o) { perror(argv[3]); return 1; }
    fwrite(h.data(), sizeof(double), n, o);
    fclose(o);
    return 0;
}

// Extensions: --directed (skip the symmetry check), --alpha-factor F (default 0.85),
// --seed S (seeded random positive start for the power iteration; default all ones),
// --eps-factor F (default 0.01), --eps-abs E (absolute per-vertex thresholds for katz/ev).
extern "C" int excent_params_main(int argc, const char* const* argv) {
    if (argc < 3) { fprintf(stderr, "usage: %s graph.egr out.env [--directed] [--alpha-factor 0.85] [--seed S] [--eps-factor 0.01] [--eps-abs E]\n", argv[0]); return 1; }
    bool directed = false;
    double afac = 0.85, efac = 0.01, eabs = -1;
    long long seed = -1;
    for (int i = 3; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--directed") directed = true;
        else if (a == "--alpha-factor" && i + 1 < argc) afac = atof(argv[++i]);
        else if (a == "--seed" && i + 1 < argc) seed = atoll(argv[++i]);
        else if (a == "--eps-factor" && i + 1 < argc) efac = atof(argv[++i]);
        else if (a == "--eps-abs" && i + 1 < argc) eabs = atof(argv[++i]);
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 1; }
    }
    int n;
    uint64_t m;
    std::vector<uint64_t> idx;
    std::vector<int32_t> adj;
    if (!load_graph(argv[1], &n, &m, idx, adj)) return 1;
    printf("graph: n=%d m=%llu\n", n, (unsigned long long)m);
    if (!directed && m) {   // tools/excent_params.cu:105-118
        std::mt19937_64 rng(42);
        int bad = 0;
        const int S = 200000;
        for (int i = 0; i < S; i++) {
            const uint64_t e = rng() % m;
            const int v = (int)(std::upper_bound(idx.begin(), idx.end(), e) - idx.begin()) - 1;
            const int u = adj[e];
            const int32_t* b = adj.data() + idx[u];
            const int32_t* en = adj.data() + idx[u + 1];
            const bool found = std::is_sorted(b, en) ? std::binary_search(b, en, v) : std::find(b, en, v) != en;
            if (!found) bad++;
        }
        if (bad) { fprintf(stderr, "graph is not symmetric (%d/%d sampled edges lack a reverse edge)\n", bad, S); return 1; }
    }
    void* Sv = nullptr;
    if (exc_fp64_create(0, n, m, idx.data(), adj.data(), &Sv)) { fprintf(stderr, "%s\n", exc_last_error()); return 1; }
    std::vector<double> start;
    if (seed >= 0) {
        std::mt19937_64 rng((uint64_t)seed);
        std::uniform_real_distribution<double> U(0.5, 1.5);
        start.resize(n);
        for (auto& s : start) s = U(rng);
    }
    double lam = 0, meanEV = 0, meanK = 0;
    int pit = 0, kit = 0;
    if (exc_fp64_lambda(Sv, start.empty() ? nullptr : start.data(), 5000, 1e-12, &lam, &meanEV, &pit) ||
        exc_fp64_katz_mean(Sv, afac / lam, 1.0, &meanK, &kit)) {
        fprintf(stderr, "%s\n", exc_last_error());
        exc_fp64_destroy(Sv);
        return 1;
    }
    exc_fp64_destroy(Sv);
    const double alpha = afac / lam;
    printf("lambda1 = %.10f after %d power iterations; eigenvector fixed-point mean = %.6g\n", lam, pit, meanEV);
    printf("Katz: alpha = %.6g/lambda1 = %.6g, beta = 1, fixed-point mean = %.6g\n", afac, alpha, meanK);
    FILE* o = fopen(argv[2], "w");
    if (!o) { perror(argv[2]); return 1; }
    fprintf(o, "# Given constants for %s (excent_params; lambda1 after %d power iterations).\n", argv[1], pit);
    fprintf(o, "# Katz alpha = %g / lambda1, beta = 1. Thresholds: %s.\n", afac,
            eabs > 0 ? "absolute per-vertex values" : "0.01 x each measure's fixed-point mean");
    fprintf(o, "LAMBDA=%.10f\nALPHA=%.6g\nBETA=1\nEPS_PR=0.01\nEPS_KATZ=%.6g\nEPS_EV=%.6g\n", lam, alpha,
            eabs > 0 ? eabs : efac * meanK, eabs > 0 ? eabs : efac * meanEV);
    fclose(o);
    printf("wrote %s\n", argv[2]);
    return 0;
}
