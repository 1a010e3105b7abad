// capi.cpp -- C entry points over the host planner and reader (include/excent.h), so
// the tests can hold the product's decisions against the oracle bit for bit.
#include <algorithm>
#include <cstring>
#include <string>

#include "egr_graph.hpp"
#include "excent.h"
#include "planner.hpp"
#include "staging.hpp"

namespace {
thread_local std::string g_err;
int err(const std::string& s) { g_err = s; return -1; }
}  // namespace

extern "C" {

const char* excent_last_error(void) { return g_err.c_str(); }

uint64_t excent_chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges) {
    return excent::chunk_cap(m, chunk_medges, target_chunks, min_chunk_medges);
}

int excent_plan_chunks(const char* path, uint64_t cap, int* v0, int* v1, uint64_t* e0, uint64_t* e1, int max) {
    try {
        excent::EgrGraph g(path, 16, false);
        auto ch = g.plan_chunks(cap);
        for (size_t c = 0; c < ch.size() && (int)c < max; c++) { v0[c] = ch[c].v0; v1[c] = ch[c].v1; e0[c] = ch[c].e0; e1[c] = ch[c].e1; }
        return (int)ch.size();
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_egr_header(const char* path, int* n, uint64_t* m, int* wide) {
    try {
        excent::EgrGraph g(path, 16, false);
        *n = g.n; *m = g.m; *wide = g.wide ? 1 : 0;
        return 0;
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_egr_offsets(const char* path, uint64_t* idx_out) {
    try {
        excent::EgrGraph g(path, 16, false);
        g.read_offsets(0, (size_t)g.n + 1, idx_out);
        return 0;
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_egr_edges(const char* path, uint64_t e0, uint64_t e1, int32_t* out) {
    try {
        excent::EgrGraph g(path, 16, false);
        if (e1 > g.m || e0 > e1) return err("edge range out of bounds");
        g.read_edges(e0, e1, out);
        return 0;
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_fetch_plan(int nc, const uint64_t* cnt, const uint8_t* resident, const uint8_t* loaded, double tau,
                      const uint64_t* chunk_edges, int hot_mode, int pass, const int* hlo, const int* hhi,
                      int8_t* mode, uint64_t* pend_v, uint64_t* demand) {
    std::vector<char> res(resident, resident + nc), ld(loaded, loaded + nc);
    std::vector<uint64_t> ce(chunk_edges, chunk_edges + nc);
    std::vector<int> lo(hlo, hlo + nc), hi(hhi, hhi + nc);
    auto o = excent::fetch_plan(nc, cnt, res, ld, tau, ce, hot_mode != 0, pass, lo, hi, mode);
    *pend_v = o.pend_v;
    *demand = o.demand;
    return o.need_bitmap ? 1 : 0;
}

int excent_pick_group(double avg, int cap) { return excent::pick_group(avg, cap); }

int excent_hot_select(int n, const uint64_t* off, uint64_t budget, const uint16_t* score,
                      int32_t* ids, uint64_t* hoff, uint32_t* bitmask, uint64_t* deg_min) {
    auto h = excent::hot_select(n, off, budget, score);
    memcpy(ids, h.ids.data(), h.ids.size() * sizeof(int32_t));
    memcpy(hoff, h.off.data(), h.off.size() * sizeof(uint64_t));
    memcpy(bitmask, h.bitmask.data(), h.bitmask.size() * sizeof(uint32_t));
    *deg_min = h.deg_min;
    return (int)h.ids.size();
}

double excent_tau(double t_stream, const double* t_g, const double* f_got) {
    return excent::tau_from_timings(t_stream, t_g, f_got);
}

int excent_prefetch_counts(const char* path, const int32_t* vids, int k, int io_threads,
                           uint64_t* pages, uint64_t* runs) {
    try {
        excent::EgrGraph g(path, io_threads, false);
        g.prefetch_lists(vids, k, pages, runs, io_threads);   // the reference splits by its thread count
        return 0;
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_stage_lists(const char* path, const int32_t* vids, int k, int dedup, int threads, uint64_t* off_out,
                       int32_t* edges_out, uint64_t* pages, uint64_t* runs) {
    try {
        excent::EgrGraph g(path, threads, false);
        for (int i = 0; i < k; i++)
            if (vids[i] < 0 || vids[i] >= g.n) return err("vertex id out of range");
        const excent::GatherStats st = excent::stage_lists(g, vids, k, off_out, edges_out, dedup != 0, threads);
        *pages = st.pages;
        *runs = st.runs;
        return 0;
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_top_in_degree(const char* path, int count, double sample, int threads, int32_t* out) {
    try {
        excent::EgrGraph g(path, threads, false);
        const std::vector<int32_t> t = excent::top_in_degree(g, count, sample, threads);
        std::copy(t.begin(), t.end(), out);
        return (int)t.size();
    } catch (const std::exception& e) { return err(e.what()); }
}

int excent_gather_list(const uint32_t* bitmap, int v0, int v1, int32_t* out) {
    return excent::gather_list(bitmap, v0, v1, out);
}

}  // extern "C"
