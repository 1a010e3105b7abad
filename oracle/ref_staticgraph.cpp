// ref_staticgraph.cpp -- C entry points over the reference's own storage reader,
// include/StaticGraph.h, compiled from the header where it lies in the reference
// tree (see oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE ONLY: it lets the
// CPU tests check the oracle's and the product's EGR parsing, chunk planner and
// gather page/run counting against the reference code itself.  Nothing here is
// copied from the reference; the header is included by path at build time.
#include "StaticGraph.h"

extern "C" {

// Opens `path` with the reference reader; returns n, m and offsets (caller buffer of n+1).
int ref_open(const char* path, int offsets_in_ram, int* n, unsigned long long* m,
             unsigned long long* idx_out, long long idx_cap) {
    StaticGraph g(path, 16, offsets_in_ram != 0);
    *n = g.n;
    *m = g.m;
    if (idx_out && idx_cap >= (long long)g.n + 1)
        for (long long v = 0; v <= g.n; v++) idx_out[v] = g.offset((size_t)v);
    return 0;
}

// StaticGraph::plan_chunks (include/StaticGraph.h:128-139): writes v0,v1,e0,e1 per chunk.
int ref_plan_chunks(const char* path, unsigned long long cap, int* v0, int* v1,
                    unsigned long long* e0, unsigned long long* e1, int max_chunks) {
    StaticGraph g(path, 16, false);
    std::vector<EdgeChunk> ch = g.plan_chunks(cap);
    int nc = (int)ch.size();
    for (int c = 0; c < nc && c < max_chunks; c++) { v0[c] = ch[c].v0; v1[c] = ch[c].v1; e0[c] = ch[c].e0; e1[c] = ch[c].e1; }
    return nc;
}

// StaticGraph::prefetch_lists (include/StaticGraph.h:90-114) page and run counts.
int ref_prefetch_counts(const char* path, const int* vids, int k, int io_threads,
                        unsigned long long* pages, unsigned long long* runs) {
    StaticGraph g(path, io_threads, false);
    g.prefetch_lists(vids, k, pages, runs);
    return 0;
}

// StaticGraph::read_edges (include/StaticGraph.h:86-88) into a caller buffer.
int ref_read_edges(const char* path, unsigned long long e0, unsigned long long e1, int* dst) {
    StaticGraph g(path, 16, false);
    g.read_edges(e0, e1, dst);
    return 0;
}

}
