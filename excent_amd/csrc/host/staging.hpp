// staging.hpp -- host side of the fetch paths: packing a gather's lists into a staging slot,
// and the break-even demand tau of the stream-vs-gather rule measured on this machine.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "egr_graph.hpp"

namespace excent {

struct GatherStats {
    uint64_t edges = 0;            // columns packed
    uint64_t pages = 0, runs = 0;  // readahead of the lists (dedup on)
};

// Packs the lists of vids[0..k) (ascending) into edges_out, with off_out[0..k] their
// offsets there (the gather path of src/excent.cu:709-740): the offsets by a blocked
// parallel scan, the page readahead of the lists when dedup (EgrGraph::prefetch_lists),
// then the columns copied out of the file mapping by `threads` workers that each take an
// equal share of the columns (a long list is split between workers).
GatherStats stage_lists(const EgrGraph& g, const int32_t* vids, int k, uint64_t* off_out, int32_t* edges_out,
                        bool dedup, int threads);

// tau calibration (the rule of src/excent.cu:515-579 is measured per machine; the model and
// the probes here are this build's).  On one chunk, with the pages of the chunk dropped from
// the page cache before every probe:
//   * storage: the time to read the whole chunk (stream) and to pack three gathers of
//     demand fractions 0.05, 0.15 and 0.4 of its columns, the lists picked by a seeded
//     degree-weighted draw over the chunk (pending lists spread over the chunk, hubs more
//     often), each through the same code as the passes;
//   * link: the time of the host-to-device copy of each of those payloads from the slot.
// The two legs of a fetch overlap across the double-buffered slots, so a fetch costs
// max(storage, link); tau is where the gather's cost crosses the stream's
// (tau_from_timings), within [0.005, 1].
struct TauProbe {
    double tau = 0;
    double t_stream = 0;        // cost of streaming the chunk
    double t_gather[3] = {};    // cost of the three gathers
    double f_got[3] = {};       // their demand: packed columns / chunk columns
    double storage_seq_gbs = 0; // chunk read, GB/s (cold)
    double link_gbs = 0;        // host-to-device copy of the chunk, GB/s
    double seconds = 0;         // whole calibration
};

// h_edges/h_vids/h_off: a staging slot's host buffers (capacity: the chunk); copy_probe(bytes)
// returns the seconds of a synchronous host-to-device copy of that many bytes of h_edges.
TauProbe calibrate_tau(const EgrGraph& g, const EdgeChunk& ch, int32_t* h_edges, int32_t* h_vids, uint64_t* h_off,
                       bool dedup, int threads, const std::function<double(uint64_t)>& copy_probe);

// The (up to) `count` vertices with the most in-edges, estimated from a sample of the column
// array: about `sample` of the columns (at least 16M, or all), read in evenly spaced 1 MiB
// blocks; ties and the order of the result do not matter (the caller privatises them).
std::vector<int32_t> top_in_degree(const EgrGraph& g, int count, double sample, int threads);

}  // namespace excent
