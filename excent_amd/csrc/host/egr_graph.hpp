// egr_graph.hpp -- the EGR graph file on the host: layout, offsets, column reads, page advice.
//
// The file format is the reference's (README.md:136-144; include/StaticGraph.h:32-55, paths
// relative to the reference root): int32 n; int32 m, or -1 followed by int64 m (wide);
// n+1 offsets as uint32 (narrow) or uint64 (wide); m int32 column ids.  Observable behaviour
// kept for drop-in parity: the chunk plan (StaticGraph.h:128-139) and the page/run counts of
// a gather's readahead (StaticGraph.h:90-114), which the pass lines report.
//
// Design: the whole file is mapped read-only (gathers copy lists straight out of the
// mapping; the offsets of a wide file are read in place when they are not copied to RAM);
// streamed chunks are copied in 4 MiB blocks that a pool of I/O threads claims from a shared
// counter, so a slow block does not hold up a fixed share of the range -- out of the mapping
// when the whole range is in the page cache, else with pread; the number of I/O threads is
// the caller's (0: every host thread OpenMP offers).  Errors throw
// egr_error; the command line turns them into its exit path.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace excent {

struct egr_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct EdgeChunk {
    int v0, v1;
    uint64_t e0, e1;
    uint64_t edges() const { return e1 - e0; }
};

// Lists of a gather are split into this many contiguous groups for readahead (at least 1024
// lists per group); page runs are merged within a group.  The reference reader uses its I/O
// thread count (16 in bin/excent) for the same split, so the page/run counts it reports
// depend on it; here the split is fixed and the I/O threads are separate.
constexpr int kAdviceGroups = 16;

class EgrGraph {
public:
    int n = 0;
    uint64_t m = 0;
    bool wide = false;

    // io_threads <= 0: omp_get_max_threads().  offsets_in_ram: copy wide offsets to RAM
    // (narrow offsets are always widened into RAM).
    EgrGraph(const std::string& path, int io_threads = 0, bool offsets_in_ram = true);
    ~EgrGraph();
    EgrGraph(const EgrGraph&) = delete;
    EgrGraph& operator=(const EgrGraph&) = delete;

    uint64_t offset(size_t v) const { return off_[v]; }
    uint64_t degree(int v) const { return off_[v + 1] - off_[v]; }
    const int32_t* edges(int v) const { return cols_ + off_[v]; }
    const int32_t* nlist() const { return cols_; }
    const uint64_t* offsets() const { return off_; }
    size_t nlist_offset() const { return cols_at_; }
    int io_threads() const { return io_threads_; }

    void read_offsets(size_t v0, size_t cnt, uint64_t* dst) const;
    // Columns [e0, e1) into dst (pread by the I/O pool).
    void read_edges(uint64_t e0, uint64_t e1, int32_t* dst) const;
    // Readahead (MADV_WILLNEED) of the pages holding the lists of vids[0..k): lists are split
    // into `groups` contiguous groups, touching or overlapping page spans are merged into runs
    // within a group, and each run is advised once.  *pages / *runs: the totals.
    void prefetch_lists(const int32_t* vids, int k, uint64_t* pages, uint64_t* runs, int groups = kAdviceGroups) const;
    // Drops the pages of columns [e0, e1) and offsets [v0, v1] from this mapping and the page
    // cache (cold-read calibration and measurements).
    void drop_cache(uint64_t e0, uint64_t e1, size_t v0, size_t v1) const;
    // Vertex-contiguous chunks, each the longest run of whole lists starting at its first
    // vertex whose columns fit in max_edges; a list longer than that is a chunk of its own.
    std::vector<EdgeChunk> plan_chunks(uint64_t max_edges) const;

private:
    std::string path_;
    int fd_ = -1;
    int io_threads_ = 1;
    size_t offsets_at_ = 0, cols_at_ = 0;   // byte positions in the file
    size_t file_bytes_ = 0;
    const char* map_ = nullptr;
    std::vector<uint64_t> ram_off_;
    const uint64_t* off_ = nullptr;
    const int32_t* cols_ = nullptr;

    bool copy_from_map_ = true;   // EXC_READ=pread: always pread (measurements)
    void load(bool offsets_in_ram);
    bool resident(size_t at, size_t bytes) const;
    void read_range(void* dst, size_t bytes, size_t at) const;
    void advise_out(size_t b0, size_t b1) const;
};

}  // namespace excent
