// egr_graph.hpp -- storage-resident EGR graph reader (mmap + parallel pread).
//
// Behaviour follows the reference reader include/StaticGraph.h:23-171 (paths relative
// to the reference root): the same header rules, offsets in RAM for the narrow format
// and in the mapping for the wide one, a whole-file read-only mapping advised
// MADV_RANDOM / POSIX_FADV_SEQUENTIAL, 8 MB pread pieces on up to io_threads OpenMP
// threads, page-run merged MADV_WILLNEED prefetch and page-cache dropping.  Errors
// throw egr_error (the reference prints and exits; the CLI does that on catch).
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

class EgrGraph {
public:
    int n = 0;
    uint64_t m = 0;
    bool wide = false;

    EgrGraph(const std::string& path, int io_threads = 16, bool offsets_in_ram = true);
    ~EgrGraph();
    EgrGraph(const EgrGraph&) = delete;
    EgrGraph& operator=(const EgrGraph&) = delete;

    uint64_t offset(size_t v) const { return off_[v]; }
    uint64_t degree(int v) const { return off_[v + 1] - off_[v]; }
    const int32_t* edges(int v) const { return nlist_ + off_[v]; }
    const int32_t* nlist() const { return nlist_; }
    const uint64_t* offsets() const { return off_; }
    size_t nlist_offset() const { return nlist_off_; }

    void read_offsets(size_t v0, size_t cnt, uint64_t* dst) const;
    void read_edges(uint64_t e0, uint64_t e1, int32_t* dst) const;
    // Prefetch the pages of the listed vertices' lists (merged runs per part).
    void prefetch_lists(const int32_t* vids, int k, uint64_t* pages, uint64_t* runs) const;
    void drop_cache(uint64_t e0, uint64_t e1, size_t v0, size_t v1) const;
    std::vector<EdgeChunk> plan_chunks(uint64_t max_edges) const;

private:
    int fd_ = -1;
    int io_threads_ = 16;
    size_t nlist_off_ = 0, idx_off_ = 0;
    std::vector<uint64_t> nindex_;
    const uint64_t* off_ = nullptr;
    void* map_ = nullptr;
    size_t map_size_ = 0;
    const int32_t* nlist_ = nullptr;
    std::string path_;

    void read_exact(void* dst, size_t bytes, size_t off) const;
    void read_parallel(void* dst, size_t bytes, size_t off) const;
};

}  // namespace excent
