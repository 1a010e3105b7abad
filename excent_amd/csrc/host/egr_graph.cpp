// egr_graph.cpp -- see egr_graph.hpp.
#include "egr_graph.hpp"

#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace excent {
namespace {

constexpr size_t kBlock = 4ull << 20;   // unit of work of the I/O pool

size_t page_size() {
    static const size_t p = (size_t)sysconf(_SC_PAGESIZE);
    return p;
}

// pread until `len` bytes are in, retrying interrupted and short reads.
void pread_all(int fd, void* dst, size_t len, size_t at, const std::string& path) {
    char* p = static_cast<char*>(dst);
    while (len) {
        const ssize_t got = pread(fd, p, len, (off_t)at);
        if (got < 0 && errno == EINTR) continue;
        if (got <= 0)
            throw egr_error("pread: " + path + (got == 0 ? ": unexpected end of file" : ": " + std::string(strerror(errno))));
        p += got;
        at += (size_t)got;
        len -= (size_t)got;
    }
}

// Page runs of one group of lists: spans are offered in list order; a span that starts at
// or before the end of the open run extends it, any other closes the run (advised with
// MADV_WILLNEED and counted) and opens a new one.
class RunMerger {
public:
    RunMerger(const char* map, size_t page) : map_(map), page_(page) {}
    void add(size_t p0, size_t p1) {
        if (open_ && p0 <= hi_) {
            hi_ = std::max(hi_, p1);
            return;
        }
        close();
        lo_ = p0;
        hi_ = p1;
        open_ = true;
    }
    void close() {
        if (!open_) return;
        madvise(const_cast<char*>(map_) + lo_ * page_, (hi_ - lo_) * page_, MADV_WILLNEED);
        pages += hi_ - lo_;
        runs++;
        open_ = false;
    }
    uint64_t pages = 0, runs = 0;

private:
    const char* map_;
    size_t page_, lo_ = 0, hi_ = 0;
    bool open_ = false;
};

}  // namespace

EgrGraph::EgrGraph(const std::string& path, int io_threads, bool offsets_in_ram) : path_(path) {
    io_threads_ = io_threads > 0 ? io_threads : std::max(1, omp_get_max_threads());
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) throw egr_error(path + ": " + strerror(errno));
    try {
        load(offsets_in_ram);
    } catch (...) {
        if (map_) munmap(const_cast<char*>(map_), file_bytes_);
        map_ = nullptr;
        ::close(fd_);
        fd_ = -1;
        throw;
    }
}

EgrGraph::~EgrGraph() {
    if (map_) munmap(const_cast<char*>(map_), file_bytes_);
    if (fd_ >= 0) ::close(fd_);
}

void EgrGraph::load(bool offsets_in_ram) {
    if (const char* r = getenv("EXC_READ")) copy_from_map_ = strcmp(r, "pread") != 0;
    struct stat st;
    if (fstat(fd_, &st) != 0) throw egr_error(path_ + ": fstat: " + strerror(errno));
    file_bytes_ = (size_t)st.st_size;
    // header: int32 n, then int32 m, or -1 and an int64 m
    unsigned char head[16] = {0};
    pread_all(fd_, head, std::min<size_t>(sizeof head, std::max<size_t>(file_bytes_, 8)), 0, path_);
    int32_t h0, h1;
    memcpy(&h0, head, 4);
    memcpy(&h1, head + 4, 4);
    n = h0;
    wide = h1 == -1;
    if (wide) {
        int64_t mw;
        memcpy(&mw, head + 8, 8);
        m = (uint64_t)mw;
    } else {
        m = (uint64_t)(int64_t)h1;   // the reference widens the int32 to unsigned long long
    }
    if (n < 1) throw egr_error("StaticGraph: bad header in " + path_);
    offsets_at_ = wide ? 16 : 8;
    const size_t off_width = wide ? sizeof(uint64_t) : sizeof(uint32_t);
    cols_at_ = offsets_at_ + ((size_t)n + 1) * off_width;
    const size_t need = cols_at_ + m * sizeof(int32_t);
    if (m > (SIZE_MAX - cols_at_) / sizeof(int32_t) || file_bytes_ < need)
        throw egr_error("StaticGraph: " + path_ + " is shorter than its header says (" + std::to_string(file_bytes_) + " < " +
                        std::to_string(need) + " bytes)");

    void* mp = mmap(nullptr, file_bytes_, PROT_READ, MAP_SHARED, fd_, 0);
    if (mp == MAP_FAILED) throw egr_error("mmap: " + path_ + ": " + strerror(errno));
    map_ = static_cast<const char*>(mp);
    // gathers touch scattered lists through the mapping (no readahead beyond what they
    // advise); streamed chunks are read in file order through the descriptor
    madvise(mp, file_bytes_, MADV_RANDOM);
    posix_fadvise(fd_, 0, 0, POSIX_FADV_SEQUENTIAL);
    cols_ = reinterpret_cast<const int32_t*>(map_ + cols_at_);

    if (wide && !offsets_in_ram) {
        off_ = reinterpret_cast<const uint64_t*>(map_ + offsets_at_);
        return;
    }
    ram_off_.resize((size_t)n + 1);
    if (wide) {
        read_range(ram_off_.data(), ram_off_.size() * sizeof(uint64_t), offsets_at_);
    } else {   // uint32 offsets, widened
        std::vector<uint32_t> narrow((size_t)n + 1);
        read_range(narrow.data(), narrow.size() * sizeof(uint32_t), offsets_at_);
        const long long cnt = (long long)narrow.size();
        uint64_t* dst = ram_off_.data();
        const uint32_t* src = narrow.data();
#pragma omp parallel for schedule(static) num_threads(io_threads_)
        for (long long i = 0; i < cnt; i++) dst[i] = src[i];
    }
    off_ = ram_off_.data();
}

// True when every page of [at, at + bytes) of the file is in the page cache (mincore over
// the mapping).
bool EgrGraph::resident(size_t at, size_t bytes) const {
    const size_t P = (size_t)sysconf(_SC_PAGESIZE);
    const size_t a0 = at / P * P, a1 = std::min(file_bytes_, at + bytes);
    if (a1 <= a0) return true;
    std::vector<unsigned char> in((a1 - a0 + P - 1) / P);   // mincore fills one byte per page begun
    if (mincore(const_cast<char*>(map_) + a0, a1 - a0, in.data()) != 0) return false;
    for (unsigned char b : in)
        if (!(b & 1)) return false;
    return true;
}

// The I/O pool: blocks of kBlock bytes are claimed from a shared counter by up to
// io_threads_ workers; the first failure stops the others and is rethrown here.
void EgrGraph::read_range(void* dst, size_t bytes, size_t at) const {
    if (!bytes) return;
    const size_t nblocks = (bytes + kBlock - 1) / kBlock;
    const int workers = (int)std::min<size_t>((size_t)io_threads_, nblocks);
    // A range already in the page cache is copied out of the mapping (one copy, no system
    // call per block: about twice the rate of pread into pinned memory on the MI355X hosts);
    // a range with pages on storage is read with pread, whose readahead serves it far better
    // than page faults on a MADV_RANDOM mapping.
    const bool from_map = copy_from_map_ && resident(at, bytes);
    if (workers == 1) {
        if (from_map) memcpy(dst, map_ + at, bytes);
        else pread_all(fd_, dst, bytes, at, path_);
        return;
    }
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    std::string first_error;
    std::mutex mu;
#pragma omp parallel num_threads(workers)
    {
        for (;;) {
            const size_t b = next.fetch_add(1, std::memory_order_relaxed);
            if (b >= nblocks || failed.load(std::memory_order_relaxed)) break;
            const size_t o = b * kBlock, len = std::min(kBlock, bytes - o);
            if (from_map) {
                memcpy(static_cast<char*>(dst) + o, map_ + at + o, len);
                continue;
            }
            try {
                pread_all(fd_, static_cast<char*>(dst) + o, len, at + o, path_);
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lock(mu);
                if (!failed.exchange(true)) first_error = e.what();
            }
        }
    }
    if (failed) throw egr_error(first_error);
}

void EgrGraph::read_offsets(size_t v0, size_t cnt, uint64_t* dst) const {
    if (off_ == ram_off_.data() && !ram_off_.empty()) std::copy(off_ + v0, off_ + v0 + cnt, dst);
    else read_range(dst, cnt * sizeof(uint64_t), offsets_at_ + v0 * sizeof(uint64_t));
}

void EgrGraph::read_edges(uint64_t e0, uint64_t e1, int32_t* dst) const {
    read_range(dst, (e1 - e0) * sizeof(int32_t), cols_at_ + e0 * sizeof(int32_t));
}

void EgrGraph::prefetch_lists(const int32_t* vids, int k, uint64_t* pages, uint64_t* runs, int groups) const {
    const size_t P = page_size();
    const int G = std::max(1, std::min(groups, k / 1024));
    std::vector<uint64_t> gp(G, 0), gr(G, 0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(std::max(1, std::min(G, io_threads_)))
    for (int g = 0; g < G; g++) {
        const int i0 = (int)((long long)k * g / G), i1 = (int)((long long)k * (g + 1) / G);
        RunMerger runs_of_group(map_, P);
        for (int i = i0; i < i1; i++) {
            const int v = vids[i];
            const uint64_t d = off_[v + 1] - off_[v];
            if (!d) continue;
            const size_t b0 = cols_at_ + off_[v] * sizeof(int32_t), b1 = b0 + d * sizeof(int32_t);
            runs_of_group.add(b0 / P, (b1 + P - 1) / P);
        }
        runs_of_group.close();
        gp[g] = runs_of_group.pages;
        gr[g] = runs_of_group.runs;
    }
    uint64_t tp = 0, tr = 0;
    for (int g = 0; g < G; g++) { tp += gp[g]; tr += gr[g]; }
    *pages = tp;
    *runs = tr;
}

// Byte range [b0, b1) widened to whole pages, dropped from this mapping (its pages would
// otherwise keep the cache entries) and from the page cache.
void EgrGraph::advise_out(size_t b0, size_t b1) const {
    const size_t P = page_size();
    const size_t a0 = b0 / P * P, a1 = std::min(file_bytes_, (b1 + P - 1) / P * P);
    if (a1 <= a0) return;
    madvise(const_cast<char*>(map_) + a0, a1 - a0, MADV_DONTNEED);
    posix_fadvise(fd_, (off_t)a0, (off_t)(a1 - a0), POSIX_FADV_DONTNEED);
}

void EgrGraph::drop_cache(uint64_t e0, uint64_t e1, size_t v0, size_t v1) const {
    advise_out(cols_at_ + e0 * sizeof(int32_t), cols_at_ + e1 * sizeof(int32_t));
    const size_t w = wide ? sizeof(uint64_t) : sizeof(uint32_t);
    advise_out(offsets_at_ + v0 * w, offsets_at_ + (v1 + 1) * w);
}

// Chunk end: the last vertex u in [v, n] with off[u] <= off[v] + max_edges, found by a
// galloping search from v (chunks are short next to n, so this stays near O(log chunk)).
std::vector<EdgeChunk> EgrGraph::plan_chunks(uint64_t max_edges) const {
    std::vector<EdgeChunk> out;
    const long long last = n;   // offsets are indexed 0..n
    for (long long v = 0; v < n;) {
        const uint64_t limit = off_[v] > UINT64_MAX - max_edges ? UINT64_MAX : off_[v] + max_edges;
        long long lo = v, step = 1;
        while (lo + step <= last && off_[lo + step] <= limit) {
            lo += step;
            step <<= 1;
        }
        long long hi = std::min(lo + step, last + 1);   // off[hi] > limit, or hi past the end
        while (hi - lo > 1) {
            const long long mid = lo + (hi - lo) / 2;
            if (off_[mid] <= limit) lo = mid; else hi = mid;
        }
        const long long v1 = lo > v ? lo : v + 1;
        out.push_back({(int)v, (int)v1, off_[v], off_[v1]});
        v = v1;
    }
    return out;
}

}  // namespace excent
