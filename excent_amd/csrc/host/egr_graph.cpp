// egr_graph.cpp -- see egr_graph.hpp.  Reference: include/StaticGraph.h:29-170.
#include "egr_graph.hpp"

#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>

namespace excent {

EgrGraph::EgrGraph(const std::string& path, int io_threads, bool offsets_in_ram)
    : io_threads_(io_threads), path_(path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) throw egr_error(path + ": " + strerror(errno));
    int32_t hdr[2];
    try {
        read_exact(hdr, sizeof hdr, 0);
    } catch (...) { ::close(fd_); fd_ = -1; throw; }
    n = hdr[0];
    size_t off = sizeof hdr;
    wide = hdr[1] == -1;
    if (wide) { read_exact(&m, sizeof m, off); off += sizeof m; }
    else m = (uint64_t)(int64_t)hdr[1];   // StaticGraph.h:38 converts the int to ull
    if (n < 1) { ::close(fd_); fd_ = -1; throw egr_error("StaticGraph: bad header in " + path); }
    idx_off_ = off;
    if (!wide) offsets_in_ram = true;   // narrow offsets are widened into RAM (StaticGraph.h:42)
    if (wide && !offsets_in_ram) {
        off += ((size_t)n + 1) * sizeof(uint64_t);
    } else if (wide) {
        nindex_.resize((size_t)n + 1);
        read_parallel(nindex_.data(), nindex_.size() * sizeof(uint64_t), off);
        off += nindex_.size() * sizeof(uint64_t);
    } else {
        nindex_.resize((size_t)n + 1);
        std::vector<int32_t> tmp((size_t)n + 1);
        read_parallel(tmp.data(), tmp.size() * sizeof(int32_t), off);
        off += tmp.size() * sizeof(int32_t);
        for (size_t i = 0; i < tmp.size(); i++) nindex_[i] = (uint64_t)(uint32_t)tmp[i];
    }
    nlist_off_ = off;
    struct stat st;
    if (fstat(fd_, &st) != 0) throw egr_error(path + ": fstat failed");
    map_size_ = (size_t)st.st_size;
    const size_t need = nlist_off_ + m * sizeof(int32_t);
    if (map_size_ < need) {
        ::close(fd_); fd_ = -1;
        throw egr_error("StaticGraph: " + path + " is shorter than its header says (" + std::to_string(map_size_) +
                        " < " + std::to_string(need) + " bytes)");
    }
    map_ = mmap(nullptr, map_size_, PROT_READ, MAP_SHARED, fd_, 0);
    if (map_ == MAP_FAILED) { map_ = nullptr; ::close(fd_); fd_ = -1; throw egr_error("mmap: " + std::string(strerror(errno))); }
    madvise(map_, map_size_, MADV_RANDOM);
    posix_fadvise(fd_, 0, 0, POSIX_FADV_SEQUENTIAL);
    nlist_ = (const int32_t*)((const char*)map_ + nlist_off_);
    off_ = offsets_in_ram ? nindex_.data() : (const uint64_t*)((const char*)map_ + idx_off_);
}

EgrGraph::~EgrGraph() {
    if (map_) munmap(map_, map_size_);
    if (fd_ >= 0) ::close(fd_);
}

void EgrGraph::read_exact(void* dst, size_t bytes, size_t off) const {
    char* p = (char*)dst;
    while (bytes) {
        ssize_t r = pread(fd_, p, bytes, (off_t)off);
        if (r <= 0) throw egr_error("pread: " + path_ + (r == 0 ? ": unexpected end of file" : ": " + std::string(strerror(errno))));
        p += r; off += (size_t)r; bytes -= (size_t)r;
    }
}

// StaticGraph.h:160-170: 8 MB pieces, one OpenMP thread each, up to io_threads.
void EgrGraph::read_parallel(void* dst, size_t bytes, size_t off) const {
    const size_t piece = 8ull << 20;
    const long long pieces = (long long)((bytes + piece - 1) / piece);
    const int threads = (int)std::min<long long>(io_threads_, std::max<long long>(1, pieces));
    std::string err;
#pragma omp parallel for num_threads(threads) schedule(static, 1)
    for (long long i = 0; i < pieces; i++) {
        const size_t b = (size_t)i * piece, len = std::min(piece, bytes - b);
        try {
            read_exact((char*)dst + b, len, off + b);
        } catch (const std::exception& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) throw egr_error(err);
}

void EgrGraph::read_offsets(size_t v0, size_t cnt, uint64_t* dst) const {
    if (!nindex_.empty()) memcpy(dst, nindex_.data() + v0, cnt * sizeof(uint64_t));
    else read_parallel(dst, cnt * sizeof(uint64_t), idx_off_ + v0 * sizeof(uint64_t));
}

void EgrGraph::read_edges(uint64_t e0, uint64_t e1, int32_t* dst) const {
    read_parallel(dst, (e1 - e0) * sizeof(int32_t), nlist_off_ + e0 * sizeof(int32_t));
}

// StaticGraph.h:90-114: the k lists are split into up to io_threads contiguous parts
// (at least 1024 lists each); each part merges overlapping or touching page runs and
// advises each run once.
void EgrGraph::prefetch_lists(const int32_t* vids, int k, uint64_t* pages, uint64_t* runs) const {
    const size_t P = (size_t)sysconf(_SC_PAGESIZE);
    uint64_t np = 0, nr = 0;
    const int parts = std::max(1, std::min(io_threads_, k / 1024));
#pragma omp parallel for num_threads(parts) reduction(+ : np, nr) schedule(static, 1)
    for (int t = 0; t < parts; t++) {
        const int i0 = (int)((long long)k * t / parts), i1 = (int)((long long)k * (t + 1) / parts);
        size_t r0 = 0, r1 = 0;
        bool open = false;
        auto flush = [&]() {
            if (!open) return;
            madvise((char*)map_ + r0 * P, (r1 - r0) * P, MADV_WILLNEED);
            np += r1 - r0;
            nr++;
        };
        for (int i = i0; i < i1; i++) {
            const uint64_t d = degree(vids[i]);
            if (!d) continue;
            const size_t b0 = nlist_off_ + off_[vids[i]] * sizeof(int32_t), b1 = b0 + d * sizeof(int32_t);
            const size_t p0 = b0 / P, p1 = (b1 + P - 1) / P;
            if (open && p0 <= r1) r1 = std::max(r1, p1);
            else { flush(); r0 = p0; r1 = p1; open = true; }
        }
        flush();
    }
    *pages = np;
    *runs = nr;
}

// StaticGraph.h:116-126
void EgrGraph::drop_cache(uint64_t e0, uint64_t e1, size_t v0, size_t v1) const {
    const size_t P = (size_t)sysconf(_SC_PAGESIZE);
    auto drop = [&](size_t b0, size_t b1) {
        const size_t a0 = b0 / P * P, a1 = std::min(map_size_, (b1 + P - 1) / P * P);
        if (a1 <= a0) return;
        madvise((char*)map_ + a0, a1 - a0, MADV_DONTNEED);
        posix_fadvise(fd_, (off_t)a0, (off_t)(a1 - a0), POSIX_FADV_DONTNEED);
    };
    drop(nlist_off_ + e0 * sizeof(int32_t), nlist_off_ + e1 * sizeof(int32_t));
    drop(idx_off_ + v0 * sizeof(uint64_t), idx_off_ + (v1 + 1) * sizeof(uint64_t));
}

// StaticGraph.h:128-139: greedy vertex-contiguous chunks of at most max_edges edges; a
// vertex larger than the cap gets a chunk of its own.
std::vector<EdgeChunk> EgrGraph::plan_chunks(uint64_t max_edges) const {
    std::vector<EdgeChunk> out;
    int v = 0;
    while (v < n) {
        const uint64_t limit = off_[v] + max_edges;
        int v1 = (int)(std::upper_bound(off_ + v + 1, off_ + n + 1, limit) - off_) - 1;
        if (v1 <= v) v1 = v + 1;
        out.push_back({v, v1, off_[v], off_[v1]});
        v = v1;
    }
    return out;
}

}  // namespace excent
