// harness.cpp -- bin/egr_info, bin/evict_file and bin/gpu_ballast: the measurement harness
// around bin/excent (a graph's shape and symmetry, a cold page cache for one file, a capped
// device memory budget held by a second process).
#include <fcntl.h>
#include <omp.h>
#include <signal.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "egr_graph.hpp"
#include "excent.h"
#include "excent_hip.h"

namespace {

// Share of a file's pages in the page cache (mincore over a read-only mapping); -1 when
// the file cannot be mapped.
double cached_share(int fd, size_t bytes) {
    if (bytes == 0) return 0.0;
    void* p = mmap(nullptr, bytes, PROT_READ, MAP_SHARED, fd, 0);
    if (p == MAP_FAILED) return -1.0;
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    const size_t pages = (bytes + page - 1) / page;
    std::vector<unsigned char> in(pages);
    double share = -1.0;
    if (mincore(p, bytes, in.data()) == 0) {
        size_t resident = 0;
        for (unsigned char c : in) resident += c & 1u;
        share = (double)resident / (double)pages;
    }
    munmap(p, bytes);
    return share;
}

volatile sig_atomic_t g_stop = 0;
void on_stop(int) { g_stop = 1; }

}  // namespace

// bin/egr_info <graph.egr> (tools/egr_info.cpp:10-29): n, m, the edge bytes, the largest
// out-degree, and a symmetry probe: 20000 edges drawn uniformly (std::mt19937_64 seeded with
// 7, edge = draw mod m, the draws of the reference tool so the counts agree), each checked for
// its reverse edge (binary search when the list is sorted, else a scan), self-loops counted.
extern "C" int excent_egr_info_main(int argc, const char* const* argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <graph.egr>\n", argc ? argv[0] : "egr_info");
        return 1;
    }
    try {
        const excent::EgrGraph g(argv[1]);
        const uint64_t* off = g.offsets();
        const int32_t* cols = g.nlist();
        const long long n = g.n;
        uint64_t maxdeg = 0;
#pragma omp parallel for reduction(max : maxdeg) schedule(static)
        for (long long v = 0; v < n; v++) maxdeg = std::max(maxdeg, off[v + 1] - off[v]);
        constexpr int kSamples = 20000;
        int missing = 0, loops = 0;
        if (g.m > 0) {
            std::mt19937_64 draw(7);
            for (int i = 0; i < kSamples; i++) {
                const uint64_t e = draw() % g.m;
                const int src = (int)(std::upper_bound(off, off + n + 1, e) - off) - 1;   // off[src] <= e < off[src+1]
                const int dst = cols[e];
                if (src == dst) loops++;
                const int32_t* lo = cols + off[dst];
                const int32_t* hi = cols + off[dst + 1];
                const bool found = std::is_sorted(lo, hi) ? std::binary_search(lo, hi, src) : std::find(lo, hi, src) != hi;
                if (!found) missing++;
            }
        }
        const int drawn = g.m > 0 ? kSamples : 0;
        printf("n=%d m=%llu (%.2f GB edges) maxdeg=%llu  reverse-edge missing %d/%d (%.1f%%) self-loops %d  m>2^32-1: %s\n",
               g.n, (unsigned long long)g.m, g.m * 4.0 / 1e9, (unsigned long long)maxdeg, missing, drawn,
               drawn ? 100.0 * missing / drawn : 0.0, loops, g.m > 4294967295ull ? "YES" : "no");
        fflush(stdout);
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}

// bin/evict_file [--report] <file>... (tools/evict_file.c:36-65): drops each file's pages
// from the page cache without root (fdatasync, then POSIX_FADV_DONTNEED) and prints the
// resident share before and after; exit status 1 when a file cannot be opened, the advice
// fails or more than 1 % stays resident.  --report only prints each file's resident %.
extern "C" int excent_evict_main(int argc, const char* const* argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s [--report] <file> [...]\n", argc ? argv[0] : "evict_file");
        return 1;
    }
    const bool report = strcmp(argv[1], "--report") == 0;
    int rc = 0;
    for (int i = report ? 2 : 1; i < argc; i++) {
        const int fd = open(argv[i], O_RDONLY | O_CLOEXEC);
        if (fd < 0) {
            perror(argv[i]);
            rc = 1;
            continue;
        }
        struct stat st;
        const size_t bytes = fstat(fd, &st) == 0 ? (size_t)st.st_size : 0;
        if (report) {
            printf("%.1f\n", 100.0 * cached_share(fd, bytes));
        } else {
            const double before = cached_share(fd, bytes);
            fdatasync(fd);
            const int err = posix_fadvise(fd, 0, 0, POSIX_FADV_DONTNEED);
            const double after = cached_share(fd, bytes);
            printf("%s: %.1f%% -> %.1f%% resident%s\n", argv[i], 100.0 * before, 100.0 * after,
                   err ? " (fadvise failed)" : "");
            if (err || after > 0.01) rc = 1;
        }
        close(fd);
    }
    fflush(stdout);
    return rc;
}

// bin/gpu_ballast <budget_GB> [--device D] (tools/gpu_ballast.cu:13-41): holds device memory
// until only budget_GB (GiB) stay free, prints "READY free=... GB (budget ... GB) of ... GB"
// and keeps it until SIGINT or SIGTERM, so a bin/excent run in another process sees a capped
// budget.  Exit status 1 when no more than the budget is free to begin with.
extern "C" int excent_ballast_main(int argc, const char* const* argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <budget_GB> [--device D]\n", argc ? argv[0] : "gpu_ballast");
        return 1;
    }
    const double gib = 1073741824.0;
    const size_t budget = (size_t)(atof(argv[1]) * gib);
    int device = 0;
    for (int i = 2; i + 1 < argc; i++)
        if (!strcmp(argv[i], "--device")) device = atoi(argv[++i]);
    // the stop signals stay blocked except inside sigsuspend, so one that arrives before the
    // wait is taken there rather than lost
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_handler = on_stop;
    sigaction(SIGINT, &sa, nullptr);
    sigaction(SIGTERM, &sa, nullptr);
    sigset_t stops, wait_mask;
    sigemptyset(&stops);
    sigaddset(&stops, SIGINT);
    sigaddset(&stops, SIGTERM);
    sigprocmask(SIG_BLOCK, &stops, &wait_mask);
    sigdelset(&wait_mask, SIGINT);
    sigdelset(&wait_mask, SIGTERM);
    size_t freeB = 0, totalB = 0;
    if (exc_mem_info(device, &freeB, &totalB) != EXC_OK) {
        fprintf(stderr, "%s\n", exc_last_error());
        return 1;
    }
    if (freeB <= budget) {
        fprintf(stderr, "only %.2f GB free, below budget\n", freeB / gib);
        return 1;
    }
    void* hold = nullptr;
    if (exc_hold_memory(device, budget, &hold, nullptr) != EXC_OK) {
        fprintf(stderr, "%s\n", exc_last_error());
        return 1;
    }
    exc_mem_info(device, &freeB, &totalB);
    printf("READY free=%.2f GB (budget %.2f GB) of %.2f GB\n", freeB / gib, budget / gib, totalB / gib);
    fflush(stdout);
    while (!g_stop) sigsuspend(&wait_mask);
    return exc_release_memory(hold) == EXC_OK ? 0 : 1;
}
