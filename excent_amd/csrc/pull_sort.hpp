// pull_sort.hpp -- setup of the source-sorted dense pull (kernels.hip): the non-hub in-edges
// of each tile of T consecutive destinations, sorted by source row.  Internal to
// libexcent_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace excent_dev {

struct SortedTileBuffers {
    uint64_t* start = nullptr;   // [ntiles + 1]: entries of tile i in [start[i], start[i + 1])
    int32_t* row = nullptr;      // p row of each entry, ascending within its tile
    uint16_t* dst = nullptr;     // the entry's destination minus the tile's first vertex
    float* w = nullptr;          // weighted graphs: the entry's coefficient (else null)
};

// The region [0, entries) of the transpose (tidx[n + 1] offsets over tadj, rows of p) holds
// the in-lists of the non-hub destinations in vertex order; tw (or null) their
// coefficients.  Row ids are < 2^row_bits.  Returns hipSuccess or the first error (buffers
// allocated so far are freed on failure); the work runs on `st` and is waited for.
hipError_t build_sorted_tiles(int n, int T, const uint64_t* tidx, const int32_t* tadj, const float* tw,
                              uint64_t entries, int row_bits, hipStream_t st, SortedTileBuffers* out);
void free_sorted_tiles(SortedTileBuffers* b);

// Sorts each run [starts[i], starts[i + 1]) of keys (row ids < 2^row_bits) ascending, in
// place, with vals (or null) following its keys: the hub in-lists, so that a wave's
// consecutive entries gather neighbouring rows.  starts: host offsets, starts[0] = 0, the
// last = count < 2^31.  Temporary buffers are freed before returning; the work runs on `st`
// and is waited for.
hipError_t sort_runs(int32_t* keys, float* vals, uint64_t count, const std::vector<int>& starts, int row_bits,
                     hipStream_t st);

}  // namespace excent_dev
