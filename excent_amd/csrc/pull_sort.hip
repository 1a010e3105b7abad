// pull_sort.hip -- see pull_sort.hpp.  The sort is hipCUB's segmented radix sort (one
// segment per tile, keys = source rows, values = entry positions); the destinations and
// coefficients follow the values.
#include "pull_sort.hpp"

#include <hipcub/hipcub.hpp>

namespace excent_dev {
namespace {

__global__ void k_tile_starts(int n, int T, long long ntiles, const uint64_t* __restrict__ tidx, uint64_t* start) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i <= ntiles; i += (long long)gridDim.x * blockDim.x) {
        const long long v = i * T < n ? i * T : n;
        start[i] = tidx[v];
    }
}

// the destination (relative to its tile) of every entry, and the entry positions 0..E-1
__global__ void k_entry_dst(int n, int T, const uint64_t* __restrict__ tidx, uint16_t* edst, uint32_t* pos) {
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < n; u += (long long)gridDim.x * blockDim.x) {
        const uint16_t local = (uint16_t)(u % T);
        for (uint64_t e = tidx[u]; e < tidx[u + 1]; e++) {
            edst[e] = local;
            pos[e] = (uint32_t)e;
        }
    }
}

__global__ void k_follow(uint64_t cnt, const uint32_t* __restrict__ perm, const uint16_t* __restrict__ edst,
                         const float* __restrict__ tw, uint16_t* dst, float* w) {
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < cnt; q += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = perm[q];
        dst[q] = edst[e];
        if (tw) w[q] = tw[e];
    }
}

}  // namespace

void free_sorted_tiles(SortedTileBuffers* b) {
    void* q[] = {b->start, b->row, b->dst, b->w};
    for (void* x : q) if (x) (void)hipFree(x);
    *b = SortedTileBuffers{};
}

hipError_t build_sorted_tiles(int n, int T, const uint64_t* tidx, const int32_t* tadj, const float* tw,
                              uint64_t entries, int row_bits, hipStream_t st, SortedTileBuffers* out) {
    SortedTileBuffers b;
    const long long ntiles = ((long long)n + T - 1) / T;
    const uint64_t E = entries ? entries : 1;
    uint16_t* edst = nullptr;
    uint32_t *pos = nullptr, *perm = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
    step(hipMalloc(&b.start, (size_t)(ntiles + 1) * sizeof(uint64_t)));
    if (e == hipSuccess) step(hipMalloc(&b.row, E * sizeof(int32_t)));
    if (e == hipSuccess) step(hipMalloc(&b.dst, E * sizeof(uint16_t)));
    if (e == hipSuccess && tw) step(hipMalloc(&b.w, E * sizeof(float)));
    if (e == hipSuccess) step(hipMalloc(&edst, E * sizeof(uint16_t)));
    if (e == hipSuccess) step(hipMalloc(&pos, E * sizeof(uint32_t)));
    if (e == hipSuccess) step(hipMalloc(&perm, E * sizeof(uint32_t)));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tile_starts, dim3((unsigned)std::min<long long>((ntiles + 256) / 256, 65535)), dim3(256), 0, st,
                           n, T, ntiles, tidx, b.start);
        hipLaunchKernelGGL(k_entry_dst, dim3((unsigned)std::min<long long>(((long long)n + 255) / 256, 1 << 16)), dim3(256),
                           0, st, n, T, tidx, edst, pos);
        step(hipGetLastError());
    }
    if (e == hipSuccess && entries) {
        step(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, temp_bytes, tadj, b.row, pos, perm, (int)entries,
                                                         (int)ntiles, b.start, b.start + 1, 0, row_bits, st));
        if (e == hipSuccess) step(hipMalloc(&temp, temp_bytes ? temp_bytes : 1));
        if (e == hipSuccess)
            step(hipcub::DeviceSegmentedRadixSort::SortPairs(temp, temp_bytes, tadj, b.row, pos, perm, (int)entries,
                                                             (int)ntiles, b.start, b.start + 1, 0, row_bits, st));
        if (e == hipSuccess) {
            const uint64_t grid = std::min<uint64_t>((entries + 255) / 256, 1 << 16);
            hipLaunchKernelGGL(k_follow, dim3((unsigned)grid), dim3(256), 0, st, entries, perm, edst, tw, b.dst, b.w);
            step(hipGetLastError());
        }
    }
    if (e == hipSuccess) step(hipStreamSynchronize(st));
    void* tmp[] = {edst, pos, perm, temp};
    for (void* x : tmp) if (x) (void)hipFree(x);
    if (e != hipSuccess) {
        free_sorted_tiles(&b);
        return e;
    }
    *out = b;
    return hipSuccess;
}

hipError_t sort_runs(int32_t* keys, float* vals, uint64_t count, const std::vector<int>& starts, int row_bits,
                     hipStream_t st) {
    const int nseg = (int)starts.size() - 1;
    if (nseg < 1 || count == 0) return hipSuccess;
    int32_t* kout = nullptr;
    float* vout = nullptr;
    int* d_starts = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
    step(hipMalloc(&kout, count * sizeof(int32_t)));
    if (e == hipSuccess && vals) step(hipMalloc(&vout, count * sizeof(float)));
    if (e == hipSuccess) step(hipMalloc(&d_starts, starts.size() * sizeof(int)));
    if (e == hipSuccess) step(hipMemcpyAsync(d_starts, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice, st));
    auto sort = [&](void* t, size_t& b) {
        return vals ? hipcub::DeviceSegmentedRadixSort::SortPairs(t, b, keys, kout, vals, vout, (int)count, nseg, d_starts,
                                                                  d_starts + 1, 0, row_bits, st)
                    : hipcub::DeviceSegmentedRadixSort::SortKeys(t, b, keys, kout, (int)count, nseg, d_starts, d_starts + 1,
                                                                 0, row_bits, st);
    };
    if (e == hipSuccess) step(sort(nullptr, temp_bytes));
    if (e == hipSuccess) step(hipMalloc(&temp, temp_bytes ? temp_bytes : 1));
    if (e == hipSuccess) step(sort(temp, temp_bytes));
    if (e == hipSuccess) step(hipMemcpyAsync(keys, kout, count * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    if (e == hipSuccess && vals) step(hipMemcpyAsync(vals, vout, count * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (e == hipSuccess) step(hipStreamSynchronize(st));
    void* tmp[] = {kout, vout, d_starts, temp};
    for (void* x : tmp) if (x) (void)hipFree(x);
    return e;
}

}  // namespace excent_dev
