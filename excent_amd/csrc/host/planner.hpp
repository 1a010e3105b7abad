// planner.hpp -- host decisions of the pass loop: chunk cap, fetch plan, launch choice,
// group width, hot-set selection, tau from calibration timings, gather list build.
// Each is a pure function so the tests can hold it against the oracle bit for bit.
#pragma once

#include <cstdint>
#include <vector>

namespace excent {

enum FetchMode : int8_t { F_SKIP = 0, F_RESIDENT = 1, F_STREAM = 2, F_GATHER = 3 };

// src/excent.cu:321-326
uint64_t chunk_cap(uint64_t m, double chunk_medges, double target_chunks, double min_chunk_medges);

// src/excent.cu:238: next power of two >= avg, within [1, cap] (the reference caps at 32;
// the wave64 group kernel accepts 64).
int pick_group(double avg, int cap = 32);

// src/excent.cu:639
inline bool use_list(uint64_t k, int nv, double list_frac) { return (double)k < list_frac * nv; }

struct PlanOut {
    uint64_t pend_v = 0, demand = 0;
    bool need_bitmap = false;
};

// src/excent.cu:666-677
PlanOut fetch_plan(int nc, const uint64_t* cnt, const std::vector<char>& resident, const std::vector<char>& loaded,
                   double tau, const std::vector<uint64_t>& chunk_edges, bool hot_mode, int pass,
                   const std::vector<int>& hlo, const std::vector<int>& hhi, int8_t* mode);

struct HotSet {
    std::vector<int32_t> ids;
    std::vector<uint64_t> off{0};
    std::vector<uint32_t> bitmask;
    uint64_t deg_min = 0;
};

// The hot set of src/excent.cu:432-491 given the budget (the selection rules are stated in
// planner.cpp).  score may be null (degree rule) or n uint16 scores (--hot-rank rule).
HotSet hot_select(int n, const uint64_t* off, uint64_t budget, const uint16_t* score, int threads = 16);

// src/excent.cu:566-572
double tau_from_timings(double t_stream, const double* t_g, const double* f_got);

// src/excent.cu:711-720: pending vertices of [v0, v1) from the host bitmap, ascending.
int gather_list(const uint32_t* bitmap, int v0, int v1, int32_t* out);

}  // namespace excent
