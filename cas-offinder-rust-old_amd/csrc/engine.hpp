// Device-side plans of libcof: the plain scan (plain.hip) and the bulge
// pipeline (bulge.hip).  Host-callable; all device memory is owned here.
#pragma once
#include <hip/hip_runtime.h>

#include "cof_internal.hpp"

namespace cof {

// Device scratch shared by all searches of one context: grown on demand,
// kept across searches and re-queries, freed with the context.
struct Workspace {
    void* plain_out = nullptr;     // PlainRaw records
    uint32_t plain_cap = 0;
    uint32_t* plain_count = nullptr;
    void* queue = nullptr;         // bulge candidates
    uint32_t q_cap = 0;
    void* bulge_out = nullptr;     // BulgeRaw records
    uint32_t bulge_cap = 0;
    void* tiles = nullptr;         // bulge filter work items
    size_t tiles_cap = 0;
    void* chunks = nullptr;        // chunk table for the device drop rule
    size_t chunks_cap = 0;
    uint32_t* counters = nullptr;  // [0] queue count, [1] record count, [2] queue cursor
    unsigned long long* sums = nullptr;  // [0] kept, [1] checksum, [2..3] diagnostics
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int cus = 0;
    ~Workspace();
    void ensure_events();
};
// grows *ptr to hold `need` elements of `elem` bytes (contents not kept)
void ws_grow(void** ptr, size_t* cap, size_t need, size_t elem);

struct PlainPlan;
PlainPlan* plain_plan_create(const Query& q);
void plain_plan_destroy(PlainPlan* pp);
bool plain_plan_empty(const PlainPlan* pp);
uint64_t plain_tile_nucl();
uint64_t plain_pad_words();
void plain_run(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
               std::vector<PlainRaw>& out, double* kernel_ms, uint64_t* n_launches);

struct BulgePlan;
BulgePlan* bulge_plan_create(const Query& q);
// count-only mode: device applies the chromosome-end drop rule and folds hits
// into stats->bulge_hits / bulge_checksum instead of returning records
void bulge_plan_destroy(BulgePlan* bp);
uint64_t bulge_pad_words();
void bulge_run(BulgePlan* bp, Workspace& ws, const Genome& g, const std::vector<Batch>& batches,
               const uint32_t* d_genome, hipStream_t st, std::vector<BulgeRaw>& out, cof_stats* stats);

}  // namespace cof
