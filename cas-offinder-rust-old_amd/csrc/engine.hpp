// Device-side plans of libcof: the plain scan (plain.hip) and the bulge
// pipeline (bulge.hip).  Host-callable; all device memory is owned here.
#pragma once
#include <hip/hip_runtime.h>

#include "cof_internal.hpp"

namespace cof {

struct PlainPlan;
PlainPlan* plain_plan_create(const Query& q);
void plain_plan_destroy(PlainPlan* pp);
bool plain_plan_empty(const PlainPlan* pp);
uint64_t plain_tile_nucl();
uint64_t plain_pad_words();
void plain_run(PlainPlan* pp, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
               std::vector<PlainRaw>& out, double* kernel_ms, uint64_t* n_launches);

struct BulgePlan;
BulgePlan* bulge_plan_create(const Query& q);
// count-only mode: device applies the chromosome-end drop rule and folds hits
// into stats->bulge_hits / bulge_checksum instead of returning records
void bulge_plan_destroy(BulgePlan* bp);
uint64_t bulge_pad_words();
void bulge_run(BulgePlan* bp, const Genome& g, const std::vector<Batch>& batches, const uint32_t* d_genome,
               hipStream_t st, std::vector<BulgeRaw>& out, cof_stats* stats);

}  // namespace cof
