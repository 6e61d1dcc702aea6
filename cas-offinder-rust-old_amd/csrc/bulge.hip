// Bulge search (D>0 or R>0) on gfx950: placeholder until the filter and
// traceback kernels land; queries with bulges are rejected.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "hip_check.hpp"

namespace cof {

struct BulgePlan {
    bool any = false;
};

BulgePlan* bulge_plan_create(const Query& q) {
    if (q.any_bulge) throw Error(COF_E_UNSUPPORTED, "bulge search not built yet");
    return new BulgePlan();
}
void bulge_plan_destroy(BulgePlan* bp) { delete bp; }
uint64_t bulge_pad_words() { return 64; }
void bulge_run(BulgePlan*, const Genome&, const std::vector<Batch>&, const uint32_t*, hipStream_t,
               std::vector<BulgeRaw>& out, cof_stats*) {
    out.clear();
}

}  // namespace cof
