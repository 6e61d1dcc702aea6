// Device-side plans of libcof: the plain scan (plain.hip) and the bulge
// pipeline (bulge.hip).  Host-callable; all device memory is owned here.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cof_internal.hpp"

namespace cof {

// Device scratch shared by all searches of one context: grown on demand,
// kept across searches and re-queries, freed with the context.
struct Workspace {
    void* plain_out = nullptr;     // PlainRaw records, 64 segments of plain_cap / 64
    uint32_t plain_cap = 0;
    uint32_t plain_cap_tuned = 0;  // the PlainTune::hit_cap the buffer was last sized to
    uint32_t* plain_count = nullptr;   // per-segment counters, 128 bytes apart
    uint32_t* plain_heads = nullptr;   // per-XCD tile queue heads of the persistent scan, 128 bytes apart
    void* plain_out2 = nullptr;    // the segments packed together
    uint32_t plain_cap2 = 0;
    uint32_t* h_counts = nullptr;  // pinned copy of the segment counters
    size_t h_counts_cap = 0;
    bool plain_pending = false;    // a launched scan awaits plain_finish
    uint64_t plain_scan_end = 0;
    // scan image of the resident genome (plain.hip): its two bit planes and
    // one flag byte per 4096-start unit, built by the first plain scan after
    // a genome load and reused by later scans
    uint32_t* img_hi = nullptr;
    uint32_t* img_lo = nullptr;
    uint8_t* img_flags = nullptr;
    uint64_t img_units = 0;        // units held (0: no valid image)
    uint64_t img_cap_units = 0;    // units allocated
    uint64_t genome_nib = 0;       // nibbles of the resident genome (the image covers them all)
    void* queue = nullptr;         // bulge candidates
    uint32_t q_cap = 0;
    void* bulge_out = nullptr;     // BulgeRaw records
    uint32_t bulge_cap = 0;
    void* tiles = nullptr;         // bulge filter work items
    size_t tiles_cap = 0;
    void* chunks = nullptr;        // chunk table for the device drop rule
    size_t chunks_cap = 0;
    uint32_t* counters = nullptr;  // [0] queue count, [1] record count, [2] queue cursor
    unsigned long long* sums = nullptr;  // [0] kept, [1] checksum, [2..4] diagnostics, [5] traced candidates
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int cus = 0;
    // bulge pipeline
    void* qs[2] = {nullptr, nullptr};          // double-buffered candidate queues
    size_t qs_cap[2] = {0, 0};
    uint32_t* qsteps[2] = {nullptr, nullptr};  // DFS transitions per queue entry (tile stats only)
    size_t qsteps_cap[2] = {0, 0};
    uint32_t* chunk_lut = nullptr;             // chunk of nibble k << 16
    size_t chunk_lut_cap = 0;
    void* q_first = nullptr;                   // queue of the first round
    size_t q_first_cap = 0;
    uint32_t* q_first_steps = nullptr;
    size_t q_first_steps_cap = 0;
    uint32_t* tile_order = nullptr;            // tile processing order
    size_t tile_order_cap = 0;
    uint32_t* tile_cand = nullptr;             // candidates per tile (last search)
    size_t tile_cand_cap = 0;
    uint32_t* tile_cand_prev = nullptr;        // the first window class's counts while the second runs
    size_t tile_cand_prev_cap = 0;
    unsigned long long* tile_steps = nullptr;  // DFS transitions per tile (last search)
    size_t tile_steps_cap = 0;
    uint32_t last_tiles = 0;
    std::vector<uint64_t> tile_j0;             // host copy of the last search's tiles
    std::vector<uint32_t> tile_n;
    std::vector<uint32_t> tile_chunk;
    uint32_t* rounds = nullptr;                // per round: queue count, cursor, overflow
    size_t rounds_cap = 0;
    uint32_t* round_ctr = nullptr;             // synchronous round: cursor, overflow
    void* sq[2] = {nullptr, nullptr};          // screened queues (one per traceback stream)
    size_t sq_cap[2] = {0, 0};
    uint32_t* screen_ctr = nullptr;            // screened queue counts, one per sq[]
    std::vector<hipEvent_t> rev;               // per round: filtered, verified
    hipStream_t st2 = nullptr, st3 = nullptr;  // traceback streams of the pipeline
    hipGraphExec_t bulge_exec[3] = {};         // per window class: the last pipeline graph, updated in place
    ~Workspace();
    void ensure_events();
    void ensure_round_events(uint32_t rounds);
    // pinned host buffers for raw hits (grown, kept across searches)
    void* h_plain = nullptr;
    size_t h_plain_cap = 0;
    void* h_bulge = nullptr;
    size_t h_bulge_cap = 0;
    void ensure_host(void** p, size_t* cap, size_t bytes);
};
// grows *ptr to hold `need` elements of `elem` bytes (contents not kept)
void ws_grow(void** ptr, size_t* cap, size_t need, size_t elem);

// Launch shape of the plain scan (results do not depend on it).
struct PlainTune {
    uint32_t threads = 64;      // workgroup size; a workgroup scans 64 starts per thread
    uint32_t pat_chunk = 0;     // patterns per workgroup (grid.y slices a group); 0 = whole group
    uint32_t hit_cap = 0;       // hit records the device buffer starts with (0: 2^20, grown on overflow)
    uint32_t queue = 0;         // 1: persistent workgroups draining per-XCD tile queues (64 threads, whole groups)
};

// Schedule of the bulge path (results do not depend on it).
struct BulgeTune {
    uint32_t first_tiles = 1024;    // filter tiles in the synchronous first round
    uint32_t round_tiles = 0;       // tiles per pipeline round; 0 = sized from the first round
    uint32_t verify_per_cu = 24;    // persistent traceback workgroups per CU
    uint32_t queue_cap = 1u << 27;  // candidates per round before it is redone in halves
    uint32_t refill = 16;           // traceback wave refills once this many lanes are idle
    uint32_t screen = 1;            // 1: budget screen between the filter and the traceback
    uint32_t flags = 0;             // COF_FLAG_NO_GRAPH / COF_FLAG_SYNC_ROUNDS
    // predicted traceback work per genome chunk (cost model); null: static
    // genome order and rounds of equal tile counts
    const std::vector<float>* chunk_cost = nullptr;
};

// Device TSV formatter (csrc/format.hip): batch buffers kept across calls.
struct TsvFormatter {
    cof_hit* d_hits = nullptr;
    size_t hits_cap = 0;
    HitAux* d_aux = nullptr;
    size_t aux_cap = 0;
    unsigned long long* d_bsum = nullptr;   // per-workgroup row bytes
    size_t bsum_cap = 0;
    unsigned long long* d_boff = nullptr;   // their exclusive prefix + batch total
    size_t boff_cap = 0;
    uint8_t* d_out = nullptr;
    size_t out_cap = 0;
    uint8_t* d_pats = nullptr;
    size_t pats_cap = 0;
    uint8_t* d_blob = nullptr;
    size_t blob_cap = 0;
    uint8_t* d_names = nullptr;
    size_t names_cap = 0;
    uint64_t* d_name_off = nullptr;
    size_t name_off_cap = 0;
    uint32_t* d_name_len = nullptr;
    size_t name_len_cap = 0;
    uint8_t* h_out[2] = {nullptr, nullptr};  // pinned, double-buffered
    size_t h_cap = 0;
    unsigned long long* h_total = nullptr;   // pinned batch totals
    hipEvent_t ev_out[2] = {nullptr, nullptr};
    hipEvent_t ev_total = nullptr;
    ~TsvFormatter();
};
// Writes the header and one reference row per hit (main.rs:130-247), rows
// formatted on the device from the hits and the resident genome; returns the
// number of rows.
uint64_t write_tsv_device(const Genome& g, const Query& q, const HitVec& hits, const AuxVec& aux,
                          const uint32_t* d_genome, FILE* f, hipStream_t st, TsvFormatter& fm);

// FASTA ingest on the device (csrc/ingest.hip): fills g (chunks, names, host
// copy of the packed genome) and encodes into alloc(n_bytes, user); returns
// false, with g empty, when the file's line layout is irregular.
bool read_fasta_device(const std::string& path, Genome& g, uint32_t* alloc(uint64_t n_bytes, void* user), void* user,
                       hipStream_t st);
// 2bit ingest on the device (csrc/ingest.hip): the host walks the record
// headers and N blocks, the file streams to the device and one kernel decodes
// every 64 Knt piece into alloc(n_bytes, user).  Always returns true.
bool read_2bit_device(const std::string& path, Genome& g, uint32_t* alloc(uint64_t n_bytes, void* user), void* user,
                      hipStream_t st);

struct PlainPlan;
PlainPlan* plain_plan_create(const Query& q);
void plain_plan_destroy(PlainPlan* pp);
bool plain_plan_empty(const PlainPlan* pp);
uint64_t plain_tile_nucl();
uint64_t plain_pad_words();
void plain_run(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
               RawView<PlainRaw>& out, double* kernel_ms, uint64_t* n_launches, const PlainTune& tune);
// plain_run in two halves: plain_launch queues the scan and returns;
// plain_finish waits for it and fills `out` (the host is free in between).
void plain_launch(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
                  uint64_t* n_launches, const PlainTune& tune);
void plain_finish(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, hipStream_t st, RawView<PlainRaw>& out,
                  double* kernel_ms, uint64_t* n_launches, const PlainTune& tune);

// Allocation-free plain search (cof_search_static).  plain_static_init, once
// per device with that device current, looks up the device's copy of the
// static scratch array; plain_search_static then scans with the hit segments,
// counters and packed hits there (throws COF_E_NOMEM when a hit segment would
// overflow), copies the counters to h_counts (kHitSegs * 32 words) and the
// hits to h_stage (plain_static_max_hits() records) and points `out` at them.
// It needs the genome's scan image, which plain_prepare_image builds when a
// genome is loaded.
constexpr uint32_t kMaxStaticDevices = 64;
void plain_static_init(int device);
uint64_t plain_static_max_hits();
void plain_search_static(PlainPlan* pp, const Workspace& ws, const uint32_t* d_genome, uint64_t scan_end,
                         hipStream_t st, const PlainTune& tune, uint32_t* h_counts, PlainRaw* h_stage,
                         RawView<PlainRaw>& out, uint64_t* n_launches);
void plain_prepare_image(Workspace& ws, const uint32_t* d_genome, hipStream_t st);

struct BulgePlan;
BulgePlan* bulge_plan_create(const Query& q);
// count-only mode: device applies the chromosome-end drop rule and folds hits
// into stats->bulge_hits / bulge_checksum instead of returning records
void bulge_plan_destroy(BulgePlan* bp);
uint64_t bulge_pad_words();
size_t bulge_plan_size(const BulgePlan* bp);   // patterns of the plan
void bulge_run(BulgePlan* bp, Workspace& ws, const Genome& g, const std::vector<Batch>& batches,
               const uint32_t* d_genome, hipStream_t st, RawView<BulgeRaw>& out, cof_stats* stats,
               const BulgeTune& tune);

}  // namespace cof
