// C ABI of libcof (include/cof.h): context, genome residency, search driver.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <sys/stat.h>

#include "cof_internal.hpp"
#include "costmodel.hpp"
#include "engine.hpp"
#include "hip_check.hpp"

struct cof_mlp {
    std::unique_ptr<cof::Mlp> m;
};

struct cof_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    cof::Query q;
    cof::Genome g;
    std::vector<cof::Batch> batches;
    uint32_t* d_genome = nullptr;
    uint64_t d_genome_words = 0;
    cof::PlainPlan* plain = nullptr;
    cof::BulgePlan* bulge = nullptr;
    cof::Workspace ws;
    cof::PlainTune plain_tune;
    // cost model: parameters (empty: static order), device copies, features
    std::vector<float> model;
    cof::FeaturePlan* feat = nullptr;
    cof::DevBuf d_model, d_feat, d_feat_chunks, d_pred;
    std::vector<float> chunk_cost;
    cof::BulgeTune bulge_tune;
    bool have_genome = false;
    // the genome is one search batch (cof_search_batch) instead of the
    // reference batcher's layout over the whole genome: -1 no, else has_head
    int single_batch = -1;
    // chunks of a genome being streamed in (cof_stream_*)
    bool streaming = false;
    cof::Genome stream_g;
    // allocation-free plain search (cof_search_static): the batch layout's
    // index, and host buffers sized when the context opens
    cof::BatchIndex bidx;
    std::unique_ptr<cof::PlainRaw[]> static_stage;
    uint32_t static_counts[64 * 32];
    bool cached = false;
    bool pending = false;          // between cof_search_async and cof_search_wait
    std::chrono::steady_clock::time_point search_t0;
    cof::HitVec hits;
    cof::AuxVec aux;
    cof::TsvFormatter fmt;
    cof_stats stats{};
    double load_ms = 0;
    bool load_device = false;
    std::string err;
};

namespace {

cof_status fail(cof_ctx* ctx, const cof::Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
}

// Entry points other than cof_search_wait refuse a context whose search is in flight.
cof_status refuse_pending(cof_ctx* ctx) {
    ctx->err = "a search started by cof_search_async is in flight: call cof_search_wait first";
    return COF_E_INVALID;
}

// Device buffer for a packed genome of n_bytes, zero-padded past the end so
// kernels need no bounds checks (one plain tile + halo, bulge window).
uint32_t* alloc_genome(cof_ctx* ctx, uint64_t n_bytes) {
    if (ctx->d_genome) { (void)hipFree(ctx->d_genome); ctx->d_genome = nullptr; }
    uint64_t words = (n_bytes + 3) / 4;
    uint64_t tile = cof::plain_tile_nucl();
    uint64_t n_tiles = (n_bytes * 2 + tile - 1) / tile + 1;
    uint64_t alloc_words = std::max<uint64_t>(words, n_tiles * tile / 8) + cof::plain_pad_words() + cof::bulge_pad_words();
    HIPCHECK(hipMalloc(&ctx->d_genome, alloc_words * 4));
    HIPCHECK(hipMemsetAsync(ctx->d_genome, 0, alloc_words * 4, ctx->stream));
    ctx->d_genome_words = alloc_words;
    return ctx->d_genome;
}

std::vector<cof::Batch> layout_batches(const cof_ctx* ctx) {
    if (ctx->single_batch < 0) return cof::make_batches(ctx->g.chunks.size(), ctx->q.batch_nucl);
    if (ctx->g.chunks.empty()) return {};
    return {cof::Batch{0, ctx->g.chunks.size(), ctx->single_batch > 0}};
}

void set_batches(cof_ctx* ctx) {
    ctx->batches = layout_batches(ctx);
    ctx->bidx = cof::BatchIndex(ctx->g, ctx->batches);
}

void finish_genome(cof_ctx* ctx) {
    ctx->ws.img_units = 0;   // the scan image follows the new genome
    ctx->ws.genome_nib = ctx->g.n_nibbles();
    ctx->g.index();
    set_batches(ctx);
    // the genome's scan image is part of making it resident, so that a plain
    // search (cof_search_static included) finds it built
    cof::plain_prepare_image(ctx->ws, ctx->d_genome, ctx->stream);
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->have_genome = true;
    ctx->cached = false;
    ctx->stats.genome_nucl = 0;
    for (auto& c : ctx->g.chunks) ctx->stats.genome_nucl += c.size();
}

void upload_genome(cof_ctx* ctx) {
    alloc_genome(ctx, ctx->g.n_bytes);
    if (ctx->g.n_bytes)
        HIPCHECK(hipMemcpyAsync(ctx->d_genome, ctx->g.packed.data(), ctx->g.n_bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    finish_genome(ctx);
}

// First byte of a regular file (-1: not a regular file or empty): '>' is
// FASTA, anything else is read as 2bit (main.rs:106-121).
int first_byte(const char* path) {
    struct stat sb;
    if (stat(path, &sb) != 0 || !S_ISREG(sb.st_mode)) return -1;
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int c = fgetc(f);
    fclose(f);
    return c;
}

// genome file -> resident genome: device ingest for FASTA files with regular
// line widths and for 2bit files unless the host reader is asked for; the
// host reader for folders and irregular FASTA files
cof_status load_file(cof_ctx* ctx, const char* path, uint32_t flags) {
    if (!ctx || !path) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    ctx->streaming = false;
    ctx->single_batch = -1;
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        bool device = false;
        const int c = (flags & COF_LOAD_HOST) ? -1 : first_byte(path);
        if (c >= 0) {
            auto alloc = [](uint64_t n, void* u) { return alloc_genome((cof_ctx*)u, n); };
            device = c == '>' ? cof::read_fasta_device(path, ctx->g, alloc, ctx, ctx->stream)
                              : cof::read_2bit_device(path, ctx->g, alloc, ctx, ctx->stream);
            if (device) finish_genome(ctx);
        }
        if (!device) {
            if (flags & COF_LOAD_DEVICE)
                throw cof::Error(COF_E_UNSUPPORTED,
                                 "device ingest needs a 2bit file or a FASTA file with regular line widths");
            cof::read_genome_file(path, ctx->g);
            upload_genome(ctx);
        }
        ctx->load_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ctx->load_device = device;
    } catch (const cof::Error& e) {
        ctx->have_genome = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

// chunk features of the resident genome for the query, on the device
void compute_features(cof_ctx* ctx) {
    if (!ctx->feat) ctx->feat = cof::feature_plan_create(ctx->q);
    const size_t nc = ctx->g.chunks.size();
    ctx->d_feat.ensure(std::max<size_t>(1, nc) * cof::kMlpIn * sizeof(float));
    cof::chunk_features(ctx->feat, ctx->g, ctx->d_genome, ctx->stream, ctx->d_feat_chunks, (float*)ctx->d_feat.p);
}

// predicted traceback work per chunk: exp(model(features)) - 1
void predict_chunk_cost(cof_ctx* ctx) {
    const size_t nc = ctx->g.chunks.size();
    compute_features(ctx);
    ctx->d_pred.ensure(std::max<size_t>(1, nc) * sizeof(float));
    cof::mlp_forward_dev((const float*)ctx->d_model.p, (const float*)ctx->d_feat.p, (uint32_t)nc,
                         (float*)ctx->d_pred.p, ctx->stream);
    ctx->chunk_cost.resize(nc);
    if (nc)
        HIPCHECK(hipMemcpyAsync(ctx->chunk_cost.data(), ctx->d_pred.p, nc * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (float& c : ctx->chunk_cost) c = std::expm1(std::min(c, 60.f));
}

cof_tuning get_tuning(const cof_ctx* ctx) {
    cof_tuning t{};
    t.plain_threads = ctx->plain_tune.threads;
    t.plain_pat_chunk = ctx->plain_tune.pat_chunk;
    t.bulge_first_tiles = ctx->bulge_tune.first_tiles;
    t.bulge_round_tiles = ctx->bulge_tune.round_tiles;
    t.verify_blocks_per_cu = ctx->bulge_tune.verify_per_cu;
    t.bulge_queue_cap = ctx->bulge_tune.queue_cap;
    t.verify_refill = ctx->bulge_tune.refill;
    t.plain_hit_cap = ctx->plain_tune.hit_cap;
    t.plain_queue = ctx->plain_tune.queue;
    t.bulge_screen = ctx->bulge_tune.screen ? 1u : 2u;
    return t;
}

void set_tuning(cof_ctx* ctx, const cof_tuning& t) {
    cof::PlainTune pt;
    cof::BulgeTune bt;
    if (t.plain_threads) {
        if (t.plain_threads != 64 && t.plain_threads != 128 && t.plain_threads != 256 && t.plain_threads != 512)
            throw cof::Error(COF_E_INVALID, "plain_threads must be 64, 128, 256 or 512");
        pt.threads = t.plain_threads;
    }
    pt.pat_chunk = t.plain_pat_chunk;
    if (t.plain_hit_cap % 64) throw cof::Error(COF_E_INVALID, "plain_hit_cap must be a multiple of 64");
    pt.hit_cap = t.plain_hit_cap;
    if (t.plain_queue > 1) throw cof::Error(COF_E_INVALID, "plain_queue must be 0 or 1");
    pt.queue = t.plain_queue;
    if (t.bulge_first_tiles) bt.first_tiles = t.bulge_first_tiles;
    bt.round_tiles = t.bulge_round_tiles;
    if (t.verify_blocks_per_cu) {
        if (t.verify_blocks_per_cu > 64) throw cof::Error(COF_E_INVALID, "verify_blocks_per_cu must be <= 64");
        bt.verify_per_cu = t.verify_blocks_per_cu;
    }
    if (t.bulge_queue_cap) {
        if (t.bulge_queue_cap < 64 || t.bulge_queue_cap > (1u << 27))
            throw cof::Error(COF_E_INVALID, "bulge_queue_cap must be in [64, 2^27]");
        bt.queue_cap = t.bulge_queue_cap;
    }
    if (t.verify_refill) {
        if (t.verify_refill > 64) throw cof::Error(COF_E_INVALID, "verify_refill must be in [1, 64]");
        bt.refill = t.verify_refill;
    }
    if (t.bulge_screen > 2) throw cof::Error(COF_E_INVALID, "bulge_screen must be 0, 1 or 2");
    bt.screen = t.bulge_screen == 2 ? 0u : 1u;
    ctx->plain_tune = pt;
    ctx->bulge_tune = bt;
    ctx->cached = false;
}

// Times each candidate shape on a genome prefix and keeps the fastest:
// plain-scan workgroup size x patterns per workgroup, and the persistent
// tile-queue scan (kernel time of a scan of the prefix), then, on the wall time of the bulge path over the prefix's
// batches, traceback workgroups per CU, the pipeline's round size and the
// traceback's lane-refill threshold.
void autotune(cof_ctx* ctx, uint64_t sample) {
    const uint64_t n = std::min<uint64_t>(sample, ctx->g.n_nibbles());
    if (!cof::plain_plan_empty(ctx->plain)) {
        cof::RawView<cof::PlainRaw> raw;
        double best = 1e300;
        cof::PlainTune best_t = ctx->plain_tune;
        for (uint32_t th : {64u, 128u, 256u, 512u}) {
            for (uint32_t pc : {0u, 16u, 64u, 1000u}) {
                // pc 1000: the persistent tile-queue scan (64-thread tiles, whole groups)
                if (pc == 1000 && th != 64) continue;
                cof::PlainTune t;
                t.threads = th;
                t.pat_chunk = pc == 1000 ? 0u : pc;
                t.queue = pc == 1000 ? 1u : 0u;
                double ms = 0, warm = 0;
                uint64_t nl = 0;
                cof::plain_run(ctx->plain, ctx->ws, ctx->d_genome, n, ctx->stream, raw, &warm, &nl, t);
                for (int r = 0; r < 2; r++)
                    cof::plain_run(ctx->plain, ctx->ws, ctx->d_genome, n, ctx->stream, raw, &ms, &nl, t);
                if (ms < best) { best = ms; best_t = t; }
            }
        }
        ctx->plain_tune = best_t;
    }
    if (ctx->bulge && cof::bulge_plan_size(ctx->bulge)) {
        // the batches of the prefix's chunks
        size_t nc = 0;
        while (nc < ctx->g.chunks.size() && ctx->g.chunks[nc].nib_begin() < n) nc++;
        std::vector<cof::Batch> sb = cof::make_batches(std::max<size_t>(nc, 1), ctx->q.batch_nucl);
        // coordinate descent: traceback workgroups per CU, then the pipeline's
        // round size (filter tiles of 4 Ki ends per round; 0 sizes rounds from
        // the first round's candidate density), then the lane-refill threshold
        cof::RawView<cof::BulgeRaw> raw;
        cof::BulgeTune best_t = ctx->bulge_tune;
        best_t.flags = ctx->q.flags;
        auto time_of = [&](const cof::BulgeTune& t) {
            cof_stats st{};
            cof::bulge_run(ctx->bulge, ctx->ws, ctx->g, sb, ctx->d_genome, ctx->stream, raw, &st, t);
            return st.bulge_ms;
        };
        double best = 1e300;
        for (uint32_t v : {4u, 8u, 12u, 16u, 24u}) {
            cof::BulgeTune t = best_t;
            t.verify_per_cu = v;
            const double ms = time_of(t);
            if (ms < best) { best = ms; best_t = t; }
        }
        for (uint32_t rt : {256u, 1024u, 4096u}) {
            cof::BulgeTune t = best_t;
            t.round_tiles = rt;
            const double ms = time_of(t);
            if (ms < best) { best = ms; best_t = t; }
        }
        for (uint32_t rf : {8u, 32u}) {
            cof::BulgeTune t = best_t;
            t.refill = rf;
            const double ms = time_of(t);
            if (ms < best) { best = ms; best_t = t; }
        }
        best_t.flags = 0;
        ctx->bulge_tune = best_t;
    }
    ctx->cached = false;
}

}  // namespace

extern "C" {

cof_status cof_open(int hip_device, const cof_query* q, cof_ctx** out) {
    if (!out) return COF_E_INVALID;
    *out = nullptr;
    auto ctx = std::make_unique<cof_ctx>();
    try {
        ctx->device = hip_device;
        HIPCHECK(hipSetDevice(hip_device));
        HIPCHECK(hipStreamCreate(&ctx->stream));
        cof::plain_static_init(hip_device);
        ctx->static_stage.reset(new cof::PlainRaw[cof::plain_static_max_hits()]);
        ctx->q = cof::prepare_query(q);
        ctx->plain = cof::plain_plan_create(ctx->q);
        ctx->bulge = cof::bulge_plan_create(ctx->q);
    } catch (const cof::Error& e) {
        cof_status s = e.code;
        cof_close(ctx.release());
        return s;
    }
    *out = ctx.release();
    return COF_OK;
}

cof_status cof_set_query(cof_ctx* ctx, const cof_query* q) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx) return COF_E_INVALID;
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        cof::Query nq = cof::prepare_query(q);
        cof::PlainPlan* np = cof::plain_plan_create(nq);
        cof::BulgePlan* nb = nullptr;
        try {
            nb = cof::bulge_plan_create(nq);
        } catch (...) {
            cof::plain_plan_destroy(np);
            throw;
        }
        cof::plain_plan_destroy(ctx->plain);
        cof::bulge_plan_destroy(ctx->bulge);
        cof::feature_plan_destroy(ctx->feat);
        ctx->feat = nullptr;
        ctx->plain = np;
        ctx->bulge = nb;
        bool batch_changed = nq.batch_nucl != ctx->q.batch_nucl;
        ctx->q = std::move(nq);
        if (ctx->have_genome && batch_changed) set_batches(ctx);
        ctx->cached = false;
        ctx->hits.clear();
        ctx->aux.clear();
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_load_genome_file(cof_ctx* ctx, const char* path) { return load_file(ctx, path, 0); }

cof_status cof_load_genome_file_ex(cof_ctx* ctx, const char* path, uint32_t flags) {

    return load_file(ctx, path, flags);
}

namespace {

// A caller-built chunk table and packed genome -> resident genome.
void load_packed(cof_ctx* ctx, const uint8_t* packed, uint64_t n_bytes, const cof_chunk* chunks, uint64_t n_chunks,
                 const char* const* names, uint32_t n_chroms) {
    HIPCHECK(hipSetDevice(ctx->device));
    cof::Genome g;
    for (uint32_t i = 0; i < n_chroms; i++) g.chrom_names.push_back(names[i] ? names[i] : "");
    uint64_t expect = 0;
    for (uint64_t i = 0; i < n_chunks; i++) {
        const cof_chunk& c = chunks[i];
        if (c.chrom >= n_chroms || c.end < c.start || c.end - c.start > cof::kChunkNucl || c.byte_offset != expect)
            throw cof::Error(COF_E_INVALID, "chunk table is not a tight packing of <=64 Knt chunks");
        if (i > 0 && !(chunks[i - 1].end == c.start || c.start == 0))
            throw cof::Error(COF_E_INVALID, "chromosome chunks out of order");
        g.chunks.push_back({c.chrom, c.start, c.end, c.byte_offset});
        expect += (c.end - c.start + 1) / 2;
    }
    if (expect != n_bytes) throw cof::Error(COF_E_INVALID, "packed size does not match the chunk table");
    g.n_bytes = n_bytes;
    g.packed.assign(packed, packed + n_bytes);
    g.packed.resize(n_bytes + 256, 0);
    ctx->g = std::move(g);
    upload_genome(ctx);
}

}  // namespace

cof_status cof_load_genome(cof_ctx* ctx, const uint8_t* packed, uint64_t n_bytes, const cof_chunk* chunks,
                           uint64_t n_chunks, const char* const* names, uint32_t n_chroms) {
    if (!ctx || (!packed && n_bytes) || (!chunks && n_chunks) || (!names && n_chroms)) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    ctx->streaming = false;
    ctx->single_batch = -1;
    try {
        load_packed(ctx, packed, n_bytes, chunks, n_chunks, names, n_chroms);
    } catch (const cof::Error& e) {
        ctx->have_genome = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_stream_begin(cof_ctx* ctx) {
    if (!ctx) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    ctx->stream_g = cof::Genome{};
    ctx->streaming = true;
    ctx->have_genome = false;
    ctx->cached = false;
    return COF_OK;
}

cof_status cof_stream_chunk(cof_ctx* ctx, const char* chr_name, uint64_t chunk_start, uint64_t chunk_end,
                            const uint8_t* data) {
    if (!ctx || !chr_name || (!data && chunk_end > chunk_start)) return COF_E_INVALID;
    if (!ctx->streaming) {
        ctx->err = "cof_stream_chunk outside cof_stream_begin / cof_stream_end";
        return COF_E_INVALID;
    }
    cof::Genome& g = ctx->stream_g;
    if (chunk_end < chunk_start || chunk_end - chunk_start > cof::kChunkNucl) {
        ctx->err = "a chunk holds at most 65536 nucleotides";
        return COF_E_INVALID;
    }
    // chunks arrive in chromosome order (search.rs:1401-1404); a chunk that
    // starts at 0 opens a chromosome
    if (chunk_start == 0) {
        g.chrom_names.push_back(chr_name);
    } else if (g.chunks.empty() || g.chunks.back().end != chunk_start || g.chrom_names.back() != chr_name) {
        ctx->err = "search expects chromosome chunks to arrive in order";
        return COF_E_INVALID;
    }
    const uint64_t nb = (chunk_end - chunk_start + 1) / 2;
    g.chunks.push_back({(uint32_t)(g.chrom_names.size() - 1), chunk_start, chunk_end, g.n_bytes});
    g.packed.insert(g.packed.end(), data, data + nb);
    if ((chunk_end - chunk_start) & 1) g.packed.back() &= 0x0F;   // the padding nibble of an odd chunk is 0
    g.n_bytes += nb;
    return COF_OK;
}

cof_status cof_stream_end(cof_ctx* ctx) {
    if (!ctx) return COF_E_INVALID;
    if (!ctx->streaming) {
        ctx->err = "cof_stream_end without cof_stream_begin";
        return COF_E_INVALID;
    }
    ctx->streaming = false;
    ctx->single_batch = -1;
    try {
        ctx->stream_g.packed.resize(ctx->stream_g.n_bytes + 256, 0);
        ctx->g = std::move(ctx->stream_g);
        ctx->stream_g = cof::Genome{};
        upload_genome(ctx);
    } catch (const cof::Error& e) {
        ctx->have_genome = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

namespace {

// First half of a search: statistics reset and the plain scan queued.
void search_begin(cof_ctx* ctx) {
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->search_t0 = std::chrono::steady_clock::now();
    cof_stats& st = ctx->stats;
    uint64_t genome_nucl = st.genome_nucl;
    st = cof_stats{};
    st.genome_nucl = genome_nucl;
    st.load_ms = ctx->load_ms;
    st.load_device = ctx->load_device ? 1 : 0;
    ctx->hits.clear();
    ctx->aux.clear();
    ctx->cached = false;
    cof::plain_launch(ctx->plain, ctx->ws, ctx->d_genome, ctx->g.n_nibbles(), ctx->stream, &st.n_plain_launches,
                      ctx->plain_tune);
    ctx->pending = true;
}

// Second half: plain hits to the host, the bulge path, conversion to final hits.
void search_end(cof_ctx* ctx) {
    ctx->pending = false;
    HIPCHECK(hipSetDevice(ctx->device));
    cof_stats& st = ctx->stats;
    cof::RawView<cof::PlainRaw> praw;
    cof::plain_finish(ctx->plain, ctx->ws, ctx->d_genome, ctx->stream, praw, &st.kernel_ms_plain,
                      &st.n_plain_launches, ctx->plain_tune);
    cof::RawView<cof::BulgeRaw> braw;
    cof::BulgeTune bt = ctx->bulge_tune;
    bt.flags = ctx->q.flags;
    if (!ctx->model.empty() && ctx->bulge && cof::bulge_plan_size(ctx->bulge)) {
        predict_chunk_cost(ctx);
        bt.chunk_cost = &ctx->chunk_cost;
    }
    cof::bulge_run(ctx->bulge, ctx->ws, ctx->g, ctx->batches, ctx->d_genome, ctx->stream, braw, &st, bt);
    st.n_raw_hits = praw.size() + braw.size();
    auto tc = std::chrono::steady_clock::now();
    cof::convert_plain(ctx->g, ctx->q, ctx->batches, ctx->bidx, praw, ctx->hits, ctx->aux);
    size_t first_bulge = ctx->hits.size();
    cof::convert_bulge(ctx->g, ctx->q, ctx->batches, ctx->bidx, braw, ctx->hits, ctx->aux);
    if (!(ctx->q.flags & COF_FLAG_COUNT_ONLY)) {
        st.bulge_hits = ctx->hits.size() - first_bulge;
        for (size_t h = first_bulge; h < ctx->hits.size(); h++) {
            const cof_hit& x = ctx->hits[h];
            uint64_t so[3];
            cof::ops_reverse(x.ops, x.n_ops, so);
            st.bulge_checksum += cof::bulge_hit_hash(ctx->aux[h].nib, cof::pattern_id_of(ctx->q.pats[x.pattern]),
                                                     x.mismatches, x.dna_bulge, x.rna_bulge, x.n_ops, so[0], so[1]);
        }
    }
    const auto t_end = std::chrono::steady_clock::now();
    st.convert_ms = std::chrono::duration<double, std::milli>(t_end - tc).count();
    st.search_ms = std::chrono::duration<double, std::milli>(t_end - ctx->search_t0).count();
    st.n_hits = ctx->hits.size();
    ctx->cached = true;
}

}  // namespace

cof_status cof_search(cof_ctx* ctx, cof_hit* out, uint64_t cap, uint64_t* n_hits) {
    if (!ctx) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    if (!ctx->have_genome) return COF_E_NOGENOME;
    try {
        // a count query (no buffer) always searches; a fetch reuses the cache
        if (!ctx->cached || !out) {
            search_begin(ctx);
            search_end(ctx);
        }
        if (n_hits) *n_hits = ctx->hits.size();
        if (cap < ctx->hits.size()) return COF_E_OVERFLOW;   // nothing written
        if (out && !ctx->hits.empty()) memcpy(out, ctx->hits.data(), ctx->hits.size() * sizeof(cof_hit));
    } catch (const cof::Error& e) {
        ctx->pending = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_search_batch(cof_ctx* ctx, const uint8_t* packed, uint64_t n_bytes, const cof_chunk* chunks,
                            uint64_t n_chunks, const char* const* names, uint32_t n_chroms, uint32_t has_head,
                            uint64_t* n_hits) {
    if (!ctx || (!packed && n_bytes) || (!chunks && n_chunks) || (!names && n_chroms) || has_head > 1)
        return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    if (has_head && n_chunks < 2) {
        ctx->err = "a batch with a head chunk holds at least two chunks";
        return COF_E_INVALID;
    }
    ctx->streaming = false;
    try {
        ctx->single_batch = (int)has_head;
        load_packed(ctx, packed, n_bytes, chunks, n_chunks, names, n_chroms);
        search_begin(ctx);
        search_end(ctx);
        if (n_hits) *n_hits = ctx->hits.size();
    } catch (const cof::Error& e) {
        ctx->pending = false;
        ctx->have_genome = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_hit_seqs(cof_ctx* ctx, uint64_t first, uint64_t n, char* dna, char* rna, size_t stride,
                        uint32_t* lens) {
    if (!ctx || (n && (!dna || !rna || !lens))) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    if (!ctx->cached || first > ctx->hits.size() || n > ctx->hits.size() - first) {
        ctx->err = "hit range outside the hits of the last search";
        return COF_E_INVALID;
    }
    // every string fits, or nothing is written
    for (uint64_t k = first; k < first + n; k++) {
        const cof_hit& h = ctx->hits[k];
        const cof::Pattern& p = ctx->q.pats[h.pattern];
        if ((p.bulge ? h.n_ops : p.L) > stride) return COF_E_OVERFLOW;
    }
    std::string d, r;
    for (uint64_t k = 0; k < n; k++) {
        cof::hit_seqs(ctx->g, ctx->q, ctx->hits[first + k], ctx->aux[first + k], d, r);
        memcpy(dna + k * stride, d.data(), d.size());
        memcpy(rna + k * stride, r.data(), r.size());
        lens[k] = (uint32_t)d.size();
    }
    return COF_OK;
}

const char* cof_chrom_name(const cof_ctx* ctx, uint32_t chrom) {
    if (!ctx || !ctx->have_genome || chrom >= ctx->g.chrom_names.size()) return nullptr;
    return ctx->g.chrom_names[chrom].c_str();
}

cof_status cof_search_static(cof_ctx* ctx, cof_hit* out, uint64_t cap, uint64_t* n_hits) {
    if (!ctx) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    if (!ctx->have_genome) return COF_E_NOGENOME;
    if (ctx->q.any_bulge) {
        ctx->err = "cof_search_static serves plain queries (every guide with D = R = 0)";
        return COF_E_UNSUPPORTED;
    }
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        cof::RawView<cof::PlainRaw> raw;
        uint64_t launches = 0;
        cof::plain_search_static(ctx->plain, ctx->ws, ctx->d_genome, ctx->g.n_nibbles(), ctx->stream, ctx->plain_tune,
                                 ctx->static_counts, ctx->static_stage.get(), raw, &launches);
        const uint64_t n = cof::convert_plain_into(ctx->g, ctx->q, ctx->batches, ctx->bidx, raw, nullptr);
        if (n_hits) *n_hits = n;
        if (out && cap < n) return COF_E_OVERFLOW;   // nothing written
        if (out) cof::convert_plain_into(ctx->g, ctx->q, ctx->batches, ctx->bidx, raw, out);
        ctx->stats.n_plain_launches = launches;
        ctx->stats.n_raw_hits = raw.size();
        ctx->stats.n_hits = n;
        ctx->stats.search_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_search_async(cof_ctx* ctx) {
    if (!ctx) return COF_E_INVALID;
    if (ctx->pending) return refuse_pending(ctx);
    if (!ctx->have_genome) return COF_E_NOGENOME;
    try {
        search_begin(ctx);
    } catch (const cof::Error& e) {
        ctx->pending = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_search_wait(cof_ctx* ctx, uint64_t* n_hits) {
    if (!ctx) return COF_E_INVALID;
    if (!ctx->pending) {
        ctx->err = "no search in flight";
        return COF_E_INVALID;
    }
    try {
        search_end(ctx);
        if (n_hits) *n_hits = ctx->hits.size();
    } catch (const cof::Error& e) {
        ctx->pending = false;
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_write_tsv_ex(cof_ctx* ctx, const char* path, uint32_t flags) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx || !path) return COF_E_INVALID;
    if (!ctx->cached) return COF_E_INVALID;
    if (flags & ~COF_WRITE_HOST) return COF_E_INVALID;
    auto t0 = std::chrono::steady_clock::now();
    FILE* f = strcmp(path, "-") == 0 ? stdout : fopen(path, "wb");
    if (!f) { ctx->err = std::string("cannot open ") + path; return COF_E_IO; }
    std::vector<char> buf(1 << 22);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    cof_status st = COF_OK;
    try {
        if (flags & COF_WRITE_HOST) {
            cof::write_tsv(ctx->g, ctx->q, ctx->hits, ctx->aux, f);
        } else {
            HIPCHECK(hipSetDevice(ctx->device));
            cof::write_tsv_device(ctx->g, ctx->q, ctx->hits, ctx->aux, ctx->d_genome, f, ctx->stream, ctx->fmt);
        }
    } catch (const cof::Error& e) {
        st = fail(ctx, e);
    }
    if (f != stdout) {
        if (fclose(f) != 0 && st == COF_OK) { ctx->err = std::string("cannot write ") + path; st = COF_E_IO; }
    } else {
        fflush(f);
        setvbuf(f, nullptr, _IOLBF, 0);
    }
    ctx->stats.write_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.write_device = (flags & COF_WRITE_HOST) ? 0 : 1;
    return st;
}

cof_status cof_write_tsv(cof_ctx* ctx, const char* path) { return cof_write_tsv_ex(ctx, path, 0); }

cof_status cof_get_stats(const cof_ctx* ctx, cof_stats* out) {
    if (!ctx || !out) return COF_E_INVALID;
    *out = ctx->stats;
    out->load_ms = ctx->load_ms;
    out->load_device = ctx->load_device ? 1 : 0;
    return COF_OK;
}

cof_status cof_set_tuning(cof_ctx* ctx, const cof_tuning* t) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx || !t) return COF_E_INVALID;
    try {
        set_tuning(ctx, *t);
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_get_tile_stats(const cof_ctx* ctx, cof_tile_stat* out, uint64_t cap, uint64_t* n) {
    if (!ctx || (!out && cap)) return COF_E_INVALID;
    try {
        const cof::Workspace& ws = ctx->ws;
        const uint64_t nt = ws.tile_j0.size();
        if (n) *n = nt;
        const uint64_t k = std::min<uint64_t>(cap, nt);
        if (k) {
            std::vector<uint32_t> cand(k);
            std::vector<unsigned long long> steps(k);
            HIPCHECK(hipSetDevice(ctx->device));
            HIPCHECK(hipMemcpy(cand.data(), ws.tile_cand, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
            HIPCHECK(hipMemcpy(steps.data(), ws.tile_steps, k * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            for (uint64_t t = 0; t < k; t++)
                out[t] = cof_tile_stat{ws.tile_j0[t], ws.tile_n[t], cand[t], steps[t], ws.tile_chunk[t], 0u};
        }
        if (cap < nt && out) return COF_E_OVERFLOW;
    } catch (const cof::Error& e) {
        return COF_E_HIP;
    }
    return COF_OK;
}

cof_status cof_chunk_features(cof_ctx* ctx, float* out, uint64_t cap_rows, uint64_t* n_rows) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx || (!out && cap_rows)) return COF_E_INVALID;
    if (!ctx->have_genome) return COF_E_NOGENOME;
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        const uint64_t nc = ctx->g.chunks.size();
        if (n_rows) *n_rows = nc;
        if (!out) return COF_OK;
        compute_features(ctx);
        const uint64_t k = std::min(cap_rows, nc);
        if (k)
            HIPCHECK(hipMemcpyAsync(out, ctx->d_feat.p, k * cof::kMlpIn * sizeof(float), hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        if (cap_rows < nc) return COF_E_OVERFLOW;
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_set_cost_model(cof_ctx* ctx, const cof_mlp* m) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx) return COF_E_INVALID;
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        if (!m) {
            ctx->model.clear();
        } else {
            ctx->model.resize(cof::kMlpParams);
            m->m->get_params(ctx->model.data());
            HIPCHECK(hipSetDevice(ctx->device));
            ctx->d_model.ensure(cof::kMlpParams * sizeof(float));
            HIPCHECK(hipMemcpy(ctx->d_model.p, ctx->model.data(), cof::kMlpParams * sizeof(float), hipMemcpyHostToDevice));
        }
        ctx->cached = false;
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

cof_status cof_mlp_create(int hip_device, uint64_t seed, cof_mlp** out) {
    if (!out) return COF_E_INVALID;
    *out = nullptr;
    try {
        auto m = std::make_unique<cof_mlp>();
        m->m = std::make_unique<cof::Mlp>(hip_device, seed);
        *out = m.release();
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

void cof_mlp_destroy(cof_mlp* m) { delete m; }

cof_status cof_mlp_get_params(const cof_mlp* m, float* out) {
    if (!m || !out) return COF_E_INVALID;
    try {
        m->m->get_params(out);
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

cof_status cof_mlp_set_params(cof_mlp* m, const float* p) {
    if (!m || !p) return COF_E_INVALID;
    try {
        m->m->set_params(p);
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

cof_status cof_mlp_forward(cof_mlp* m, const float* x, uint64_t n_rows, float* y) {
    if (!m || (n_rows && (!x || !y)) || n_rows > 0xFFFFFFFFull) return COF_E_INVALID;
    try {
        m->m->forward(x, n_rows, y);
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

cof_status cof_mlp_loss_grad(cof_mlp* m, const float* x, const float* t, uint64_t n_rows, float* loss, float* grad) {
    if (!m || !x || !t || !n_rows || n_rows > 0xFFFFFFFFull) return COF_E_INVALID;
    try {
        m->m->loss_grad(x, t, n_rows, loss, grad);
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

cof_status cof_mlp_train(cof_mlp* m, const float* x, const float* t, uint64_t n_rows, uint32_t steps, float lr,
                         uint32_t flags, float* loss) {
    if (!m || !x || !t || !n_rows || n_rows > 0xFFFFFFFFull || !(lr > 0)) return COF_E_INVALID;
    try {
        m->m->train(x, t, n_rows, steps, lr, (flags & COF_MLP_EAGER) != 0, loss);
    } catch (const cof::Error& e) {
        return e.code;
    }
    return COF_OK;
}

cof_status cof_get_tuning(const cof_ctx* ctx, cof_tuning* out) {
    if (!ctx || !out) return COF_E_INVALID;
    *out = get_tuning(ctx);
    return COF_OK;
}

cof_status cof_autotune(cof_ctx* ctx, uint64_t sample_nucl, cof_tuning* chosen) {
    if (ctx && ctx->pending) return refuse_pending(ctx);

    if (!ctx) return COF_E_INVALID;
    if (!ctx->have_genome) return COF_E_NOGENOME;
    try {
        HIPCHECK(hipSetDevice(ctx->device));
        autotune(ctx, sample_nucl ? sample_nucl : (64ull << 20));
        if (chosen) *chosen = get_tuning(ctx);
    } catch (const cof::Error& e) {
        return fail(ctx, e);
    }
    return COF_OK;
}

void cof_close(cof_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->pending) (void)hipStreamSynchronize(ctx->stream);   // a search left in flight
    cof::plain_plan_destroy(ctx->plain);
    cof::bulge_plan_destroy(ctx->bulge);
    cof::feature_plan_destroy(ctx->feat);
    if (ctx->d_genome) (void)hipFree(ctx->d_genome);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

cof_status cof_device_count(int* n) {
    if (!n) return COF_E_INVALID;
    *n = 0;
    if (hipGetDeviceCount(n) != hipSuccess) {
        *n = 0;
        return COF_E_HIP;
    }
    return COF_OK;
}

cof_status cof_device_name(int device, char* buf, size_t cap) {
    if (!buf || cap == 0) return COF_E_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return COF_E_INVALID;
    snprintf(buf, cap, "%s (%s, %d CUs, %.0f GiB)", prop.name, prop.gcnArchName, prop.multiProcessorCount,
             prop.totalGlobalMem / 1073741824.0);
    return COF_OK;
}

const char* cof_strerror(cof_status s) {
    switch (s) {
        case COF_OK: return "ok";
        case COF_E_INVALID: return "invalid argument or query";
        case COF_E_HIP: return "HIP runtime error";
        case COF_E_OVERFLOW: return "output buffer too small";
        case COF_E_IO: return "genome file unreadable or malformed";
        case COF_E_NOGENOME: return "no genome loaded";
        case COF_E_NOMEM: return "out of memory";
        case COF_E_UNSUPPORTED: return "query not supported by the kernels";
        default: return "unknown status";
    }
}

const char* cof_last_error(const cof_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

}  // extern "C"
