/*
 * cof.h — C ABI of the MI355X (gfx950) off-target search library (libcof.so).
 *
 * The library replaces the GPU backend of cas-offinder-lib: the batched
 * search behind the public `search()` entry point
 * (cas-offinder-lib/src/search.rs:1490-1663), i.e. the backend seam
 * `search_chunk_cuda` / `search_chunk_cuda_myers` (search.rs:322-366,
 * 1193-1314) together with the host code whose output defines the hit set:
 * `convert_matches` / `convert_matches_myers` (search.rs:1316-1374,
 * 1428-1488) and the CLI writer (cas-offinder-cli/src/main.rs:130-247).
 *
 * Conventions: status codes, never aborts; the caller owns host buffers, the
 * library owns device memory; one context per device per host thread (not
 * thread-safe), as the reference runs one host thread per device
 * (search.rs:344-361).  Output order is unspecified; compare sorted rows.
 */
#ifndef COF_H
#define COF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int cof_status;
#define COF_OK 0
#define COF_E_INVALID 1     /* bad query / argument (reference: CliError::ArgumentError, assert!) */
#define COF_E_HIP 2         /* HIP runtime error (reference: DriverError -> panic!) */
#define COF_E_OVERFLOW 3    /* caller buffer too small; *n_hits holds the count needed */
#define COF_E_IO 4          /* genome file unreadable / malformed (reference: CliError::BadFileFormat) */
#define COF_E_NOGENOME 5    /* cof_search before a genome was loaded */
#define COF_E_NOMEM 6       /* device or host allocation failed */
#define COF_E_UNSUPPORTED 7 /* query outside what the kernels implement (e.g. Myers pattern > 64 nt) */

typedef struct cof_ctx cof_ctx;

/*
 * A search query.  One entry per forward guide; the library appends the
 * reverse complement of every guide exactly as main.rs:253-259 does, so
 * pattern indices run 0..n_guides-1 (forward) then n_guides..2*n_guides-1.
 *
 * Reference mode (what one reference run can express, cli_utils.rs:37-125):
 * every guide has the same filter, length and limits.  The library also
 * accepts per-guide filters, lengths and limits; results then equal the
 * union of one reference run per (filter, length, M, D, R) group.
 *
 * Guides whose D and R are both 0 use the plain mismatch semantics
 * (search.rs:1513, the "legacy popcount" path plus main.rs:158-222);
 * guides with D>0 or R>0 use the bulge semantics (PAM quick check, Myers
 * filter, all-alignment DP traceback).
 */
typedef struct {
    uint32_t n_guides;
    const char* const* guides;       /* ASCII IUPAC, guide i has length len(filters[i]) */
    const char* const* filters;      /* ASCII IUPAC filter (PAM mask) per guide */
    const uint32_t* max_mismatches;  /* per guide */
    const uint32_t* max_dna_bulges;  /* per guide, NULL = all 0 */
    const uint32_t* max_rna_bulges;  /* per guide, NULL = all 0 */
    uint64_t batch_nucl;             /* virtual batch size in nt, multiple of 65536;
                                        0 = 1<<30 (reference GPU, search.rs:22,44-50);
                                        1<<22 reproduces the reference CPU path */
    uint32_t flags;                  /* COF_FLAG_* */
} cof_query;

/* Bulge-path hits are counted and folded into an order-independent checksum
 * on the device instead of being returned (for outputs too large to hold:
 * a 3 Gbp bulge search can produce ~1e10 alignments).  cof_search then
 * reports plain hits only in `out`; cof_stats carries the bulge count and
 * checksum, which equal those of a record-mode search. */
#define COF_FLAG_COUNT_ONLY 1u
/* Bulge path schedule (results are identical in every mode):
 * default — after a first synchronous round that measures candidate density,
 *   tile rounds run as a two-stream filter/traceback pipeline with
 *   double-buffered candidate queues, captured into one hipGraph;
 * COF_FLAG_NO_GRAPH — the same pipeline launched eagerly;
 * COF_FLAG_SYNC_ROUNDS — one stream, filter then traceback per round with a
 *   host synchronisation in between (the reference's batch loop shape). */
#define COF_FLAG_NO_GRAPH 2u
#define COF_FLAG_SYNC_ROUNDS 4u
/* Also account traceback DFS transitions per filter tile (cof_tile_stat.dfs_steps;
 * one atomic per candidate, so it costs time). */
#define COF_FLAG_TILE_STATS 8u

/* One chunk of the tight-packed genome (chrom_chunk.rs:1-12, search.rs:57-79). */
typedef struct {
    uint32_t chrom;        /* index into chrom_names */
    uint32_t _pad;
    uint64_t start, end;   /* nucleotide range within the chromosome */
    uint64_t byte_offset;  /* first byte of the chunk in the packed buffer */
} cof_chunk;

/* One hit after chunk->chromosome conversion and the writer's filters
 * (Match, chrom_chunk.rs:15-25, with the writer's recomputed mismatches). */
typedef struct {
    uint64_t chrom_pos;    /* Match.chrom_idx: leftmost + strand coordinate (before the
                              reverse-strand shift of main.rs:229-236) */
    uint32_t chrom;        /* chromosome index */
    uint32_t pattern;      /* pattern index (see cof_query) */
    uint16_t mismatches;
    uint16_t dna_bulge;
    uint16_t rna_bulge;
    uint16_t n_ops;        /* bulge path: alignment length; plain path: 0 */
    uint64_t ops[3];       /* bulge path: 2 bits per op in alignment order, op k at bits
                              2*(k%32) of ops[k/32]; 0 match, 1 mismatch, 2 DNA bulge, 3 RNA bulge */
} cof_hit;

typedef struct {
    double search_ms;          /* wall time of the last search: device work, D2H of hits, conversion */
    double kernel_ms_plain;    /* summed HIP-event time of plain-scan launches */
    double kernel_ms_filter;   /* summed HIP-event time of PAM+Myers filter launches of the
                                  bulge path's synchronous rounds (all rounds with COF_FLAG_SYNC_ROUNDS) */
    double kernel_ms_verify;   /* the same for DP traceback launches */
    uint64_t n_plain_launches;
    uint64_t n_candidates;     /* Myers survivors of the bulge filter */
    uint64_t n_raw_hits;       /* hits before the host-side seam/drop rules */
    uint64_t n_hits;           /* final hits */
    uint64_t genome_nucl;      /* nucleotides in the resident genome (incl. N) */
    uint64_t bulge_hits;       /* final bulge-path hits (both modes) */
    uint64_t bulge_checksum;   /* sum over final bulge hits of a 64-bit hash of
                                  (start nibble, pattern, mm, db, rb, ops) */
    uint64_t verify_steps;     /* diagnostics: DFS transitions of the traceback kernel */
    uint64_t verify_slots;     /* diagnostics: lane slots of its DFS loop (64 per wave iteration) */
    uint64_t n_pipeline_rounds;/* bulge rounds run by the pipeline */
    double bulge_ms;           /* wall time of the bulge path (filter + traceback, all rounds) */
    uint64_t n_redo_rounds;    /* bulge rounds whose queue overflowed and were redone in halves */
    uint64_t verify_busy;      /* diagnostics: slots whose lane held a candidate (busy / slots = lane use) */
    double convert_ms;         /* host conversion of raw device hits into final hits */
    double load_ms;            /* wall time of the last genome file load */
    uint64_t load_device;      /* 1 if that load encoded the FASTA on the device */
    double write_ms;           /* wall time of the last cof_write_tsv{,_ex} */
    uint64_t write_device;     /* 1 if those rows were formatted on the device */
    uint64_t graph_instantiations; /* bulge pipeline graphs instantiated by the last search */
    uint64_t graph_updates;    /* bulge pipeline graphs replayed after an in-place update (same
                                  pipeline shape as the previous search of this context) */
    uint64_t n_traced;         /* Myers survivors that entered the DP traceback (after the
                                  budget screen; n_candidates without it) */
} cof_stats;

/* Launch shapes and schedule.  No reference counterpart: the reference fixes
 * them at kernel-compile time (`-D` constants, search.rs:322-343,1193-1281).
 * Results never depend on them; 0 selects the default of each field. */
typedef struct {
    uint32_t plain_threads;        /* plain-scan workgroup: 64 (default), 128, 256 or 512 threads,
                                      64 start positions per thread */
    uint32_t plain_pat_chunk;      /* patterns per plain-scan workgroup; a larger group is split
                                      over grid rows (0: whole group).  The matrix-core compare
                                      serves 0 and multiples of 16; other values use the plane
                                      compare (same results) */
    uint32_t bulge_first_tiles;    /* filter tiles (4096 end positions each) in the synchronous
                                      first round of the bulge path (default 1024) */
    uint32_t bulge_round_tiles;    /* tiles per later round (0: sized from the first round's
                                      candidate density: about 4 rounds, each within half a queue) */
    uint32_t verify_blocks_per_cu; /* persistent traceback workgroups per CU (default 24) */
    uint32_t bulge_queue_cap;      /* candidate queue capacity per round (default 2^27 = 4 GiB of
                                      32-byte records, min 64); a round that overflows it is redone
                                      in halves */
    uint32_t verify_refill;        /* a traceback wave takes new candidates once this many of its
                                      64 lanes are idle (1-64, default 16) */
    uint32_t plain_hit_cap;        /* plain-scan hit records the device buffer starts with (a
                                      multiple of 64; 0: 2^20).  A scan whose hits do not fit grows
                                      the buffer and runs again */
    uint32_t plain_queue;          /* 1: the plain scan runs as persistent two-wave workgroups that
                                      claim 4,096-start tiles from per-XCD device queues (64-thread
                                      shape, whole groups only); 0: one workgroup per tile */
    uint32_t bulge_screen;         /* 1 (default, also 0): Myers survivors pass an exact budget
                                      screen (an alignment within the M / D / R caps and filter
                                      cells exists) before the DP traceback; 2: every survivor is
                                      traced.  Results are the same either way */
} cof_tuning;

/* Per filter tile of the last bulge search (tiles in genome order). */
typedef struct {
    uint64_t first_end;    /* first alignment end (nibble index in the resident genome) */
    uint32_t n_ends;       /* end positions in the tile (<= 4096) */
    uint32_t candidates;   /* Myers survivors sent to the traceback */
    uint64_t dfs_steps;    /* traceback DFS transitions spent on them (COF_FLAG_TILE_STATS, else 0) */
    uint32_t chunk;        /* genome chunk holding first_end (row of cof_chunk_features) */
    uint32_t _pad;
} cof_tile_stat;

/* Replaces search_chunk_cuda{,_myers}'s set-up (search.rs:322-343,1193-1281):
 * validates the query, builds per-pattern tables and uploads them. */
cof_status cof_open(int hip_device, const cof_query* q, cof_ctx** out);

/* Replaces the query of an open context; the resident genome stays loaded, so
 * many queries can run against one upload (each reference run re-reads the
 * genome, main.rs:106-121). */
cof_status cof_set_query(cof_ctx* ctx, const cof_query* q);

/* Makes a genome resident in HBM from a FASTA file, a 2bit file or a folder of
 * *.fa files, with the reference readers' chunking (read_fasta.rs:9-71,
 * read_2bit.rs:33-118, read_fasta_folder.rs:8-23, main.rs:106-121). */
cof_status cof_load_genome_file(cof_ctx* ctx, const char* path);

/* cof_load_genome_file with a choice of reader: COF_LOAD_HOST forces the
 * host reader; COF_LOAD_DEVICE requires device ingest (the FASTA encoder or
 * the 2bit decoder; COF_E_UNSUPPORTED for folders or FASTA files whose lines
 * in a chunk do not share one width); 0 = device for 2bit and regular FASTA,
 * else host.  The result is the same genome either way.  No reference
 * counterpart (the reference reads on one host thread, read_fasta.rs:9-71,
 * read_2bit.rs:33-118). */
#define COF_LOAD_HOST 1u
#define COF_LOAD_DEVICE 2u
cof_status cof_load_genome_file_ex(cof_ctx* ctx, const char* path, uint32_t flags);

/* Makes a caller-built genome resident: `packed` is the concatenation of the
 * chunks in bit4 (search.rs:1376-1427 layout: ceil(size/2) bytes per chunk,
 * even nucleotide in the low nibble). */
cof_status cof_load_genome(cof_ctx* ctx, const uint8_t* packed, uint64_t n_bytes,
                           const cof_chunk* chunks, uint64_t n_chunks,
                           const char* const* chrom_names, uint32_t n_chroms);

/* Runs the search over the resident genome and writes the hit count to
 * *n_hits.  With out == NULL the search always runs (a count query); with a
 * buffer, the hits of the last search are returned when they are cached,
 * otherwise the search runs first.  Returns COF_E_OVERFLOW when cap < total
 * and then writes nothing to `out` (the cache is kept, so a retry with a
 * larger buffer does not search again).  Replaces the per-batch loop of
 * search_device_cuda{,_myers} (search.rs:249-320,1054-1191) and its overflow
 * bisection (search.rs:1140-1152). */
cof_status cof_search(cof_ctx* ctx, cof_hit* out, uint64_t cap, uint64_t* n_hits);

/* The rest of the reference's Match record (cas-offinder-lib/src/chrom_chunk.rs:19-29)
 * for hits [first, first + n) of the last search: Match.dna_seq and
 * Match.rna_seq byte for byte as convert_matches (search.rs:1456-1475: the
 * window of the N-masked batch, genome N and bytes past the batch as NUL,
 * and the pattern, both decoded) and convert_matches_myers (search.rs:1352-1372:
 * the aligned strings with '-' at bulges) build them, reverse-complemented
 * for reverse-strand hits.  Hit k's strings go to dna + k * stride and
 * rna + k * stride (not NUL-terminated: a plain DNA string may hold NUL),
 * their common length to lens[k].  COF_E_OVERFLOW, with nothing written,
 * when a string is longer than stride (plain: pattern length; bulge: n_ops). */
cof_status cof_hit_seqs(cof_ctx* ctx, uint64_t first, uint64_t n, char* dna, char* rna, size_t stride,
                        uint32_t* lens);
/* Match.chr_name: name of chromosome `chrom` (cof_hit.chrom) of the resident
 * genome, valid until the genome is replaced; NULL when out of range. */
const char* cof_chrom_name(const cof_ctx* ctx, uint32_t chrom);

/* Genome by chunks, for a caller that receives them one by one as
 * search(recv: Receiver<ChromChunkInfo>, ...) does (search.rs:1490-1500,
 * 1525-1572): cof_stream_begin, then one cof_stream_chunk per
 * ChromChunkInfo (chrom_chunk.rs:4-11: chr_name, chunk_start, chunk_end,
 * and data = ceil((end - start) / 2) bytes of bit4, even nucleotide in the
 * low nibble), in chromosome order (a chunk starting at 0 opens a
 * chromosome), then cof_stream_end makes the genome resident.  The result
 * equals cof_load_genome of the same chunks; searches run after the end. */
cof_status cof_stream_begin(cof_ctx* ctx);
cof_status cof_stream_chunk(cof_ctx* ctx, const char* chr_name, uint64_t chunk_start, uint64_t chunk_end,
                            const uint8_t* data);
cof_status cof_stream_end(cof_ctx* ctx);

/* One batch of the backend seam search_chunk_cuda{,_myers} (search.rs:322-329,
 * 1193-1203): the batch is one SearchChunkInfo (chunks_to_searchchunk,
 * search.rs:1376-1427) given as a chunk table over its tight-packed bytes,
 * has_head = SearchChunkMeta.has_overlap_head (the first chunk only gives
 * backward context: plain starts before chunk 1 - (L - 1) and bulge ends
 * before chunk 1 are not reported, search.rs:540-546, 921-935).  The batch
 * replaces the resident genome and is searched as the reference searches
 * one batch, converted as convert_matches{,_myers} convert its result; the
 * hits are then read with cof_search (cached), cof_hit_seqs and
 * cof_chrom_name.  *n_hits = number of hits. */
cof_status cof_search_batch(cof_ctx* ctx, const uint8_t* packed, uint64_t n_bytes, const cof_chunk* chunks,
                            uint64_t n_chunks, const char* const* chrom_names, uint32_t n_chroms, uint32_t has_head,
                            uint64_t* n_hits);

/* cof_search for plain queries (every guide with D = R = 0) without any
 * allocation in the call: no hipMalloc, hipFree, pinned or pageable host
 * allocation.  No reference counterpart; it is for callers that cannot
 * allocate (serving loops, harnesses that forbid allocation while they run).
 * The scan's hit segments, their counters and the packed hits live in one
 * static device array per GPU (16 MiB + 8 KiB, held until the process
 * ends): a search whose hits put more than 8,192 records into one of its 64
 * segments (at most 524,288 raw hits in all) fails with COF_E_NOMEM and a
 * message in cof_last_error, where cof_search has no limit.  Every call
 * searches (nothing is cached): the hit count goes to *n_hits; with out !=
 * NULL and cap >= that count the hits (the records cof_search returns, in
 * unspecified order) are written to out, else COF_E_OVERFLOW and nothing is
 * written.  Calls on one GPU share its array: calls that overlap in time on
 * one GPU (from several host threads) race.  COF_E_UNSUPPORTED for a query
 * with bulge guides. */
cof_status cof_search_static(cof_ctx* ctx, cof_hit* out, uint64_t cap, uint64_t* n_hits);

/* cof_search (count query) in two calls: cof_search_async queues the device
 * work of a search and returns; cof_search_wait completes it (hits copied
 * back and converted, any bulge pattern searched) and writes the count.  In
 * between the host is free, so a caller with several contexts overlaps one
 * context's host work with the next one's kernels, as the reference overlaps
 * its reader, search and writer threads (main.rs:61-274, search.rs:1490-1663).
 * Until cof_search_wait returns, other calls on the context fail with
 * COF_E_INVALID (cof_get_stats and cof_close excepted). */
cof_status cof_search_async(cof_ctx* ctx);
cof_status cof_search_wait(cof_ctx* ctx, uint64_t* n_hits);

/* Writes the cached hits of the last search as reference TSV rows, header
 * included (main.rs:130-247).  path "-" = stdout.  The rows are formatted on
 * the device from the hits and the resident genome and streamed to the file. */
cof_status cof_write_tsv(cof_ctx* ctx, const char* path);
/* The same with flags: COF_WRITE_HOST formats the rows on the host (one
 * thread, the reference writer's shape).  Both produce identical bytes. */
#define COF_WRITE_HOST 1u
cof_status cof_write_tsv_ex(cof_ctx* ctx, const char* path, uint32_t flags);

cof_status cof_get_stats(const cof_ctx* ctx, cof_stats* out);

/* Sets the launch shapes (validated; COF_E_INVALID names the bad field).
 * Build-only, no reference counterpart. */
cof_status cof_set_tuning(cof_ctx* ctx, const cof_tuning* t);
/* Per-tile work of the last bulge search: writes min(cap, tiles) records,
 * *n = number of tiles.  Build-only, no reference counterpart. */
cof_status cof_get_tile_stats(const cof_ctx* ctx, cof_tile_stat* out, uint64_t cap, uint64_t* n);
/* ---- Learned cost model of the bulge scheduler (no reference counterpart) ----
 * An FP32 MLP 12 -> 64 -> 64 -> 1 (ReLU) maps 12 features of a genome chunk
 * to a predicted log(1 + traceback work).  With a model set, the bulge path
 * orders its filter tiles heavy-first and cuts rounds of equal predicted work;
 * hits never depend on it.  Parameters (COF_MLP_PARAMS floats): W1[64][12],
 * b1[64], W2[64][64], b2[64], w3[64], b3. */
#define COF_MLP_IN 12
#define COF_MLP_PARAMS 5057
typedef struct cof_mlp cof_mlp;
/* New model on a HIP device, He-uniform weights from `seed`, zero biases. */
cof_status cof_mlp_create(int hip_device, uint64_t seed, cof_mlp** out);
void cof_mlp_destroy(cof_mlp* m);
cof_status cof_mlp_get_params(const cof_mlp* m, float* out);
cof_status cof_mlp_set_params(cof_mlp* m, const float* params);   /* also resets Adam */
/* y[r] = model(x[r * 12 .. r * 12 + 11]) */
cof_status cof_mlp_forward(cof_mlp* m, const float* x, uint64_t n_rows, float* y);
/* loss = sum_r (y[r] - t[r])^2 / n_rows and its gradient (COF_MLP_PARAMS floats) */
cof_status cof_mlp_loss_grad(cof_mlp* m, const float* x, const float* t, uint64_t n_rows, float* loss, float* grad);
/* `steps` full-batch Adam steps at learning rate `lr`; *loss = loss afterwards.
 * flags: COF_MLP_EAGER launches each step's kernels directly instead of
 * replaying one captured hipGraph (same kernels, same results). */
#define COF_MLP_EAGER 1u
cof_status cof_mlp_train(cof_mlp* m, const float* x, const float* t, uint64_t n_rows, uint32_t steps, float lr,
                         uint32_t flags, float* loss);
/* Features of every genome chunk for the context's query (row c = chunk c,
 * COF_MLP_IN floats per row), computed on the device. */
cof_status cof_chunk_features(cof_ctx* ctx, float* out, uint64_t cap_rows, uint64_t* n_rows);
/* Uses a copy of the model's current parameters for later searches (NULL:
 * static genome order). */
cof_status cof_set_cost_model(cof_ctx* ctx, const cof_mlp* m);

/* Effective launch shapes, defaults filled in. */
cof_status cof_get_tuning(const cof_ctx* ctx, cof_tuning* out);
/* Times candidate shapes on the first `sample_nucl` nucleotides of the
 * resident genome (0: 64 Mnt) with the current query, keeps the fastest and
 * returns it in `chosen` (may be NULL).  Needs a genome.  Build-only, no
 * reference counterpart (the reference compiles one shape per run). */
cof_status cof_autotune(cof_ctx* ctx, uint64_t sample_nucl, cof_tuning* chosen);
void cof_close(cof_ctx* ctx);
/* HIP devices visible to the library and their marketing names; replaces the
 * device discovery and get_device_strs of OclRunConfig (run_config.rs:30-74),
 * which the CLI's usage text lists.  Neither call needs a context. */
cof_status cof_device_count(int* n);
cof_status cof_device_name(int device, char* buf, size_t cap);
const char* cof_strerror(cof_status s);
const char* cof_last_error(const cof_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* COF_H */
