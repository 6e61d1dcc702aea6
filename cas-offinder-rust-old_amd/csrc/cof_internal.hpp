// Internal declarations shared by the host side of libcof (C++17) and the HIP
// translation units.  Nothing here is part of the C ABI (include/cof.h).
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <memory>
#include <utility>
#include <vector>

#include "cof.h"

namespace cof {

// ---------------------------------------------------------------------------
// bit4 nucleotide code (same table values as cas-offinder-lib/src/bit4ops.rs
// :1-121, because the packed genome layout is part of the boundary):
// T=1 C=2 A=4 G=8, IUPAC = OR, N = 0xF, lowercase == uppercase, other = 0.
// ---------------------------------------------------------------------------
struct Codec {
    uint8_t enc_mixed[256];   // any IUPAC letter
    uint8_t enc_plain[256];   // A/C/G/T only (cmp_chars' DNA side)
    char dec[16];             // bit4 -> uppercase letter, 0 -> NUL
    Codec();
};
extern const Codec kCodec;

inline uint8_t nib(const uint8_t* packed, uint64_t i) {
    uint8_t b = packed[i >> 1];
    return (i & 1) ? (uint8_t)(b >> 4) : (uint8_t)(b & 0xF);
}
inline uint8_t comp4(uint8_t v) { return (uint8_t)(((v << 2) | (v >> 2)) & 0xF); }
char comp_char(char c);                                   // keeps case, maps IUPAC
std::string revcomp(const std::string& s);
inline bool is_n_char(char c) { return c == 'N' || c == 'n'; }
// cmp_chars(dna, rna) of bit4ops.rs:233-237 on characters
inline bool cmp_chars(uint8_t dna, uint8_t rna) { return (kCodec.enc_plain[dna] & kCodec.enc_mixed[rna]) != 0; }
// one-bit (A/C/G/T) nibble test
inline bool single_base(uint8_t v) { return v != 0 && (v & (v - 1)) == 0; }

struct Error : std::runtime_error {
    cof_status code;
    Error(cof_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------------------
// Resident genome: tight-packed 64 Knt chunks (search.rs:1376-1427 layout).
// ---------------------------------------------------------------------------
static const uint64_t kChunkNucl = 1u << 16;

struct Chunk {
    uint32_t chrom;
    uint64_t start, end;      // within chromosome
    uint64_t byte_off;        // in packed buffer
    uint64_t size() const { return end - start; }
    uint64_t nib_begin() const { return byte_off * 2; }
};

struct Genome {
    std::vector<std::string> chrom_names;
    std::vector<Chunk> chunks;
    std::vector<uint8_t> packed;   // host copy, + zero padding
    uint64_t n_bytes = 0;          // valid bytes
    uint64_t n_nibbles() const { return n_bytes * 2; }
    // lut[k] = chunk holding byte k << 15 (chunks hold at most 32 KB, so a
    // bucket has few chunk starts); built by index() once the table is final
    std::vector<uint32_t> lut;
    void index();
    // chunk containing global nibble i (the last chunk starting at or before it)
    size_t chunk_of(uint64_t nibble) const;
};

void read_genome_file(const std::string& path, Genome& g);   // FASTA / 2bit / folder

// Virtual batches of the reference batcher (search.rs:1525-1572): batch b
// holds chunks [first, last), the first being a head carried over from the
// previous batch when has_head.
struct Batch {
    size_t first, last;
    bool has_head;
};
std::vector<Batch> make_batches(size_t n_chunks, uint64_t batch_nucl);

// ---------------------------------------------------------------------------
// Query preparation.
// ---------------------------------------------------------------------------
struct Pattern {
    std::string ascii;          // as given (forward) or reverse complement
    std::string eff_filter;     // filter (forward) or revcomp(filter) (search.rs:1213-1225)
    std::vector<uint8_t> b4;    // one nibble per position
    uint32_t L, M, D, R;
    uint32_t guide;             // forward guide index
    bool forward;
    uint32_t pam_len_out;       // trailing Ns of the first guide of its run (main.rs:80-91)
    bool bulge;                 // D>0 || R>0
};

struct Query {
    std::vector<Pattern> pats;  // n fwd then n rev
    uint32_t n_guides = 0;
    uint64_t batch_nucl = 1ull << 30;
    uint32_t flags = 0;
    bool any_plain = false, any_bulge = false;
};
Query prepare_query(const cof_query* q);

// Raw device hit of the plain scan: global start nibble, pattern and the
// writer's recounted mismatches on the resident genome.
struct PlainRaw {
    uint64_t pos;
    uint32_t pattern;
    uint32_t flags;   // recounted mismatches
};

// Raw device hit of the bulge path: alignment start nibble, pattern, counts, ops.
struct BulgeRaw {
    uint64_t start;
    uint32_t pattern;
    uint8_t mm, db, rb, n_ops;
    uint64_t ops[3];          // DFS stack order: bits 2k of op k counted from the alignment's end
};

// Stack order <-> alignment order of n 2-bit ops (an involution).
inline void ops_reverse(const uint64_t* in, uint32_t n, uint64_t* out) {
    uint64_t o[3] = {0, 0, 0};
    for (uint32_t k = 0; k < n; k++) {
        uint64_t op = (in[k / 32] >> (2 * (k % 32))) & 3u;
        uint32_t r = n - 1 - k;
        o[r / 32] |= op << (2 * (r % 32));
    }
    out[0] = o[0]; out[1] = o[1]; out[2] = o[2];
}

#if defined(__HIPCC__)
#define COF_HD __host__ __device__
#else
#define COF_HD
#endif

// Order-independent checksum term of one bulge hit (summed over hits).
COF_HD inline uint64_t cof_mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}
// pattern_id: a hash of the pattern's sequence and strand (pattern_id_of), so
// the checksum does not depend on how a query numbers its patterns
// Hash of one bulge hit; ops in stack order (at most 64 ops: two words).
COF_HD inline uint64_t bulge_hit_hash(uint64_t start, uint64_t pattern_id, uint32_t mm, uint32_t db, uint32_t rb,
                                      uint32_t n_ops, uint64_t ops0, uint64_t ops1) {
    const uint64_t meta = ((uint64_t)mm << 24) | ((uint64_t)db << 16) | ((uint64_t)rb << 8) | n_ops;
    return cof_mix64((start * 0x9E3779B97F4A7C15ull) ^ pattern_id ^ (meta << 40) ^ (ops0 * 0xD6E8FEB86659FD93ull) ^
                     (ops1 * 0xA0761D6478BD642Full));
}

inline uint64_t pattern_id_of(const Pattern& p) {
    uint64_t h = 0xcbf29ce484222325ull;                 // FNV-1a over the bit4 codes
    for (uint8_t v : p.b4) { h ^= v; h *= 0x100000001b3ull; }
    return cof_mix64(h ^ (p.forward ? 0x5bd1e995ull : 0x27d4eb2full) ^ ((uint64_t)p.M << 48) ^
                     ((uint64_t)p.D << 52) ^ ((uint64_t)p.R << 56));
}

// Host-side conversion (chunk lookup, batch-seam and chromosome-end drop
// rules, writer predicate) of raw device hits into final hits.
// Per-hit data the row writer needs beyond cof_hit: the global start nibble
// and the end of the batch view the hit was reported from.
struct HitAux {
    uint64_t nib;
    uint64_t view_end;
};
// Vector that leaves new elements uninitialised (hit buffers are filled by
// the converters; zero-filling tens of MB per search would only cost time).
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <typename U>
    void construct(U* p) { ::new ((void*)p) U; }
    template <typename U, typename... A>
    void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};
using HitVec = std::vector<cof_hit, NoInitAlloc<cof_hit>>;
using AuxVec = std::vector<HitAux, NoInitAlloc<HitAux>>;

// Raw hits in (pinned) host memory, as copied from the device.
template <typename T>
struct RawView {
    const T* p = nullptr;
    size_t n = 0;
    size_t size() const { return n; }
    const T& operator[](size_t i) const { return p[i]; }
};

// Per chunk and per batch facts of a batch layout the converters look up.
struct BatchIndex {
    std::vector<int64_t> nonhead_batch;    // per chunk: the batch holding it as a member (-1: none)
    std::vector<int64_t> head_batch;       // per chunk: the batch holding it as its head chunk (-1: none)
    std::vector<uint64_t> end_nib;         // per batch: end of its last chunk's bytes (nibbles)
    std::vector<uint64_t> active_begin;    // per batch: first end position scanned (bulge path)
    BatchIndex() = default;
    BatchIndex(const Genome& g, const std::vector<Batch>& bs);
};
// One raw plain hit -> final hit: the batch that reports it, the seam and
// drop rules, the writer predicate on that batch's view (false: not reported).
bool place_plain_hit(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                     const PlainRaw& r, cof_hit& h, HitAux& a);
// The same over raw hits in order, one thread, into out (null: count only);
// returns the number of final hits.  Allocates nothing.
uint64_t convert_plain_into(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                            RawView<PlainRaw> raw, cof_hit* out);
void convert_plain(const Genome& g, const Query& q, const std::vector<Batch>& batches, const BatchIndex& bi,
                   RawView<PlainRaw> raw, HitVec& out, AuxVec& aux);
void convert_bulge(const Genome& g, const Query& q, const std::vector<Batch>& batches, const BatchIndex& bi,
                   RawView<BulgeRaw> raw, HitVec& out, AuxVec& aux);
void write_tsv(const Genome& g, const Query& q, const HitVec& hits, const AuxVec& aux, FILE* f);
// Match.dna_seq / Match.rna_seq of one final hit, before the writer's case
// marking (plain: N-masked window, NUL for genome N; bulge: aligned strings).
void hit_seqs(const Genome& g, const Query& q, const cof_hit& h, const HitAux& aux, std::string& dna,
              std::string& rna);

}  // namespace cof
