// cof-ref: CPU restatement of the reference's off-target search path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle (and the timed CPU
// baseline) for the MI355X library.  It is never linked into, loaded by or
// called from the product; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker.
//
// It restates, function by function, the CPU path of the Rust reference
// (cas-offinder-lib + cas-offinder-cli).  Each function cites the reference
// file:line it follows.  GPU-only quirks of the reference (the CUDA plain
// kernel skipping the last <=30 starts of a batch, search.rs:280-284, and the
// 32-op record limit, kernel_myers.cu:310-312) are NOT reproduced: the CPU
// functions are normative.
//
// Build: see oracle/Makefile (g++ -O3, std::thread, no other dependencies).
// Parity status of each part is documented at the head of each section.

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <memory>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>
#include <chrono>

namespace ref {

// ---------------------------------------------------------------------------
// bit4 codec — cas-offinder-lib/src/bit4ops.rs:1-121
// T=1 C=2 A=4 G=8, IUPAC = OR of members, N = 0xF.  Lowercase maps like
// uppercase (apply_lower, bit4ops.rs:38-47).  Pinned by the reader
// known-answer tests (file_test.rs:34-73) and closed-form round trips.
// ---------------------------------------------------------------------------
static uint8_t STR2BIT4[2][256];   // [mixed_base][ascii]
static uint8_t BIT4_TO_STR[256];   // invert_chrmap(STR2BIT4[1]), bit4ops.rs:48-59,118

static void init_tables() {
    static bool done = false;
    if (done) return;
    done = true;
    memset(STR2BIT4, 0, sizeof STR2BIT4);
    for (int mixed = 0; mixed < 2; mixed++) {
        uint8_t up[256] = {0};
        up['G'] = 8; up['C'] = 2; up['A'] = 4; up['T'] = 1;       // bit4ops.rs:19-22
        if (mixed) {                                               // bit4ops.rs:23-35
            up['R'] = 4 | 8; up['Y'] = 2 | 1; up['S'] = 8 | 2; up['W'] = 4 | 1;
            up['K'] = 8 | 1; up['M'] = 4 | 2; up['B'] = 2 | 8 | 1; up['D'] = 4 | 8 | 1;
            up['H'] = 4 | 2 | 1; up['V'] = 4 | 2 | 8; up['N'] = 15;
        }
        for (int i = 1; i <= 26; i++) {                            // apply_lower
            STR2BIT4[mixed][i + 96] = up[i + 64];
            STR2BIT4[mixed][i + 64] = up[i + 64];
        }
    }
    memset(BIT4_TO_STR, 0, sizeof BIT4_TO_STR);
    // invert_chrmap walks idx from 255 down to 0, so the lowest ASCII code
    // mapping to a value wins (uppercase beats lowercase); value 0 stays NUL.
    for (int i = 0; i < 256; i++) {
        int idx = 255 - i;
        uint8_t v = STR2BIT4[1][idx];
        if (v != 0) BIT4_TO_STR[v] = (uint8_t)idx;
    }
}

static inline uint8_t get_bit4(const uint8_t* data, size_t i) {      // bit4ops.rs:281-288
    uint8_t b = data[i / 2];
    return (i % 2 == 0) ? (b & 0x0F) : ((b >> 4) & 0x0F);
}
static inline uint8_t bit4_to_char(uint8_t b) { return BIT4_TO_STR[b & 0x0F]; }  // :293-295
static inline uint8_t char_to_bit4(uint8_t c) { return STR2BIT4[1][c]; }         // :274-276

// bit4_to_string, bit4ops.rs:123-144.  The pair table DBIT4_TO_STR decodes a
// byte as (low nibble, high nibble) with the same per-nibble map, so a
// per-nibble decode is equivalent.
static void bit4_to_string(uint8_t* out, const uint8_t* data, size_t read_offset, size_t n) {
    for (size_t k = 0; k < n; k++) out[k] = bit4_to_char(get_bit4(data, read_offset + k));
}

// string_to_bit4, bit4ops.rs:145-165 (ORs into a zeroed destination).
static void string_to_bit4(uint8_t* out, const uint8_t* s, size_t n, size_t write_offset, bool mixed) {
    for (size_t k = 0; k < n; k++) {
        size_t i = write_offset + k;
        uint8_t v = STR2BIT4[mixed ? 1 : 0][s[k]];
        if (i % 2 == 0) out[i / 2] |= v;
        else out[i / 2] |= (uint8_t)(v << 4);
    }
}

// memsetbit4, bit4ops.rs:184-200
static void memsetbit4(uint8_t* dest, uint8_t v, size_t start, size_t end) {
    for (size_t i = start; i < end; i++) {
        if (i % 2 == 0) dest[i / 2] = (uint8_t)((dest[i / 2] & 0xF0) | v);
        else dest[i / 2] = (uint8_t)((dest[i / 2] & 0x0F) | (v << 4));
    }
}

// bit2_to_bit4, bit4ops.rs:60-75,166-183.  2bit code T=0 C=1 A=2 G=3, first
// base in the two most significant bits; bit4 = 1 << code.
static void bit2_to_bit4(uint8_t* out, const uint8_t* data, size_t n_chrs) {
    size_t n_blks = (n_chrs + 3) / 4;
    for (size_t i = 0; i < n_blks; i++) {
        uint16_t val = 0;
        for (int j = 0; j < 4; j++) {
            int b2 = (data[i] >> ((4 - j - 1) * 2)) & 0x3;
            val |= (uint16_t)((1u << b2) << (j * 4));
        }
        out[2 * i] = (uint8_t)(val & 0xFF);
        out[2 * i + 1] = (uint8_t)(val >> 8);
    }
    if (n_chrs % 4 != 0) memsetbit4(out, 0, n_chrs, n_blks * 4);
}

static bool is_mixedbase_str(const std::string& s) {              // bit4ops.rs:201-206
    for (unsigned char c : s) if (STR2BIT4[1][c] == 0) return false;
    return true;
}

// compliment_char / reverse_compliment_char_i, bit4ops.rs:207-232
static uint8_t compliment_char(uint8_t c) {
    uint8_t b4 = STR2BIT4[1][c];
    uint8_t rev = (uint8_t)(((b4 << 2) | (b4 >> 2)) & 0xF);
    uint8_t rc = b4 != 0 ? BIT4_TO_STR[rev] : c;
    return (uint8_t)(rc | (c & 0x20));
}
static void reverse_compliment_i(std::vector<uint8_t>& v) {
    for (auto& c : v) c = compliment_char(c);
    std::reverse(v.begin(), v.end());
}

static inline bool cmp_chars(uint8_t dna, uint8_t rna) {          // bit4ops.rs:233-237
    return (STR2BIT4[0][dna] & STR2BIT4[1][rna]) != 0;
}

// ---------------------------------------------------------------------------
// Genome readers — read_fasta.rs:9-71, read_2bit.rs:33-118,
// read_fasta_folder.rs:8-23, chrom_chunk.rs:1-29.
// ---------------------------------------------------------------------------
static const size_t CHUNK_SIZE = 1 << 16;
static const size_t CHUNK_SIZE_BYTES = CHUNK_SIZE / 2;

struct ChromChunk {
    std::string name;
    std::vector<uint8_t> data;   // CHUNK_SIZE_BYTES, zero-initialised
    uint64_t start = 0, end = 0;
    size_t size() const { return (size_t)(end - start); }
};

struct Err : std::runtime_error { using std::runtime_error::runtime_error; };

static void read_fasta(std::vector<ChromChunk>& out, const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Err("cannot open " + path);
    bool started = false;
    ChromChunk cur;
    cur.data.assign(CHUNK_SIZE_BYTES, 0);
    std::string line;
    std::vector<char> buf(1 << 20);
    auto handle_line = [&](std::string& ln) {
        // BufRead::lines strips "\n" and a trailing "\r".
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (!ln.empty() && ln[0] == '>') {
            ChromChunk next;
            next.name = ln.substr(1);
            next.data.assign(CHUNK_SIZE_BYTES, 0);
            if (cur.end != cur.start) out.push_back(std::move(cur));
            cur = std::move(next);
            started = true;
        } else {
            if (!started) throw Err("fasta file needs to start with >");
            if (cur.name.empty()) throw Err("> must be followed by chromosome name");
            if (ln.size() + cur.size() > CHUNK_SIZE) {
                ChromChunk next;
                next.name = cur.name;
                next.start = cur.end;
                next.end = cur.end;
                next.data.assign(CHUNK_SIZE_BYTES, 0);
                out.push_back(std::move(cur));
                cur = std::move(next);
            }
            if (ln.size() > CHUNK_SIZE) throw Err("line in fasta too long");
            string_to_bit4(cur.data.data(), (const uint8_t*)ln.data(), ln.size(), cur.size(), true);
            cur.end += ln.size();
        }
    };
    size_t n;
    bool have = false;
    while ((n = fread(buf.data(), 1, buf.size(), f)) > 0) {
        size_t s = 0;
        for (size_t i = 0; i < n; i++) {
            if (buf[i] == '\n') {
                line.append(buf.data() + s, i - s);
                handle_line(line);
                line.clear();
                have = false;
                s = i + 1;
            }
        }
        if (s < n) { line.append(buf.data() + s, n - s); have = true; }
    }
    if (have) handle_line(line);
    fclose(f);
    if (cur.size() > 0) out.push_back(std::move(cur));
}

static uint32_t rd_u32(FILE* f) {
    uint32_t v;
    if (fread(&v, 4, 1, f) != 1) throw Err("2bit: short read");
    return v;
}

static void read_2bit(std::vector<ChromChunk>& out, const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Err("cannot open " + path);
    if (rd_u32(f) != 0x1A412743u) throw Err(".2bit file badly formatted header");
    if (rd_u32(f) != 0) throw Err("only supports version 0 of .2bit format");
    uint32_t chrcnt = rd_u32(f);
    fseek(f, 4, SEEK_CUR);
    std::vector<std::string> names;
    for (uint32_t c = 0; c < chrcnt; c++) {
        int len = fgetc(f);
        if (len == EOF) throw Err("2bit: short read");
        std::string nm(len, '\0');
        if (len && fread(&nm[0], 1, len, f) != (size_t)len) throw Err("2bit: short read");
        names.push_back(nm);
        fseek(f, 4, SEEK_CUR);
    }
    std::vector<uint8_t> raw(CHUNK_SIZE / 4);
    for (auto& name : names) {
        size_t chrlen = rd_u32(f);
        size_t nbc = rd_u32(f);
        std::vector<uint32_t> st(nbc), sz(nbc);
        for (auto& v : st) v = rd_u32(f);
        for (auto& v : sz) v = rd_u32(f);
        std::vector<std::pair<uint32_t, uint32_t>> nb;
        for (size_t i = 0; i < nbc; i++) nb.push_back({st[i], sz[i]});
        std::stable_sort(nb.begin(), nb.end(), [](auto& a, auto& b) { return a.first < b.first; });
        uint32_t mbc = rd_u32(f);
        fseek(f, (long)mbc * 8 + 4, SEEK_CUR);
        size_t read_pos = 0;
        int64_t bmi = 0;
        while (read_pos < chrlen) {
            size_t read_size = std::min(chrlen - read_pos, CHUNK_SIZE);
            size_t nb_bytes = (read_size + 3) / 4;
            if (fread(raw.data(), 1, nb_bytes, f) != nb_bytes) throw Err("2bit: short read");
            ChromChunk ch;
            ch.name = name;
            ch.data.assign(CHUNK_SIZE_BYTES, 0);
            bit2_to_bit4(ch.data.data(), raw.data(), read_size);
            bmi = std::max<int64_t>(bmi - 1, 0);                    // read_2bit.rs:163
            while (bmi < (int64_t)nb.size()) {
                int64_t bs = (int64_t)nb[bmi].first - (int64_t)read_pos;
                int64_t be = (int64_t)(uint32_t)(nb[bmi].first + nb[bmi].second) - (int64_t)read_pos;
                if (bs > (int64_t)read_size) break;
                size_t a = (size_t)std::max<int64_t>(0, bs);
                size_t b = std::min((size_t)std::max<int64_t>(0, be), read_size);
                if (a < b) memsetbit4(ch.data.data(), 0xF, a, b);
                bmi++;
            }
            ch.start = read_pos;
            ch.end = read_pos + read_size;
            out.push_back(std::move(ch));
            read_pos += read_size;
        }
    }
    fclose(f);
}

static void read_genome(std::vector<ChromChunk>& out, const std::string& path) {  // main.rs:106-121
    struct stat stt;
    if (stat(path.c_str(), &stt) != 0) throw Err("cannot stat " + path);
    if (S_ISDIR(stt.st_mode)) {
        // read_fasta_folder: regular files ending in ".fa", in directory order.
        DIR* d = opendir(path.c_str());
        if (!d) throw Err("cannot open dir " + path);
        struct dirent* e;
        while ((e = readdir(d)) != nullptr) {
            std::string nm = e->d_name;
            std::string full = path + "/" + nm;
            struct stat s2;
            if (stat(full.c_str(), &s2) != 0 || !S_ISREG(s2.st_mode)) continue;
            if (nm.size() >= 3 && nm.compare(nm.size() - 3, 3, ".fa") == 0) read_fasta(out, full);
        }
        closedir(d);
        return;
    }
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Err("cannot open " + path);
    int c = fgetc(f);
    fclose(f);
    if (c == '>') read_fasta(out, path);
    else read_2bit(out, path);
}

// ---------------------------------------------------------------------------
// Batcher — search.rs:1376-1427 (chunks_to_searchchunk) and :1525-1572.
// ---------------------------------------------------------------------------
struct Batch {
    std::vector<uint8_t> data;            // search_chunk_bytes, zero tail
    std::vector<size_t> chunk_idx;        // global chunk index of each member
    std::vector<uint32_t> byte_off;       // + sentinel
    bool has_head = false;
};

static void make_batches(const std::vector<ChromChunk>& chunks, size_t batch_nucl,
                         std::vector<std::vector<size_t>>& members, std::vector<bool>& heads) {
    size_t cps = batch_nucl / CHUNK_SIZE;
    std::vector<size_t> buf;
    bool has_head = false;
    for (size_t i = 0; i < chunks.size(); i++) {
        buf.push_back(i);
        if (buf.size() == cps) {
            members.push_back(buf);
            heads.push_back(has_head);
            size_t last = buf.back();
            buf.clear();
            buf.push_back(last);
            has_head = true;
        }
    }
    if (!buf.empty() && !(has_head && buf.size() == 1)) {
        members.push_back(buf);
        heads.push_back(has_head);
    }
}

static Batch build_batch(const std::vector<ChromChunk>& chunks, const std::vector<size_t>& mem,
                         bool has_head, size_t batch_nucl) {
    Batch b;
    b.data.assign(batch_nucl / 2 + 64, 0);   // +64: aligned u64 reads past the end see zeros
    b.has_head = has_head;
    size_t cur = 0;
    for (size_t k = 0; k < mem.size(); k++) {
        const ChromChunk& c = chunks[mem[k]];
        if (k > 0) {
            const ChromChunk& p = chunks[mem[k - 1]];
            if (!(p.end == c.start || c.start == 0)) throw Err("search expects chromosome chunks to arrive in order");
        }
        size_t vb = (c.size() + 1) / 2;
        if (cur + vb > batch_nucl / 2) throw Err("tight-packed chunks exceed batch budget");
        b.byte_off.push_back((uint32_t)cur);
        memcpy(b.data.data() + cur, c.data.data(), vb);
        b.chunk_idx.push_back(mem[k]);
        cur += vb;
    }
    b.byte_off.push_back((uint32_t)cur);
    return b;
}

// Output record after chunk→chromosome conversion (chrom_chunk.rs:15-25).
struct Match {
    size_t chunk;            // global chunk index (for the chromosome name)
    uint64_t chrom_idx;
    uint32_t pattern_idx;
    uint32_t mismatches;
    bool is_forward;
    std::vector<uint8_t> dna, rna;
    uint32_t db = 0, rb = 0;
};

// partition_point(|off| off <= byte_pos) - 1 over the batch offsets (incl. sentinel)
static size_t locate_chunk(const Batch& b, size_t nucl_pos) {
    size_t byte_pos = nucl_pos / 2;
    size_t pp = std::upper_bound(b.byte_off.begin(), b.byte_off.end(), (uint32_t)byte_pos) - b.byte_off.begin();
    return pp == 0 ? 0 : pp - 1;
}

// ---------------------------------------------------------------------------
// Plain mismatch scan (bulge = 0).  search.rs:378-463 (search_chunk_cpu),
// :465-477 (pack_patterns), :512-531 (mask_genome_n_to_zero_inplace),
// :540-546 (legacy_match_start_min_nucl), :1428-1488 (convert_matches).
// Pinned by the 117-row reference count and by a brute-force restatement.
// ---------------------------------------------------------------------------
struct SearchMatch { uint32_t nucl; uint32_t pattern_idx; uint32_t mm; };

static void mask_n_to_zero(std::vector<uint8_t>& d) {
    for (auto& b : d) {
        uint8_t hi = (b >> 4) & 0xF, lo = b & 0xF;
        if (hi == 0xF) hi = 0;
        if (lo == 0xF) lo = 0;
        b = (uint8_t)((hi << 4) | lo);
    }
}

static std::vector<SearchMatch> search_chunk_cpu(uint32_t max_mm, size_t L, const std::vector<uint64_t>& pat_blocks,
                                                 size_t n_patterns, const uint8_t* data, size_t data_len,
                                                 uint32_t min_start) {
    std::vector<SearchMatch> out;
    const size_t NPB = 16;                       // nucleotides per u64 block
    const size_t BPE = 4;                        // BLOCKS_PER_EXEC
    size_t genome_blocks = data_len / 8;
    size_t pattern_blocks = (L + 15) / 16;
    size_t pps = (pattern_blocks + 1) / 2 * 2;   // roundup(pattern_blocks, 2)
    std::vector<uint64_t> g(genome_blocks);
    memcpy(g.data(), data, genome_blocks * 8);
    std::vector<uint64_t> sh(BPE + pattern_blocks + 1);
    for (size_t gb = 0; gb < genome_blocks / BPE; gb++) {
        size_t gi = gb * BPE;
        std::fill(sh.begin(), sh.end(), 0);
        size_t n_copy = std::min(BPE + pattern_blocks + 1, genome_blocks - gi);
        for (size_t k = 0; k < n_copy; k++) sh[k] = g[gi + k];
        for (size_t l = 0; l < NPB; l++) {
            for (size_t j = 0; j < n_patterns; j++) {
                uint32_t nm[BPE] = {0, 0, 0, 0};
                for (size_t k = 0; k < pattern_blocks; k++)
                    for (size_t o = 0; o < BPE; o++)
                        nm[o] += (uint32_t)__builtin_popcountll(sh[k + o] & pat_blocks[j * pps + k]);
                for (size_t o = 0; o < BPE; o++) {
                    int32_t mmi = (int32_t)L - (int32_t)nm[o];
                    if (mmi <= (int32_t)max_mm) {
                        uint32_t start = (uint32_t)((gi + o) * NPB + l);
                        if (start < min_start) continue;
                        out.push_back({start, (uint32_t)j, (uint32_t)std::max(mmi, 0)});
                    }
                }
            }
            for (size_t k = 0; k < pattern_blocks + BPE - 1; k++) {
                sh[k] >>= 4;
                sh[k] |= sh[k + 1] << (4 * (NPB - 1));
            }
            sh[pattern_blocks + BPE - 1] >>= 4;
        }
    }
    return out;
}

static std::vector<uint64_t> pack_patterns(const std::vector<std::vector<uint8_t>>& pats_bit4, size_t L) {
    size_t pb = (L + 1) / 2;
    size_t padded = (pb + 15) / 16 * 16;
    std::vector<uint8_t> bytes;
    for (auto& p : pats_bit4) {
        bytes.insert(bytes.end(), p.begin(), p.end());
        bytes.insert(bytes.end(), padded - pb, 0);
    }
    std::vector<uint64_t> out(bytes.size() / 8);
    memcpy(out.data(), bytes.data(), bytes.size());
    return out;
}

static void convert_matches(const Batch& b, const std::vector<ChromChunk>& chunks, size_t L,
                            const std::vector<std::vector<uint8_t>>& pats_bit4,
                            const std::vector<SearchMatch>& sm, std::vector<Match>& out) {
    size_t nb = b.chunk_idx.size();
    for (auto& m : sm) {
        size_t idx = locate_chunk(b, m.nucl);
        if (idx >= nb) throw Err("match outside any chunk (reference panics here)");
        const ChromChunk& c = chunks[b.chunk_idx[idx]];
        size_t offset = m.nucl - (size_t)b.byte_off[idx] * 2;
        uint64_t pos = c.start + offset;
        bool is_last_chunk = false;   // has_overlap_tail is always false (search.rs:1547,1567)
        bool is_end_chrom = idx == nb - 1 || chunks[b.chunk_idx[idx + 1]].start == 0;
        bool past_end = pos + L > c.end;
        if (is_last_chunk || (is_end_chrom && past_end)) continue;
        Match r;
        r.chunk = b.chunk_idx[idx];
        r.chrom_idx = pos;
        r.pattern_idx = m.pattern_idx;
        r.mismatches = m.mm;
        r.is_forward = m.pattern_idx < pats_bit4.size() / 2;
        r.dna.resize(L);
        r.rna.resize(L);
        bit4_to_string(r.dna.data(), b.data.data(), m.nucl, L);
        bit4_to_string(r.rna.data(), pats_bit4[m.pattern_idx].data(), 0, L);
        if (!r.is_forward) {
            reverse_compliment_i(r.dna);
            reverse_compliment_i(r.rna);
        }
        out.push_back(std::move(r));
    }
}

// ---------------------------------------------------------------------------
// Myers bulge path.  myers.rs:11-50 (build_peq), search.rs:597-688
// (classify_myers_candidate), :703-755 (check_pam_quick), :761-777
// (precompute_pam_precheck), :779-897 (search_chunk_myers), :951-1031,
// :1316-1374 (convert_matches_myers); traceback.rs:129-134, 269-469
// (cmp_loose_bit4, traceback_all, enumerate).
// Pinned by: Myers score == textbook O(mn) semi-global DP; enumeration ==
// brute-force enumeration of every op string on small cases; plan-doc vectors.
// ---------------------------------------------------------------------------
struct Peq { uint64_t peq[4]; uint64_t peq_n; };

static Peq build_peq(const std::vector<uint8_t>& pat) {          // myers.rs:20-50
    Peq t{};
    for (size_t i = 0; i < pat.size(); i++) {
        uint8_t u = (uint8_t)toupper(pat[i]);
        int idx = u == 'A' ? 0 : u == 'C' ? 1 : u == 'G' ? 2 : u == 'T' ? 3 : 4;
        if (idx < 4) t.peq[idx] |= 1ull << i;
        else {
            for (auto& p : t.peq) p |= 1ull << i;
            t.peq_n |= 1ull << i;
        }
    }
    return t;
}

static inline bool cmp_loose_bit4(uint8_t p, uint8_t t) {         // traceback.rs:129-134
    if (t == 0) return false;
    if (p == 0xF) return true;
    if (t == 0xF) return false;
    return (p & t) != 0;
}

enum Op : uint8_t { OP_M = 0, OP_S = 1, OP_DNA = 2, OP_RNA = 3 };

struct Alignment {
    std::vector<uint8_t> ops;      // forward order
    std::vector<uint8_t> pa, ta;
    size_t text_start;
    uint32_t mm, db, rb;
};

struct EnumCtx {
    const uint8_t* pat; size_t m;
    const uint8_t* text; size_t n;
    const std::vector<std::vector<uint32_t>>* dp;
    uint32_t max_edits, max_dna, max_rna, max_mm;
    const std::vector<uint8_t>* pam_is_n;
    std::vector<Alignment>* results;
    std::vector<uint8_t> ops;
};

static const uint32_t INF = 0x7FFFFFFFu;   // u32::MAX / 2
static inline uint32_t sat_add(uint32_t a, uint32_t b) { uint64_t s = (uint64_t)a + b; return s > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s; }

static void enumerate(EnumCtx& c, size_t i, size_t j, uint32_t cost, uint32_t mm, uint32_t db, uint32_t rb) {
    if (sat_add(cost, (*c.dp)[i][j]) > c.max_edits) return;
    if (mm > c.max_mm || db > c.max_dna || rb > c.max_rna) return;
    const auto& isn = *c.pam_is_n;
    if (i == 0) {
        Alignment a;
        a.ops.assign(c.ops.rbegin(), c.ops.rend());
        size_t pi = 0, ti = j;
        for (uint8_t op : a.ops) {
            if (op == OP_M || op == OP_S) { a.pa.push_back(bit4_to_char(c.pat[pi++])); a.ta.push_back(bit4_to_char(c.text[ti++])); }
            else if (op == OP_RNA) { a.pa.push_back(bit4_to_char(c.pat[pi++])); a.ta.push_back('-'); }
            else { a.pa.push_back('-'); a.ta.push_back(bit4_to_char(c.text[ti++])); }
        }
        a.text_start = j;
        a.mm = mm; a.db = db; a.rb = rb;
        c.results->push_back(std::move(a));
        return;
    }
    if (j > 0) {                                                   // diagonal
        uint32_t mc = cmp_loose_bit4(c.pat[i - 1], c.text[j - 1]) ? 0 : 1;
        bool touches_n = mc == 1 && isn[i - 1];
        if (!touches_n) {
            c.ops.push_back(mc == 0 ? OP_M : OP_S);
            enumerate(c, i - 1, j - 1, cost + mc, mm + mc, db, rb);
            c.ops.pop_back();
        }
    }
    if (rb < c.max_rna && c.max_rna > 0) {                          // RNA bulge
        bool pos_n = isn[i - 1];
        bool next_pam = i < c.m && isn[i];
        bool prev_pam = i >= 2 && isn[i - 2];
        if (!pos_n && !next_pam && !prev_pam) {
            c.ops.push_back(OP_RNA);
            enumerate(c, i - 1, j, cost + 1, mm, db, rb + 1);
            c.ops.pop_back();
        }
    }
    if (db < c.max_dna && c.max_dna > 0 && j > 0 && i < c.m) {     // DNA bulge
        bool prev_n = i > 0 && isn[i - 1];
        bool next_n = isn[i];
        if (!(prev_n && next_n)) {
            c.ops.push_back(OP_DNA);
            enumerate(c, i, j - 1, cost + 1, mm, db + 1, rb);
            c.ops.pop_back();
        }
    }
}

static std::vector<Alignment> traceback_all(const std::vector<uint8_t>& pat, const std::vector<uint8_t>& text,
                                            uint32_t max_edits, uint32_t max_dna, uint32_t max_rna,
                                            uint32_t max_mm, const std::vector<uint8_t>& pam_is_n) {
    size_t m = pat.size(), n = text.size();
    uint32_t dna_cost = max_dna == 0 ? INF : 1;
    uint32_t rna_cost = max_rna == 0 ? INF : 1;
    std::vector<std::vector<uint32_t>> dp(m + 1, std::vector<uint32_t>(n + 1, 0));
    for (size_t i = 0; i <= m; i++) {
        uint64_t v = (uint64_t)i * rna_cost;
        uint32_t s = v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
        dp[i][0] = std::min(s, INF);
    }
    for (size_t i = 1; i <= m; i++)
        for (size_t j = 1; j <= n; j++) {
            if (text[j - 1] == 0) { dp[i][j] = INF; continue; }
            uint32_t mc = cmp_loose_bit4(pat[i - 1], text[j - 1]) ? 0 : 1;
            dp[i][j] = std::min(sat_add(dp[i - 1][j - 1], mc),
                                std::min(sat_add(dp[i - 1][j], rna_cost), sat_add(dp[i][j - 1], dna_cost)));
        }
    std::vector<Alignment> res;
    if (dp[m][n] > max_edits) return res;
    EnumCtx c{pat.data(), m, text.data(), n, &dp, max_edits, max_dna, max_rna, max_mm, &pam_is_n, &res, {}};
    enumerate(c, m, n, 0, 0, 0, 0);
    return res;
}

struct MyersQuery {
    size_t L;
    uint32_t M, D, R;
    std::vector<std::vector<uint8_t>> pats_ascii;     // fwd then revcomp
    std::vector<std::vector<uint8_t>> pats_bit4;      // one nibble per byte
    std::vector<std::vector<uint8_t>> is_n;
    std::vector<std::vector<uint8_t>> eff_filter;     // bit4 per position
    std::vector<size_t> pam_off;
    std::vector<std::vector<uint8_t>> pam_filter;
    std::vector<Peq> peqs;
};

static MyersQuery prepare_myers(const std::vector<std::vector<uint8_t>>& pats, const std::string& filter,
                                uint32_t M, uint32_t D, uint32_t R) {           // search.rs:961-999
    MyersQuery q;
    q.L = pats[0].size(); q.M = M; q.D = D; q.R = R;
    q.pats_ascii = pats;
    size_t n_orig = pats.size() / 2;
    std::vector<uint8_t> f(filter.begin(), filter.end()), frc = f;
    reverse_compliment_i(frc);
    for (size_t p = 0; p < pats.size(); p++) {
        q.peqs.push_back(build_peq(pats[p]));
        std::vector<uint8_t> isn, ef, pb;
        for (uint8_t ch : pats[p]) { isn.push_back(ch == 'N' || ch == 'n'); pb.push_back(char_to_bit4(ch)); }
        const auto& fa = p < n_orig ? f : frc;
        for (uint8_t ch : fa) ef.push_back(char_to_bit4(ch));
        size_t ps = std::find(isn.begin(), isn.end(), 1) - isn.begin();
        std::vector<uint8_t> pf;
        for (size_t i = ps; i < isn.size(); i++) if (isn[i]) pf.push_back(ef[i]);
        q.is_n.push_back(isn); q.eff_filter.push_back(ef); q.pats_bit4.push_back(pb);
        q.pam_off.push_back(ps); q.pam_filter.push_back(pf);
    }
    return q;
}

static bool check_pam_quick(const uint8_t* g, size_t total, size_t end_pos, size_t L, size_t pam_off,
                            const std::vector<uint8_t>& pf, uint32_t D, uint32_t R) {
    int64_t dbr = pam_off == 0 ? (int64_t)D + 1 : 1, rbr = pam_off == 0 ? (int64_t)R + 1 : 1;
    for (int64_t bd = 0; bd < dbr; bd++)
        for (int64_t br = 0; br < rbr; br++) {
            int64_t span = (int64_t)L + bd - br;
            if (span <= 0 || span > (int64_t)end_pos + 1) continue;
            size_t as = end_pos + 1 - (size_t)span;
            bool ok = true;
            for (size_t k = 0; k < pf.size(); k++) {
                size_t gp = as + pam_off + k;
                if (gp >= total) { ok = false; break; }
                uint8_t gv = get_bit4(g, gp), fv = pf[k];
                bool pos_ok = fv == 0xF ? gv != 0 : (gv == 0 || gv == 0xF) ? false : (gv & fv) != 0;
                if (!pos_ok) { ok = false; break; }
            }
            if (ok) return true;
        }
    return false;
}

static int myers_window_score(const Peq& peq, size_t L, const uint8_t* g, size_t text_start, size_t t_pos) {
    uint64_t mask = L < 64 ? (1ull << L) - 1 : ~0ull;
    uint64_t last = 1ull << (L - 1);
    uint64_t vp = mask, vn = 0;
    int score = (int)L;
    for (size_t t = text_start; t <= t_pos; t++) {
        uint8_t b4 = get_bit4(g, t);
        uint64_t eq;
        if (b4 == 0) eq = 0;
        else if (b4 == 0xF) eq = peq.peq_n;
        else {
            eq = 0;
            if (b4 & 4) eq |= peq.peq[0];
            if (b4 & 2) eq |= peq.peq[1];
            if (b4 & 8) eq |= peq.peq[2];
            if (b4 & 1) eq |= peq.peq[3];
        }
        uint64_t x = eq | vn;
        uint64_t d0 = (((x & vp) + vp) ^ vp) | x;
        uint64_t hn = vp & d0;
        uint64_t hp = vn | ~(vp | d0);
        uint64_t xs = hp << 1;
        vn = xs & d0;
        vp = (hn << 1) | ~(xs | d0);
        vp &= mask; vn &= mask;
        if (hp & last) score++;
        if (hn & last) score--;
    }
    return score;
}

struct MyersHit { uint32_t start; uint32_t pattern_idx; uint32_t mm, db, rb; std::vector<uint8_t> pa, ta; };

static std::vector<MyersHit> search_chunk_myers(const MyersQuery& q, const uint8_t* data, size_t data_nucl,
                                                size_t active_start, size_t active_end, uint64_t* n_cand) {
    std::vector<MyersHit> out;
    size_t total = std::min(active_end, data_nucl);
    size_t W = q.L + q.D;
    uint32_t E = q.M + q.D + q.R;
    for (size_t p = 0; p < q.pats_ascii.size(); p++) {
        for (size_t t = active_start; t < total; t++) {
            if (t + 1 < q.L) continue;
            if (!check_pam_quick(data, total, t, q.L, q.pam_off[p], q.pam_filter[p], q.D, q.R)) continue;
            size_t ts = t + 1 >= W ? t + 1 - W : 0;
            int score = myers_window_score(q.peqs[p], q.L, data, ts, t);
            if (score < 0 || (uint32_t)score > E) continue;
            if (n_cand) (*n_cand)++;
            // classify_myers_candidate, search.rs:597-688
            std::vector<uint8_t> text;
            for (size_t i = ts; i <= t; i++) text.push_back(get_bit4(data, i));
            size_t text_len = text.size();
            auto al = traceback_all(q.pats_bit4[p], text, E, q.D, q.R, q.M, q.is_n[p]);
            for (auto& a : al) {
                size_t span = 0;
                for (uint8_t ch : a.ta) if (ch != '-') span++;
                size_t off = text_len - span;
                size_t pip = 0, go = off;
                bool ok = true;
                for (uint8_t op : a.ops) {
                    if (op == OP_M || op == OP_S) {
                        uint8_t gv = get_bit4(data, ts + go), fv = q.eff_filter[p][pip];
                        bool fo = fv == 0xF ? gv != 0 : (gv == 0 || gv == 0xF) ? false : (gv & fv) != 0;
                        if (!fo) { ok = false; break; }
                        pip++; go++;
                    } else if (op == OP_DNA) go++;
                    else pip++;
                }
                if (!ok) continue;
                MyersHit h;
                h.start = (uint32_t)(t + 1 - span);
                h.pattern_idx = (uint32_t)p;
                h.mm = a.mm; h.db = a.db; h.rb = a.rb;
                h.pa = a.pa; h.ta = a.ta;
                out.push_back(std::move(h));
            }
        }
    }
    return out;
}

static void convert_matches_myers(const Batch& b, const std::vector<ChromChunk>& chunks, size_t L, size_t n_pat,
                                  std::vector<MyersHit>& hits, std::vector<Match>& out) {
    size_t n_orig = n_pat / 2;
    size_t nb = b.chunk_idx.size();
    for (auto& h : hits) {
        size_t idx = locate_chunk(b, h.start);
        if (idx >= nb) throw Err("match outside any chunk");
        const ChromChunk& c = chunks[b.chunk_idx[idx]];
        uint64_t pos = c.start + (h.start - (size_t)b.byte_off[idx] * 2);
        bool is_end_chrom = idx == nb - 1 || chunks[b.chunk_idx[idx + 1]].start == 0;
        uint64_t extent = (uint64_t)L + h.db - h.rb;
        bool past_end = pos + extent > c.end;
        if (is_end_chrom && past_end) continue;
        Match r;
        r.chunk = b.chunk_idx[idx];
        r.chrom_idx = pos;
        r.pattern_idx = h.pattern_idx;
        r.mismatches = h.mm;
        r.is_forward = h.pattern_idx < n_orig;
        r.dna = h.ta;
        r.rna = h.pa;
        if (!r.is_forward) { reverse_compliment_i(r.dna); reverse_compliment_i(r.rna); }
        r.db = h.db; r.rb = h.rb;
        out.push_back(std::move(r));
    }
}

// ---------------------------------------------------------------------------
// Input parsing — cas-offinder-cli/src/cli_utils.rs:37-125.
// Writer — cas-offinder-cli/src/main.rs:80-91, 130-247.
// ---------------------------------------------------------------------------
struct RunInfo {
    std::string genome_path;
    std::string filter;
    std::vector<std::vector<uint8_t>> patterns;
    size_t L = 0;
    uint32_t M = 0, D = 0, R = 0;
};

static std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

static bool parse_u32(const std::string& s, uint32_t& v) {
    if (s.empty() || s.size() > 10) return false;
    uint64_t x = 0;
    for (char c : s) { if (c < '0' || c > '9') return false; x = x * 10 + (c - '0'); }
    if (x > 0xFFFFFFFFull) return false;
    v = (uint32_t)x;
    return true;
}

static RunInfo parse_input(const std::string& text) {
    std::vector<std::string> lines;
    size_t s = 0;
    for (size_t i = 0; i <= text.size(); i++) {
        if (i == text.size() || text[i] == '\n') {
            if (i == text.size() && s == i) break;
            std::string ln = text.substr(s, i - s);
            if (!ln.empty() && ln.back() == '\r') ln.pop_back();
            lines.push_back(ln);
            s = i + 1;
        }
    }
    if (lines.size() < 2) throw Err("Input file must contain at least 3 lines");
    RunInfo ri;
    ri.genome_path = lines[0];
    auto parts = split_ws(lines[1]);
    if (parts.size() == 1) { ri.filter = parts[0]; }
    else if (parts.size() == 3) {
        ri.filter = parts[0];
        if (!parse_u32(parts[1], ri.D)) throw Err("bulge_dna must be a non-negative integer");
        if (!parse_u32(parts[2], ri.R)) throw Err("bulge_rna must be a non-negative integer");
    } else throw Err("2nd line must be: <search_filter> [<bulge_dna> <bulge_rna>]");
    if (!is_mixedbase_str(ri.filter)) throw Err("Pattern in input file needs to be a mixed base string");
    ri.L = ri.filter.size();
    bool have_mm = false, have_info = false, using_info = false;
    for (size_t li = 2; li < lines.size(); li++) {
        auto lp = split_ws(lines[li]);
        if (lp.size() != 2 && lp.size() != 3) throw Err("Pattern line must have following elements: <patterns> <mismatches> [<label>]");
        if (!is_mixedbase_str(lp[0])) throw Err("Pattern in input file needs to be a mixed base string");
        uint32_t cm;
        if (!parse_u32(lp[1], cm)) throw Err("2nd element of each pattern in mismatches must be an unsigned integer");
        if (!have_mm) { ri.M = cm; have_mm = true; }
        if (!have_info) { using_info = lp.size() == 3; have_info = true; }
        if (ri.M != cm) throw Err("In this version of cas-offinder, all mismatches on every line of input file must be the same");
        if (using_info && lp.size() < 3) throw Err("Pattern lines in input file must be consistently have either 2 or 3 elements, no mixing and matching allowed");
        if (lp[0].size() != ri.L) throw Err("All patters in input file must be same length");
        ri.patterns.emplace_back(lp[0].begin(), lp[0].end());
    }
    if (!have_mm) throw Err("Input file must contain at least 1 pattern line");
    return ri;
}

struct Options {
    size_t batch_nucl = 1 << 22;   // CPU default, search.rs:21
    unsigned threads = 0;          // 0 = hardware_concurrency (available_parallelism)
    double search_seconds = 0;
    uint64_t n_candidates = 0;     // Myers survivors (diagnostic)
};

static void write_rows(FILE* out, const RunInfo& ri, bool use_myers, const std::vector<ChromChunk>& chunks,
                       std::vector<Match>& ms, uint64_t& total) {
    uint64_t pam_len = 0;                                          // main.rs:80-91
    for (auto it = ri.patterns[0].rbegin(); it != ri.patterns[0].rend(); ++it) {
        if (*it == 'N' || *it == 'n') pam_len++; else break;
    }
    std::string line;
    for (auto& m : ms) {
        const char* btype;
        uint32_t bsize;
        std::vector<uint8_t> dna_out;
        if (use_myers) {
            uint32_t tot = m.db + m.rb;
            btype = tot == 0 ? "X" : m.db > 0 ? "DNA" : "RNA";
            bsize = tot;
            dna_out = m.dna;
        } else {
            bool pam_ok = true;
            uint32_t amm = 0;
            for (size_t p = 0; p < m.dna.size(); p++) {
                uint8_t fc = (uint8_t)ri.filter[p];
                if (!(fc == 'N' || fc == 'n')) {
                    if (!cmp_chars(m.dna[p], fc)) { pam_ok = false; break; }
                }
                uint8_t rc = m.rna[p];
                if (!(rc == 'N' || rc == 'n') && !cmp_chars(m.dna[p], rc)) amm++;
            }
            if (!pam_ok || amm > ri.M) continue;
            m.mismatches = amm;
            dna_out = m.dna;
            for (size_t p = 0; p < dna_out.size(); p++) {
                uint8_t& d = dna_out[p];
                uint8_t r = m.rna[p];
                bool rn = r == 'N' || r == 'n';
                bool dn = d == 'N' || d == 'n';
                if (d == 0) d = rn ? 'N' : 'n';
                else if (rn) {}
                else if (dn) d = 'n';
                else if (!cmp_chars(d, r)) d |= 0x20;
            }
            btype = "X";
            bsize = 0;
        }
        uint64_t pos = m.is_forward ? m.chrom_idx
                                    : m.chrom_idx + (pam_len > 0 ? pam_len - 1 : 0) - std::min<uint64_t>(pam_len > 0 ? pam_len - 1 : 0, m.db);
        line.clear();
        line += btype; line += '\t';
        line.append((const char*)m.rna.data(), m.rna.size()); line += '\t';
        line.append((const char*)dna_out.data(), dna_out.size()); line += '\t';
        line += chunks[m.chunk].name; line += '\t';
        line += std::to_string(pos); line += '\t';
        line += m.is_forward ? '+' : '-'; line += '\t';
        line += std::to_string(m.mismatches); line += '\t';
        line += std::to_string(bsize); line += '\n';
        fwrite(line.data(), 1, line.size(), out);
        total++;
    }
}

// Full run: parse → read → batches → per-batch search (thread pool) → convert → write.
static uint64_t run(const std::string& in_text, FILE* out, Options& opt) {
    init_tables();
    RunInfo ri = parse_input(in_text);
    std::vector<ChromChunk> chunks;
    read_genome(chunks, ri.genome_path);
    std::vector<std::vector<uint8_t>> all = ri.patterns;             // main.rs:253-259
    for (auto& p : ri.patterns) { auto r = p; reverse_compliment_i(r); all.push_back(r); }
    bool use_myers = ri.D > 0 || ri.R > 0;
    size_t L = ri.L;
    std::vector<std::vector<size_t>> members;
    std::vector<bool> heads;
    make_batches(chunks, opt.batch_nucl, members, heads);
    unsigned nt = opt.threads ? opt.threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<Match>> results(members.size());
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> ncand{0};
    std::vector<std::vector<uint8_t>> pats_b4;                      // search.rs:1516-1523
    for (auto& p : all) {
        std::vector<uint8_t> b((L + 1) / 2, 0);
        string_to_bit4(b.data(), p.data(), p.size(), 0, true);
        pats_b4.push_back(b);
    }
    MyersQuery mq;
    if (use_myers) mq = prepare_myers(all, ri.filter, ri.M, ri.D, ri.R);
    std::vector<uint64_t> packed = pack_patterns(pats_b4, L);
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        for (;;) {
            size_t bi = next.fetch_add(1);
            if (bi >= members.size()) return;
            Batch b = build_batch(chunks, members[bi], heads[bi], opt.batch_nucl);
            std::vector<Match> res;
            if (!use_myers) {
                uint32_t min_start = 0;
                if (b.has_head) {
                    size_t c1 = (size_t)b.byte_off[1] * 2;
                    min_start = (uint32_t)(c1 >= L - 1 ? c1 - (L - 1) : 0);
                }
                mask_n_to_zero(b.data);
                auto sm = search_chunk_cpu(ri.M, L, packed, all.size(), b.data.data(), opt.batch_nucl / 2, min_start);
                convert_matches(b, chunks, L, pats_b4, sm, res);
            } else {
                size_t n = b.chunk_idx.size();
                size_t as = b.has_head ? (size_t)b.byte_off[1] * 2 : 0;
                size_t ae = (size_t)b.byte_off[n] * 2;
                uint64_t nc = 0;
                auto hits = search_chunk_myers(mq, b.data.data(), opt.batch_nucl, as, ae, &nc);
                ncand += nc;
                convert_matches_myers(b, chunks, L, all.size(), hits, res);
            }
            results[bi] = std::move(res);
        }
    };
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; i++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    opt.search_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    opt.n_candidates = ncand.load();
    fputs("#Bulge type\tcrRNA\tDNA\tChromosome\tPosition\tDirection\tMismatches\tBulge Size\n", out);
    uint64_t total = 0;
    for (auto& r : results) write_rows(out, ri, use_myers, chunks, r, total);
    return total;
}

}  // namespace ref

// ---------------------------------------------------------------------------
// C entry points for the tests (ctypes).  Strings are ASCII.
// ---------------------------------------------------------------------------
extern "C" {

// Runs a whole search: in_text is the content of a reference-format input
// file; rows go to out_path.  Returns the row count, or -1 on error (message
// in err, up to err_cap bytes).
int64_t cofref_run(const char* in_text, const char* out_path, uint64_t batch_nucl, unsigned threads,
                   double* search_seconds, uint64_t* n_candidates, char* err, size_t err_cap) {
    try {
        ref::Options o;
        if (batch_nucl) o.batch_nucl = batch_nucl;
        o.threads = threads;
        FILE* f = fopen(out_path, "wb");
        if (!f) throw ref::Err("cannot open output");
        uint64_t n = ref::run(in_text, f, o);
        fclose(f);
        if (search_seconds) *search_seconds = o.search_seconds;
        if (n_candidates) *n_candidates = o.n_candidates;
        return (int64_t)n;
    } catch (std::exception& e) {
        if (err && err_cap) { strncpy(err, e.what(), err_cap - 1); err[err_cap - 1] = 0; }
        return -1;
    }
}

// Decodes a genome (FASTA / 2bit / folder) into the concatenated chunk text
// (file_test.rs:17-25) and the chunk table "name\tstart\tend\n".
int64_t cofref_decode_genome(const char* path, char* seq_out, size_t seq_cap, char* table_out, size_t table_cap) {
    try {
        ref::init_tables();
        std::vector<ref::ChromChunk> ch;
        ref::read_genome(ch, path);
        size_t k = 0;
        std::string tab;
        for (auto& c : ch) {
            if (k + c.size() > seq_cap) return -2;
            ref::bit4_to_string((uint8_t*)seq_out + k, c.data.data(), 0, c.size());
            k += c.size();
            tab += c.name + "\t" + std::to_string(c.start) + "\t" + std::to_string(c.end) + "\n";
        }
        if (table_out) {
            if (tab.size() + 1 > table_cap) return -3;
            memcpy(table_out, tab.c_str(), tab.size() + 1);
        }
        return (int64_t)k;
    } catch (...) { return -1; }
}

// Windowed Myers score (search.rs:841-877) of an ASCII pattern over an ASCII
// text (genome encoding: mixed-base, N -> 0xF, other chars -> 0).
int cofref_myers_score(const char* pattern, const char* text) {
    ref::init_tables();
    std::vector<uint8_t> p(pattern, pattern + strlen(pattern));
    size_t n = strlen(text);
    std::vector<uint8_t> g((n + 1) / 2 + 8, 0);
    ref::string_to_bit4(g.data(), (const uint8_t*)text, n, 0, true);
    if (n == 0) return (int)p.size();
    return ref::myers_window_score(ref::build_peq(p), p.size(), g.data(), 0, n - 1);
}

// traceback_all (traceback.rs:269-327) on ASCII pattern / text; pam_is_n from
// the pattern's ASCII N.  Writes "pa\tta\tmm\tdb\trb\ttext_start\n" lines.
int64_t cofref_traceback_all(const char* pattern, const char* text, uint32_t max_edits, uint32_t D, uint32_t R,
                             uint32_t M, char* out, size_t cap) {
    ref::init_tables();
    std::vector<uint8_t> p, t, isn;
    for (const char* c = pattern; *c; c++) { p.push_back(ref::char_to_bit4((uint8_t)*c)); isn.push_back(*c == 'N' || *c == 'n'); }
    for (const char* c = text; *c; c++) t.push_back(ref::STR2BIT4[1][(uint8_t)*c]);
    auto al = ref::traceback_all(p, t, max_edits, D, R, M, isn);
    std::string s;
    for (auto& a : al) {
        s.append((const char*)a.pa.data(), a.pa.size()); s += '\t';
        s.append((const char*)a.ta.data(), a.ta.size()); s += '\t';
        s += std::to_string(a.mm) + "\t" + std::to_string(a.db) + "\t" + std::to_string(a.rb) + "\t" + std::to_string(a.text_start) + "\n";
    }
    if (s.size() + 1 > cap) return -2;
    memcpy(out, s.c_str(), s.size() + 1);
    return (int64_t)al.size();
}

}  // extern "C"

#ifdef COF_REF_MAIN
// CLI twin of cas-offinder-cli (main.rs:55-275): cof-ref <in|-> C <out|->
// [--batch-nucl N] [--threads T].  Prints "Completed in Xs" to stderr.
int main(int argc, char** argv) {
    if (argc < 4) { fprintf(stderr, "usage: cof-ref <input|-> C <output|-> [--batch-nucl N] [--threads T]\n"); return 2; }
    auto t0 = std::chrono::steady_clock::now();
    ref::Options o;
    for (int i = 4; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--batch-nucl")) o.batch_nucl = strtoull(argv[i + 1], nullptr, 10);
        else if (!strcmp(argv[i], "--threads")) o.threads = (unsigned)atoi(argv[i + 1]);
    }
    std::string in;
    {
        FILE* f = strcmp(argv[1], "-") ? fopen(argv[1], "rb") : stdin;
        if (!f) { fprintf(stderr, "cannot open input\n"); return 1; }
        char b[65536];
        size_t n;
        while ((n = fread(b, 1, sizeof b, f)) > 0) in.append(b, n);
        if (f != stdin) fclose(f);
    }
    FILE* out = strcmp(argv[3], "-") ? fopen(argv[3], "wb") : stdout;
    if (!out) { fprintf(stderr, "cannot open output\n"); return 1; }
    try {
        uint64_t n = ref::run(in, out, o);
        if (out != stdout) fclose(out);
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fprintf(stderr, "Completed in %gs (search %gs, %llu rows, %llu myers candidates)\n", s, o.search_seconds,
                (unsigned long long)n, (unsigned long long)o.n_candidates);
    } catch (std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
#endif
