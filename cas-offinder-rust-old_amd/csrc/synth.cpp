// libcofsynth: seeded synthetic genomes for tests and benchmarks.
//
// Produces the tight-packed bit4 chunk layout that reading a FASTA file with
// fixed-width lines yields (a chunk ends when the next line would pass 64 Knt),
// so a genome can be made resident directly (cof_load_genome) or written as
// FASTA for a file-based run.  Composition, N gaps, repeat families and
// planted guide sites follow the benchmark workloads:
//   bases A=T=0.295, C=G=0.205 (41% GC); N in log-uniform gaps;
//   repeat elements copied with a given divergence in place of random
//   sequence; planted sites = guide + sampled PAM with 0..k substitutions in
//   the guide part, random strand, non-overlapping, away from N.
// PRNG: splitmix64 seeding xoshiro256**, one stream per chunk / feature.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <set>
#include <thread>
#include <vector>

#include "cof.h"

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        for (auto& v : s) v = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const uint8_t kBases[4] = {4, 2, 8, 1};  // A C G T as bit4

inline uint8_t draw_base(uint16_t u) {
    // A=T=0.295, C=G=0.205 with 16-bit resolution
    if (u < 19333) return 4;       // A
    if (u < 32768) return 2;       // C (13435)
    if (u < 46203) return 8;       // G (13435)
    return 1;                      // T
}

uint8_t enc(char c) {
    switch (c | 0x20) {
        case 'a': return 4; case 'c': return 2; case 'g': return 8; case 't': return 1;
        case 'r': return 12; case 'y': return 3; case 's': return 10; case 'w': return 5;
        case 'k': return 9; case 'm': return 6; case 'b': return 11; case 'd': return 13;
        case 'h': return 7; case 'v': return 14; case 'n': return 15;
        default: return 0;
    }
}
const char kDec[17] = "\0TCYAWMHGKSBRDVN";

struct Plant {
    std::string site;
    uint32_t guide_off, guide_len, max_subs, count;
};

}  // namespace

struct cofsynth {
    uint64_t seed;
    std::vector<uint64_t> lens;
    uint32_t width;
    double n_frac = 0;
    uint64_t gap_min = 0, gap_max = 0;
    double rep_frac = 0, rep_div = 0;
    uint32_t rep_elems = 0, rep_len = 0;
    double iupac_frac = 0;
    std::vector<Plant> plants;
    std::vector<std::string> names;
    std::vector<cof_chunk> chunks;
    std::vector<uint8_t> packed;
    std::vector<uint64_t> chrom_first_chunk;
    uint64_t chunk_nt = 0;
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> n_gaps;    // per chrom, sorted
    std::vector<std::set<std::pair<uint64_t, uint64_t>>> occupied;     // planted intervals

    uint64_t gnib(uint32_t chrom, uint64_t pos) const {
        // every chunk but a chromosome's last holds exactly chunk_nt
        uint64_t n_chunks = chrom_first_chunk[chrom + 1] - chrom_first_chunk[chrom];
        uint64_t k = std::min<uint64_t>(pos / chunk_nt, n_chunks - 1);
        const cof_chunk& c = chunks[chrom_first_chunk[chrom] + k];
        return c.byte_offset * 2 + (pos - c.start);
    }
    void put(uint32_t chrom, uint64_t pos, uint8_t v) {
        uint64_t i = gnib(chrom, pos);
        uint8_t& b = packed[i >> 1];
        b = (i & 1) ? (uint8_t)((b & 0x0F) | (v << 4)) : (uint8_t)((b & 0xF0) | v);
    }
    uint8_t get(uint32_t chrom, uint64_t pos) const {
        uint64_t i = gnib(chrom, pos);
        uint8_t b = packed[i >> 1];
        return (i & 1) ? (uint8_t)(b >> 4) : (uint8_t)(b & 0xF);
    }
};

extern "C" {

cofsynth* cofsynth_new(uint64_t seed, uint32_t n_chroms, const uint64_t* lens, uint32_t line_width) {
    if (!lens || n_chroms == 0 || line_width == 0 || line_width > 65536) return nullptr;
    auto* s = new cofsynth();
    s->seed = seed;
    s->lens.assign(lens, lens + n_chroms);
    s->width = line_width;
    for (uint32_t i = 0; i < n_chroms; i++) s->names.push_back("chr" + std::to_string(i + 1));
    return s;
}

void cofsynth_set_n(cofsynth* s, double frac, uint64_t gap_min, uint64_t gap_max) {
    s->n_frac = frac; s->gap_min = gap_min; s->gap_max = std::max(gap_min, gap_max);
}
void cofsynth_set_repeats(cofsynth* s, double frac, uint32_t n_elems, uint32_t elem_len, double div) {
    s->rep_frac = frac; s->rep_elems = n_elems; s->rep_len = elem_len; s->rep_div = div;
}
void cofsynth_set_iupac(cofsynth* s, double frac) { s->iupac_frac = frac; }

int cofsynth_plant(cofsynth* s, const char* site, uint32_t guide_off, uint32_t guide_len, uint32_t max_subs,
                   uint32_t count) {
    std::string st = site;
    if (guide_off + guide_len > st.size()) return -1;
    s->plants.push_back({st, guide_off, guide_len, max_subs, count});
    return 0;
}

// Places `count` non-overlapping intervals of `len` in [0, total) uniformly:
// pick sorted offsets in the free length, then shift by earlier intervals.
static std::vector<uint64_t> spread(Rng& r, uint64_t total, uint64_t len, uint64_t count) {
    std::vector<uint64_t> out;
    if (count == 0 || len * count > total) return out;
    uint64_t free_len = total - len * count;
    for (uint64_t i = 0; i < count; i++) out.push_back(r.below(free_len + 1));
    std::sort(out.begin(), out.end());
    for (uint64_t i = 0; i < count; i++) out[i] += i * len;
    return out;
}

int cofsynth_build(cofsynth* s) {
    // chunk layout of a FASTA with fixed-width lines
    s->chunk_nt = (65536 / s->width) * s->width;
    uint64_t off = 0;
    s->chunks.clear();
    s->chrom_first_chunk.clear();
    // A FASTA reader flushes a chunk when the next line would pass 64 Knt, so
    // full chunks hold chunk_nt and a short last line joins the previous chunk
    // when it still fits.
    for (uint32_t c = 0; c < s->lens.size(); c++) {
        s->chrom_first_chunk.push_back(s->chunks.size());
        uint64_t len = s->lens[c], full = len / s->width * s->width, p = 0;
        std::vector<std::pair<uint64_t, uint64_t>> spans;
        while (p < full) {
            uint64_t e = std::min(p + s->chunk_nt, full);
            spans.push_back({p, e});
            p = e;
        }
        if (len > full) {
            if (!spans.empty() && spans.back().second - spans.back().first + (len - full) <= 65536) spans.back().second = len;
            else spans.push_back({full, len});
        }
        for (auto& sp : spans) {
            cof_chunk ch{};
            ch.chrom = c;
            ch.start = sp.first;
            ch.end = sp.second;
            ch.byte_offset = off;
            off += (ch.end - ch.start + 1) / 2;
            s->chunks.push_back(ch);
        }
    }
    s->chrom_first_chunk.push_back(s->chunks.size());
    s->packed.assign(off, 0);
    // random bases, chunk-parallel
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([s, t, nt]() {
            for (size_t ci = t; ci < s->chunks.size(); ci += nt) {
                const cof_chunk& c = s->chunks[ci];
                Rng r(s->seed * 0x100000001B3ull + ci * 0x9E3779B97F4A7C15ull + 1);
                uint64_t n = c.end - c.start;
                uint8_t* out = s->packed.data() + c.byte_offset;
                uint64_t bits = 0;
                int avail = 0;
                for (uint64_t i = 0; i < n; i += 2) {
                    if (avail < 2) { bits = r.next(); avail = 4; }
                    uint8_t a = draw_base((uint16_t)bits);
                    uint8_t b = i + 1 < n ? draw_base((uint16_t)(bits >> 16)) : 0;
                    bits >>= 32;
                    avail -= 2;
                    out[i / 2] = (uint8_t)(a | (b << 4));
                }
            }
        });
    }
    for (auto& x : th) x.join();
    Rng r(s->seed ^ 0x5EEDF00Dull);
    // repeat families written in place of random sequence
    if (s->rep_frac > 0 && s->rep_elems && s->rep_len) {
        std::vector<std::vector<uint8_t>> elems(s->rep_elems);
        for (auto& e : elems)
            for (uint32_t i = 0; i < s->rep_len; i++) e.push_back(draw_base((uint16_t)r.next()));
        for (uint32_t c = 0; c < s->lens.size(); c++) {
            uint64_t copies = (uint64_t)(s->rep_frac * (double)s->lens[c] / s->rep_len);
            for (uint64_t p : spread(r, s->lens[c], s->rep_len, copies)) {
                const auto& e = elems[r.below(s->rep_elems)];
                for (uint32_t i = 0; i < s->rep_len; i++) {
                    uint8_t v = e[i];
                    if (r.unit() < s->rep_div) v = kBases[(std::find(kBases, kBases + 4, v) - kBases + 1 + r.below(3)) % 4];
                    s->put(c, p + i, v);
                }
            }
        }
    }
    // N gaps (log-uniform sizes)
    s->n_gaps.assign(s->lens.size(), {});
    if (s->n_frac > 0 && s->gap_min > 0) {
        for (uint32_t c = 0; c < s->lens.size(); c++) {
            uint64_t target = (uint64_t)(s->n_frac * (double)s->lens[c]);
            std::vector<uint64_t> sizes;
            uint64_t tot = 0;
            while (tot < target) {
                double lo = std::log((double)s->gap_min), hi = std::log((double)s->gap_max);
                uint64_t g = (uint64_t)std::exp(lo + (hi - lo) * r.unit());
                g = std::max<uint64_t>(1, std::min(g, target - tot));
                sizes.push_back(g);
                tot += g;
            }
            if (tot >= s->lens[c]) continue;
            uint64_t free_len = s->lens[c] - tot;
            std::vector<uint64_t> at;
            for (size_t i = 0; i < sizes.size(); i++) at.push_back(r.below(free_len + 1));
            std::sort(at.begin(), at.end());
            uint64_t shift = 0;
            for (size_t i = 0; i < sizes.size(); i++) {
                uint64_t p = at[i] + shift;
                for (uint64_t k = 0; k < sizes[i]; k++) s->put(c, p + k, 0xF);
                s->n_gaps[c].push_back({p, p + sizes[i]});
                shift += sizes[i];
            }
        }
    }
    // sprinkled IUPAC letters (exercise the exact-predicate paths)
    if (s->iupac_frac > 0) {
        const uint8_t iu[] = {12, 3, 10, 5, 9, 6, 11, 13, 7, 14};
        for (uint32_t c = 0; c < s->lens.size(); c++) {
            uint64_t k = (uint64_t)(s->iupac_frac * (double)s->lens[c]);
            for (uint64_t i = 0; i < k; i++) s->put(c, r.below(s->lens[c]), iu[r.below(10)]);
        }
    }
    // planted sites
    uint64_t total = 0;
    for (auto l : s->lens) total += l;
    s->occupied.assign(s->lens.size(), {});
    for (const auto& pl : s->plants) {
        uint64_t len = pl.site.size();
        for (uint32_t k = 0; k < pl.count; k++) {
            for (int attempt = 0; attempt < 64; attempt++) {
                uint64_t x = r.below(total);
                uint32_t c = 0;
                while (x >= s->lens[c]) { x -= s->lens[c]; c++; }
                if (x + len > s->lens[c]) continue;
                bool bad = false;
                for (uint64_t i = 0; i < len && !bad; i++) bad = s->get(c, x + i) == 0xF;
                if (!bad) {
                    // neighbours in the start-sorted interval set
                    auto it = s->occupied[c].lower_bound({x, 0});
                    if (it != s->occupied[c].end() && it->first < x + len) bad = true;
                    if (it != s->occupied[c].begin() && std::prev(it)->second > x) bad = true;
                }
                if (bad) continue;
                std::vector<uint8_t> v(len);
                for (uint64_t i = 0; i < len; i++) {
                    uint8_t m = enc(pl.site[i]);
                    std::vector<uint8_t> opts;
                    for (uint8_t b : kBases) if (m & b) opts.push_back(b);
                    v[i] = opts.empty() ? 4 : opts[r.below(opts.size())];
                }
                uint32_t subs = (uint32_t)r.below(pl.max_subs + 1);
                for (uint32_t t = 0; t < subs; t++) {
                    uint64_t i = pl.guide_off + r.below(pl.guide_len);
                    uint8_t old = v[i];
                    v[i] = kBases[(std::find(kBases, kBases + 4, old) - kBases + 1 + r.below(3)) % 4];
                }
                bool rev = r.below(2);
                for (uint64_t i = 0; i < len; i++) {
                    uint8_t b = rev ? v[len - 1 - i] : v[i];
                    if (rev) b = (uint8_t)(((b << 2) | (b >> 2)) & 0xF);
                    s->put(c, x + i, b);
                }
                s->occupied[c].insert({x, x + len});
                break;
            }
        }
    }
    return 0;
}

uint64_t cofsynth_packed_bytes(const cofsynth* s) { return s->packed.size(); }
const uint8_t* cofsynth_packed(const cofsynth* s) { return s->packed.data(); }
uint64_t cofsynth_n_chunks(const cofsynth* s) { return s->chunks.size(); }
const cof_chunk* cofsynth_chunks(const cofsynth* s) { return s->chunks.data(); }
uint32_t cofsynth_n_chroms(const cofsynth* s) { return (uint32_t)s->names.size(); }
const char* cofsynth_chrom_name(const cofsynth* s, uint32_t i) { return s->names[i].c_str(); }

// Writes the genome as FASTA with the configured line width; reading it back
// gives exactly the same chunks.
int cofsynth_write_fasta(const cofsynth* s, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf(1 << 22);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    std::string line;
    for (uint32_t c = 0; c < s->lens.size(); c++) {
        fprintf(f, ">%s\n", s->names[c].c_str());
        for (uint64_t p = 0; p < s->lens[c]; p += s->width) {
            uint64_t e = std::min<uint64_t>(p + s->width, s->lens[c]);
            line.clear();
            for (uint64_t i = p; i < e; i++) line.push_back(kDec[s->get(c, i)]);
            line.push_back('\n');
            fwrite(line.data(), 1, line.size(), f);
        }
    }
    fclose(f);
    return 0;
}

// Writes the genome as a version-0 .2bit file: runs of N become N blocks,
// A/C/G/T pack at 2 bits per base (T C A G = 0 1 2 3, first base in the top
// bits); other IUPAC letters pack as T, as in UCSC's converter, and so do
// not read back identically (use a genome without them).
int cofsynth_write_2bit(const cofsynth* s, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf(1 << 22);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    auto put32 = [&](uint32_t v) { fwrite(&v, 4, 1, f); };
    const uint32_t n = (uint32_t)s->lens.size();
    put32(0x1A412743u);
    put32(0);
    put32(n);
    put32(0);
    // N blocks first: the index needs every record's size
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> blocks(n);
    uint64_t off = 16;
    for (uint32_t c = 0; c < n; c++) off += 1 + s->names[c].size() + 4;
    std::vector<uint32_t> rec_off(n);
    for (uint32_t c = 0; c < n; c++) {
        for (uint64_t i = 0; i < s->lens[c];) {
            if (s->get(c, i) != 0xF) { i++; continue; }
            uint64_t j = i;
            while (j < s->lens[c] && s->get(c, j) == 0xF) j++;
            blocks[c].push_back({(uint32_t)i, (uint32_t)(j - i)});
            i = j;
        }
        rec_off[c] = (uint32_t)off;
        off += 4 + 4 + 8 * blocks[c].size() + 4 + 4 + (s->lens[c] + 3) / 4;
    }
    for (uint32_t c = 0; c < n; c++) {
        const uint8_t len = (uint8_t)s->names[c].size();
        fwrite(&len, 1, 1, f);
        fwrite(s->names[c].data(), 1, len, f);
        put32(rec_off[c]);
    }
    std::vector<uint8_t> dna;
    for (uint32_t c = 0; c < n; c++) {
        put32((uint32_t)s->lens[c]);
        put32((uint32_t)blocks[c].size());
        for (auto& b : blocks[c]) put32(b.first);
        for (auto& b : blocks[c]) put32(b.second);
        put32(0);   // mask blocks
        put32(0);   // reserved
        dna.assign((s->lens[c] + 3) / 4, 0);
        for (uint64_t i = 0; i < s->lens[c]; i++) {
            const uint8_t v = s->get(c, i);
            const uint8_t code = v == 2 ? 1 : v == 4 ? 2 : v == 8 ? 3 : 0;
            dna[i >> 2] |= (uint8_t)(code << (6 - 2 * (i & 3)));
        }
        fwrite(dna.data(), 1, dna.size(), f);
    }
    return fclose(f) == 0 ? 0 : -1;
}

void cofsynth_free(cofsynth* s) { delete s; }

}  // extern "C"
