// Genome ingest for libcof: codec tables, the three reference readers and the
// virtual batch layout.  Host C++; the result is a tight-packed bit4 buffer
// that the search kernels keep resident in HBM.
#include <algorithm>
#include <cstdio>
#include <dirent.h>
#include <sys/stat.h>

#include "cof_internal.hpp"

namespace cof {

Codec::Codec() {
    memset(enc_mixed, 0, sizeof enc_mixed);
    memset(enc_plain, 0, sizeof enc_plain);
    memset(dec, 0, sizeof dec);
    const char* letters = "TCAGRYSWKMBDHVN";
    const uint8_t codes[] = {1, 2, 4, 8, 4 | 8, 2 | 1, 8 | 2, 4 | 1, 8 | 1, 4 | 2, 2 | 8 | 1, 4 | 8 | 1, 4 | 2 | 1, 4 | 2 | 8, 15};
    for (int i = 0; letters[i]; i++) {
        unsigned char up = (unsigned char)letters[i], lo = (unsigned char)(up | 0x20);
        enc_mixed[up] = enc_mixed[lo] = codes[i];
        if (i < 4) enc_plain[up] = enc_plain[lo] = codes[i];
        dec[codes[i]] = letters[i];
    }
}
const Codec kCodec;

char comp_char(char c) {                       // bit4ops.rs:207-221 semantics
    uint8_t v = kCodec.enc_mixed[(uint8_t)c];
    if (v == 0) return c;
    return (char)(kCodec.dec[comp4(v)] | (c & 0x20));
}
std::string revcomp(const std::string& s) {
    std::string r(s.rbegin(), s.rend());
    for (auto& c : r) c = comp_char(c);
    return r;
}

void Genome::index() {
    lut.assign((size_t)(n_bytes >> 15) + 2, 0);
    size_t c = 0;
    for (size_t k = 0; k < lut.size(); k++) {
        const uint64_t byte = (uint64_t)k << 15;
        while (c + 1 < chunks.size() && chunks[c + 1].byte_off <= byte) c++;
        lut[k] = (uint32_t)c;
    }
}

size_t Genome::chunk_of(uint64_t nibble) const {
    const uint64_t byte = nibble >> 1;
    size_t lo = 0, hi = chunks.size();
    const uint64_t k = byte >> 15;
    if (k + 1 < lut.size()) {                  // the chunks starting in this bucket
        lo = lut[k];
        hi = std::min<size_t>(chunks.size(), (size_t)lut[k + 1] + 1);
    }
    while (lo < hi) {                          // first chunk with byte_off > byte
        size_t mid = (lo + hi) / 2;
        if (chunks[mid].byte_off <= byte) lo = mid + 1; else hi = mid;
    }
    return lo == 0 ? 0 : lo - 1;
}

namespace {

// Appends chunks to a Genome as they are produced by a reader.
struct Builder {
    Genome& g;
    std::vector<uint8_t> cur;      // CHUNK bytes being filled
    std::string cur_name;
    uint64_t cur_start = 0, cur_end = 0;
    explicit Builder(Genome& gg) : g(gg), cur(kChunkNucl / 2, 0) {}
    uint32_t chrom_id(const std::string& name) {
        if (g.chrom_names.empty() || g.chrom_names.back() != name) g.chrom_names.push_back(name);
        return (uint32_t)g.chrom_names.size() - 1;
    }
    void emit() {
        uint64_t sz = cur_end - cur_start;
        uint64_t vb = (sz + 1) / 2;
        Chunk c;
        c.chrom = chrom_id(cur_name);
        c.start = cur_start;
        c.end = cur_end;
        c.byte_off = g.n_bytes;
        g.packed.insert(g.packed.end(), cur.begin(), cur.begin() + vb);
        g.n_bytes += vb;
        g.chunks.push_back(c);
    }
    void reset(const std::string& name, uint64_t start) {
        std::fill(cur.begin(), cur.end(), 0);
        cur_name = name;
        cur_start = cur_end = start;
    }
    void put_line(const char* s, size_t n) {
        uint64_t off = cur_end - cur_start;
        for (size_t k = 0; k < n; k++, off++) {
            uint8_t v = kCodec.enc_mixed[(uint8_t)s[k]];
            cur[off >> 1] |= (off & 1) ? (uint8_t)(v << 4) : v;
        }
        cur_end += n;
    }
};

// FASTA: chunk flush when the next line would overflow 64 Knt, so chunk
// boundaries fall on line boundaries (read_fasta.rs:45-64).
void read_fasta(const std::string& path, Genome& g) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Error(COF_E_IO, "cannot open " + path);
    Builder b(g);
    bool started = false;
    std::string line;
    std::vector<char> buf(1 << 22);
    auto handle = [&](std::string& ln) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        if (!ln.empty() && ln[0] == '>') {
            if (b.cur_end != b.cur_start) b.emit();
            b.reset(ln.substr(1), 0);
            started = true;
            return;
        }
        if (!started) throw Error(COF_E_IO, "fasta file needs to start with >");
        if (b.cur_name.empty()) throw Error(COF_E_IO, "> must be followed by chromosome name");
        if (ln.size() + (b.cur_end - b.cur_start) > kChunkNucl) {
            uint64_t s = b.cur_end;
            std::string nm = b.cur_name;
            b.emit();
            b.reset(nm, s);
        }
        if (ln.size() > kChunkNucl) throw Error(COF_E_IO, "line in fasta too long");
        b.put_line(ln.data(), ln.size());
    };
    size_t n;
    bool pending = false;
    while ((n = fread(buf.data(), 1, buf.size(), f)) > 0) {
        size_t s = 0;
        const char* p = buf.data();
        while (true) {
            const char* nl = (const char*)memchr(p + s, '\n', n - s);
            if (!nl) break;
            size_t i = (size_t)(nl - p);
            line.append(p + s, i - s);
            handle(line);
            line.clear();
            pending = false;
            s = i + 1;
        }
        if (s < n) { line.append(p + s, n - s); pending = true; }
    }
    if (pending) handle(line);
    fclose(f);
    if (b.cur_end > b.cur_start) b.emit();
}

uint32_t u32le(FILE* f) {
    uint32_t v;
    if (fread(&v, 4, 1, f) != 1) throw Error(COF_E_IO, "2bit: truncated file");
    return v;
}

// 2bit: 64 Knt pieces per chromosome; N blocks become 0xF, soft-mask blocks
// are ignored (read_2bit.rs:108-193).
void read_2bit(const std::string& path, Genome& g) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Error(COF_E_IO, "cannot open " + path);
    if (u32le(f) != 0x1A412743u) throw Error(COF_E_IO, ".2bit file badly formatted header");
    if (u32le(f) != 0) throw Error(COF_E_IO, "only supports version 0 of .2bit format");
    uint32_t n = u32le(f);
    fseek(f, 4, SEEK_CUR);
    std::vector<std::string> names(n);
    for (auto& nm : names) {
        int len = fgetc(f);
        if (len == EOF) throw Error(COF_E_IO, "2bit: truncated file");
        nm.resize((size_t)len);
        if (len && fread(&nm[0], 1, (size_t)len, f) != (size_t)len) throw Error(COF_E_IO, "2bit: truncated file");
        fseek(f, 4, SEEK_CUR);
    }
    Builder b(g);
    std::vector<uint8_t> raw(kChunkNucl / 4);
    for (auto& nm : names) {
        uint64_t len = u32le(f);
        uint32_t nb = u32le(f);
        std::vector<uint32_t> st(nb), sz(nb);
        for (auto& v : st) v = u32le(f);
        for (auto& v : sz) v = u32le(f);
        std::vector<std::pair<uint32_t, uint32_t>> blocks(nb);
        for (uint32_t i = 0; i < nb; i++) blocks[i] = {st[i], sz[i]};
        std::stable_sort(blocks.begin(), blocks.end(), [](auto& a, auto& c) { return a.first < c.first; });
        uint32_t nm_blocks = u32le(f);
        fseek(f, (long)nm_blocks * 8 + 4, SEEK_CUR);
        uint64_t pos = 0;
        int64_t bi = 0;
        while (pos < len) {
            uint64_t piece = std::min<uint64_t>(len - pos, kChunkNucl);
            size_t nbytes = (size_t)((piece + 3) / 4);
            if (fread(raw.data(), 1, nbytes, f) != nbytes) throw Error(COF_E_IO, "2bit: truncated file");
            b.reset(nm, pos);
            for (uint64_t i = 0; i < piece; i++) {
                uint8_t code = (raw[i >> 2] >> (6 - 2 * (i & 3))) & 3;     // T C A G
                uint8_t v = (uint8_t)(1u << code);
                b.cur[i >> 1] |= (i & 1) ? (uint8_t)(v << 4) : v;
            }
            // N blocks; the scan restarts one block back so a block spanning
            // pieces is applied to each (read_2bit.rs:162-182).
            bi = std::max<int64_t>(bi - 1, 0);
            while (bi < (int64_t)blocks.size()) {
                int64_t bs = (int64_t)blocks[bi].first - (int64_t)pos;
                int64_t be = (int64_t)(uint32_t)(blocks[bi].first + blocks[bi].second) - (int64_t)pos;
                if (bs > (int64_t)piece) break;
                uint64_t a = (uint64_t)std::max<int64_t>(0, bs);
                uint64_t e = std::min<uint64_t>((uint64_t)std::max<int64_t>(0, be), piece);
                for (uint64_t i = a; i < e; i++) b.cur[i >> 1] |= (i & 1) ? 0xF0 : 0x0F;
                bi++;
            }
            b.cur_end = pos + piece;
            b.emit();
            pos += piece;
        }
    }
    fclose(f);
}

}  // namespace

void read_genome_file(const std::string& path, Genome& g) {
    g = Genome();
    struct stat st;
    if (stat(path.c_str(), &st) != 0) throw Error(COF_E_IO, "cannot stat " + path);
    if (S_ISDIR(st.st_mode)) {
        DIR* d = opendir(path.c_str());
        if (!d) throw Error(COF_E_IO, "cannot open directory " + path);
        while (struct dirent* e = readdir(d)) {
            std::string nm = e->d_name, full = path + "/" + nm;
            struct stat s2;
            if (stat(full.c_str(), &s2) != 0 || !S_ISREG(s2.st_mode)) continue;
            if (nm.size() >= 3 && nm.compare(nm.size() - 3, 3, ".fa") == 0) read_fasta(full, g);
        }
        closedir(d);
    } else {
        FILE* f = fopen(path.c_str(), "rb");
        if (!f) throw Error(COF_E_IO, "cannot open " + path);
        int c = fgetc(f);
        fclose(f);
        if (c == '>') read_fasta(path, g); else read_2bit(path, g);
    }
    // chunk order invariant of chunks_to_searchchunk (search.rs:1396-1399)
    for (size_t i = 1; i < g.chunks.size(); i++) {
        const Chunk &p = g.chunks[i - 1], &c = g.chunks[i];
        if (!(p.end == c.start || c.start == 0)) throw Error(COF_E_IO, "chromosome chunks out of order");
    }
    g.packed.resize(g.n_bytes + 256, 0);
}

std::vector<Batch> make_batches(size_t n_chunks, uint64_t batch_nucl) {
    std::vector<Batch> out;
    size_t cps = (size_t)(batch_nucl / kChunkNucl);
    if (n_chunks == 0) return out;
    // batch 0: [0, cps); batch k: [k(cps-1), k(cps-1)+cps) with a head chunk;
    // the remainder forms a last batch unless it is only the carried head.
    size_t first = 0;
    bool head = false;
    while (true) {
        size_t last = std::min(first + cps, n_chunks);
        if (last - first == cps) {
            out.push_back({first, last, head});
            first = last - 1;
            head = true;
            if (first + 1 == n_chunks) break;   // only the carried head is left
            continue;
        }
        if (!(head && last - first == 1)) out.push_back({first, last, head});
        break;
    }
    return out;
}

}  // namespace cof
