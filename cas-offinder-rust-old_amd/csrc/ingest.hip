// Device-side FASTA ingest: the text streams through pinned host blocks to
// the device while the host indexes its lines and forms the reference's chunks (read_fasta.rs:9-71:
// a chunk is flushed when the next line would overflow 64 Knt, so chunks end
// on line boundaries), and the bit4 encode + tight nibble packing of every
// chunk (search.rs:1376-1427 layout, bit4ops.rs:123-153 codes) runs on the
// GPU straight into the resident genome buffer.
//
// The encode kernel maps a chunk's nucleotide n to its file byte with one
// division: within a chunk all lines but the last have the same width and the
// same distance between line starts.  A file whose chunks do not have that
// shape (varying line widths) is left to the host reader.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "cof_internal.hpp"
#include "engine.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

struct IngestChunk {
    uint64_t file_off;     // first nucleotide's byte in the file
    uint64_t byte_off;     // packed output offset
    uint32_t size;         // nucleotides
    uint32_t width;        // nucleotides per line (all lines but the last)
    uint32_t stride;       // bytes between line starts
    uint32_t pad;
};

__global__ __launch_bounds__(256) void fasta_encode_kernel(const uint8_t* __restrict__ text,
                                                           const IngestChunk* __restrict__ chunks,
                                                           const uint8_t* __restrict__ codes,
                                                           uint8_t* __restrict__ out) {
    __shared__ uint8_t lut[256];
    lut[threadIdx.x] = codes[threadIdx.x];
    __syncthreads();
    const IngestChunk C = chunks[blockIdx.x];
    const uint32_t n_bytes = (C.size + 1) / 2;
    for (uint32_t k = threadIdx.x; k < n_bytes; k += blockDim.x) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
            const uint32_t n = 2 * k + h;
            if (n < C.size) {
                const uint32_t line = n / C.width, col = n - line * C.width;
                v |= (uint32_t)lut[text[C.file_off + (uint64_t)line * C.stride + col]] << (4 * h);
            }
        }
        out[C.byte_off + k] = (uint8_t)v;
    }
}

// One 64 Knt piece of a 2bit chromosome (read_2bit.rs:77-114): its packed
// 2-bit DNA in the file, its place in the resident genome, and its N spans.
struct TwoBitPiece {
    uint64_t src;          // first packed-DNA byte of the piece in the file
    uint64_t byte_off;     // packed bit4 output offset
    uint32_t size;         // nucleotides
    uint32_t n_first;      // N spans [n_first, n_first + n_count)
    uint32_t n_count;
    uint32_t pad;
};

struct NSpan {
    uint32_t a, e;         // nucleotides [a, e) of the piece become N (0xF)
};

// One workgroup per piece: decode 4 nucleotides per file byte (2 bits each,
// most significant first, T C A G -> 1 2 4 8) into a 32 KB LDS image of the
// packed chunk, OR the N spans in, write the chunk's ceil(size/2) bytes.
__global__ __launch_bounds__(256) void twobit_decode_kernel(const uint8_t* __restrict__ text,
                                                            const TwoBitPiece* __restrict__ pieces,
                                                            const NSpan* __restrict__ spans,
                                                            uint8_t* __restrict__ out) {
    __shared__ uint32_t w[kChunkNucl / 8];
    const TwoBitPiece P = pieces[blockIdx.x];
    const uint32_t n_words = (P.size + 7) / 8;
    for (uint32_t k = threadIdx.x; k < n_words; k += blockDim.x) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
            const uint32_t rb = 2 * k + h;
            if (4 * rb < P.size) {
                const uint32_t raw = text[P.src + rb];
#pragma unroll
                for (uint32_t q = 0; q < 4; q++)
                    if (4 * rb + q < P.size) v |= (1u << ((raw >> (6 - 2 * q)) & 3u)) << (4 * (4 * h + q));
            }
        }
        w[k] = v;
    }
    __syncthreads();
    for (uint32_t sidx = P.n_first; sidx < P.n_first + P.n_count; sidx++) {
        const NSpan sp = spans[sidx];
        for (uint32_t k = sp.a / 8 + threadIdx.x; 8 * k < sp.e; k += blockDim.x) {
            const uint32_t lo = sp.a > 8 * k ? sp.a - 8 * k : 0;
            const uint32_t hi = min(sp.e - 8 * k, 8u);
            const uint32_t upto = hi == 8 ? 0xFFFFFFFFu : (1u << (4 * hi)) - 1;
            atomicOr(&w[k], upto & ~((1u << (4 * lo)) - 1));
        }
    }
    __syncthreads();
    const uint32_t n_bytes = (P.size + 1) / 2;
    const uint8_t* wb = (const uint8_t*)w;
    for (uint32_t b = threadIdx.x; b < n_bytes; b += blockDim.x) out[P.byte_off + b] = wb[b];
}

// The whole file streams to the device through two pinned blocks: block k
// is handed to `on_block` on the host while block k-1 is still being copied.
// The device copy stays valid until the guard is destroyed.
struct FileOnDevice {
    FILE* f = nullptr;
    uint64_t fsize = 0;
    uint8_t* stage[2] = {nullptr, nullptr};
    uint8_t* d_text = nullptr;
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    FileOnDevice(const std::string& path, hipStream_t stream) : st(stream) {
        struct stat sb;
        if (stat(path.c_str(), &sb) != 0) throw Error(COF_E_IO, "cannot stat " + path);
        fsize = (uint64_t)sb.st_size;
        f = fopen(path.c_str(), "rb");
        if (!f) throw Error(COF_E_IO, "cannot open " + path);
    }
    ~FileOnDevice() {
        (void)hipStreamSynchronize(st);
        for (int k = 0; k < 2; k++) {
            if (stage[k]) (void)hipHostFree(stage[k]);
            if (copied[k]) (void)hipEventDestroy(copied[k]);
        }
        if (d_text) (void)hipFree(d_text);
        if (f) fclose(f);
    }
    template <typename Fn>
    void stream(const std::string& path, Fn on_block) {
        constexpr uint64_t kBlock = 64ull << 20;
        for (int k = 0; k < 2; k++) {
            HIPCHECK(hipHostMalloc((void**)&stage[k], kBlock, hipHostMallocDefault));
            HIPCHECK(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
        }
        HIPCHECK(hipMalloc(&d_text, std::max<uint64_t>(fsize, 1)));
        if (fseek(f, 0, SEEK_SET) != 0) throw Error(COF_E_IO, "cannot read " + path);
        uint64_t off = 0;
        for (int k = 0; off < fsize; k ^= 1) {
            const uint64_t n = std::min(kBlock, fsize - off);
            HIPCHECK(hipEventSynchronize(copied[k]));   // the block's previous copy has left
            for (uint64_t got = 0; got < n;) {
                size_t r = fread(stage[k] + got, 1, (size_t)(n - got), f);
                if (r == 0) throw Error(COF_E_IO, "cannot read " + path);
                got += r;
            }
            HIPCHECK(hipMemcpyAsync(d_text + off, stage[k], n, hipMemcpyHostToDevice, st));
            HIPCHECK(hipEventRecord(copied[k], st));
            on_block(off, stage[k], n);
            off += n;
        }
    }
};

}  // namespace

bool read_fasta_device(const std::string& path, Genome& g, uint32_t* d_genome_alloc(uint64_t n_bytes, void* user),
                       void* user, hipStream_t st) {
    g = Genome();
    FileOnDevice fd(path, st);

    // line index -> chunks, with the host reader's rules and errors
    std::vector<IngestChunk> ic;
    bool regular = true;
    bool started = false;
    std::string name;
    uint64_t cur_start = 0, cur_end = 0;     // nucleotides of the open chunk (within the chromosome)
    uint64_t c_file = 0, c_last_line = 0;    // open chunk: first line start, last line start
    uint32_t c_width = 0, c_stride = 0, c_lines = 0, c_last_w = 0;
    auto chrom_id = [&](const std::string& nm) {
        if (g.chrom_names.empty() || g.chrom_names.back() != nm) g.chrom_names.push_back(nm);
        return (uint32_t)g.chrom_names.size() - 1;
    };
    auto emit = [&]() {
        const uint64_t sz = cur_end - cur_start;
        if (sz == 0) return;
        Chunk c;
        c.chrom = chrom_id(name);
        c.start = cur_start;
        c.end = cur_end;
        c.byte_off = g.n_bytes;
        g.chunks.push_back(c);
        ic.push_back(IngestChunk{c_file, g.n_bytes, (uint32_t)sz, std::max(c_width, 1u), c_stride ? c_stride : 1u, 0u});
        g.n_bytes += (sz + 1) / 2;
    };
    auto reset = [&](uint64_t start) {
        cur_start = cur_end = start;
        c_lines = 0;
        c_width = c_stride = c_last_w = 0;
    };
    // one line: file offset pos, its bytes s[0, n) without the '\n'
    auto handle = [&](uint64_t pos, const uint8_t* s, uint64_t n) {
        uint64_t len = n;
        if (len > 0 && s[len - 1] == '\r') len--;
        if (len > 0 && s[0] == '>') {
            emit();
            name.assign((const char*)s + 1, (size_t)(len - 1));
            reset(0);
            started = true;
            return;
        }
        if (!started) throw Error(COF_E_IO, "fasta file needs to start with >");
        if (name.empty()) throw Error(COF_E_IO, "> must be followed by chromosome name");
        if (len + (cur_end - cur_start) > kChunkNucl) {
            emit();
            reset(cur_end);
        }
        if (len > kChunkNucl) throw Error(COF_E_IO, "line in fasta too long");
        if (len == 0) return;
        if (c_lines == 0) {
            c_file = pos;
            c_width = (uint32_t)len;
        } else {
            // every line but the last must have the first line's width, and
            // line starts must be evenly spaced
            if (c_last_w != c_width) regular = false;
            const uint64_t stride = pos - c_last_line;
            if (c_lines == 1) c_stride = (uint32_t)stride;
            else if (stride != c_stride) regular = false;
            if (len > c_width) regular = false;
        }
        c_last_line = pos;
        c_last_w = (uint32_t)len;
        c_lines++;
        cur_end += len;
    };

    std::vector<uint8_t> carry;              // a line cut by a block boundary
    uint64_t carry_pos = 0;
    fd.stream(path, [&](uint64_t off, const uint8_t* b, uint64_t n) {
        uint64_t p = 0;
        if (!carry.empty() || carry_pos == off) {
            // finish the carried line with this block's first line
            const uint8_t* nl = (const uint8_t*)memchr(b, '\n', (size_t)n);
            const uint64_t e = nl ? (uint64_t)(nl - b) : n;
            carry.insert(carry.end(), b, b + e);
            if (nl) {
                handle(carry_pos, carry.data(), carry.size());
                carry.clear();
                p = e + 1;
                carry_pos = off + p;
            } else {
                p = n;
            }
        }
        while (p < n) {
            const uint8_t* nl = (const uint8_t*)memchr(b + p, '\n', (size_t)(n - p));
            if (!nl) {
                carry.assign(b + p, b + n);
                carry_pos = off + p;
                break;
            }
            const uint64_t e = (uint64_t)(nl - b);
            handle(off + p, b + p, e - p);
            p = e + 1;
            carry_pos = off + p;
        }
    });
    if (!carry.empty()) handle(carry_pos, carry.data(), carry.size());
    emit();
    for (size_t i = 1; i < g.chunks.size(); i++) {
        const Chunk &p = g.chunks[i - 1], &c = g.chunks[i];
        if (!(p.end == c.start || c.start == 0)) throw Error(COF_E_IO, "chromosome chunks out of order");
    }
    if (!regular) {
        g = Genome();
        return false;
    }
    // device encode straight into the resident genome
    uint32_t* d_out = d_genome_alloc(g.n_bytes, user);
    if (!ic.empty()) {
        IngestChunk* d_ic = nullptr;
        uint8_t* d_codes = nullptr;
        HIPCHECK(hipMalloc(&d_ic, ic.size() * sizeof(IngestChunk)));
        HIPCHECK(hipMalloc(&d_codes, 256));
        struct DFree {
            void* a;
            void* b;
            ~DFree() {
                if (a) (void)hipFree(a);
                if (b) (void)hipFree(b);
            }
        } dguard{d_ic, d_codes};
        HIPCHECK(hipMemcpyAsync(d_ic, ic.data(), ic.size() * sizeof(IngestChunk), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(d_codes, kCodec.enc_mixed, 256, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(fasta_encode_kernel, dim3((uint32_t)ic.size()), dim3(256), 0, st, fd.d_text, d_ic, d_codes,
                           (uint8_t*)d_out);
        HIPCHECK(hipGetLastError());
        // host copy of the packed genome (row strings, seam re-evaluation)
        g.packed.resize(g.n_bytes + 256, 0);
        if (g.n_bytes) HIPCHECK(hipMemcpyAsync(g.packed.data(), d_out, g.n_bytes, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    } else {
        g.packed.resize(256, 0);
    }
    return true;
}

bool read_2bit_device(const std::string& path, Genome& g, uint32_t* d_genome_alloc(uint64_t n_bytes, void* user),
                      void* user, hipStream_t st) {
    g = Genome();
    FileOnDevice fd(path, st);
    FILE* f = fd.f;
    auto u32 = [&]() {
        uint32_t v;
        if (fread(&v, 4, 1, f) != 1) throw Error(COF_E_IO, "2bit: truncated file");
        return v;
    };
    auto skip = [&](uint64_t n) {
        if (fseek(f, (long)n, SEEK_CUR) != 0) throw Error(COF_E_IO, "2bit: truncated file");
    };
    // header and index (read_2bit.rs:33-57); records follow in index order
    if (u32() != 0x1A412743u) throw Error(COF_E_IO, ".2bit file badly formatted header");
    if (u32() != 0) throw Error(COF_E_IO, "only supports version 0 of .2bit format");
    const uint32_t n = u32();
    skip(4);
    std::vector<std::string> names(n);
    for (auto& nm : names) {
        int len = fgetc(f);
        if (len == EOF) throw Error(COF_E_IO, "2bit: truncated file");
        nm.resize((size_t)len);
        if (len && fread(&nm[0], 1, (size_t)len, f) != (size_t)len) throw Error(COF_E_IO, "2bit: truncated file");
        skip(4);
    }
    // per record: N blocks and where its packed DNA lies; pieces of 64 Knt
    // with the reader's N-block walk (one block back at each piece,
    // read_2bit.rs:99-112), evaluated here so every piece gets its spans
    std::vector<TwoBitPiece> pieces;
    std::vector<NSpan> spans;
    for (auto& nm : names) {
        const uint64_t len = u32();
        const uint32_t nb = u32();
        std::vector<uint32_t> bst(nb), bsz(nb);
        for (auto& v : bst) v = u32();
        for (auto& v : bsz) v = u32();
        std::vector<std::pair<uint32_t, uint32_t>> blocks(nb);
        for (uint32_t i = 0; i < nb; i++) blocks[i] = {bst[i], bsz[i]};
        std::stable_sort(blocks.begin(), blocks.end(), [](auto& a, auto& c) { return a.first < c.first; });
        const uint32_t nm_blocks = u32();
        skip((uint64_t)nm_blocks * 8 + 4);
        const uint64_t dna_off = (uint64_t)ftell(f);
        const uint64_t dna_bytes = (len + 3) / 4;
        if (dna_off + dna_bytes > fd.fsize) throw Error(COF_E_IO, "2bit: truncated file");
        int64_t bi = 0;
        uint32_t chrom = 0;
        for (uint64_t pos = 0; pos < len; pos += kChunkNucl) {
            const uint64_t piece = std::min<uint64_t>(len - pos, kChunkNucl);
            TwoBitPiece P{dna_off + pos / 4, g.n_bytes, (uint32_t)piece, (uint32_t)spans.size(), 0, 0};
            bi = std::max<int64_t>(bi - 1, 0);
            while (bi < (int64_t)blocks.size()) {
                const int64_t bs = (int64_t)blocks[bi].first - (int64_t)pos;
                const int64_t be = (int64_t)(uint32_t)(blocks[bi].first + blocks[bi].second) - (int64_t)pos;
                if (bs > (int64_t)piece) break;
                const uint64_t a = (uint64_t)std::max<int64_t>(0, bs);
                const uint64_t e = std::min<uint64_t>((uint64_t)std::max<int64_t>(0, be), piece);
                if (e > a) spans.push_back(NSpan{(uint32_t)a, (uint32_t)e});
                bi++;
            }
            P.n_count = (uint32_t)spans.size() - P.n_first;
            if (pos == 0) {
                if (g.chrom_names.empty() || g.chrom_names.back() != nm) g.chrom_names.push_back(nm);
                chrom = (uint32_t)g.chrom_names.size() - 1;
            }
            Chunk c;
            c.chrom = chrom;
            c.start = pos;
            c.end = pos + piece;
            c.byte_off = g.n_bytes;
            g.chunks.push_back(c);
            pieces.push_back(P);
            g.n_bytes += (piece + 1) / 2;
        }
        skip(dna_bytes);
    }
    // the file to the device, then one decode launch into the resident genome
    fd.stream(path, [](uint64_t, const uint8_t*, uint64_t) {});
    uint32_t* d_out = d_genome_alloc(g.n_bytes, user);
    if (!pieces.empty()) {
        TwoBitPiece* d_pieces = nullptr;
        NSpan* d_spans = nullptr;
        HIPCHECK(hipMalloc(&d_pieces, pieces.size() * sizeof(TwoBitPiece)));
        HIPCHECK(hipMalloc(&d_spans, std::max<size_t>(spans.size(), 1) * sizeof(NSpan)));
        struct DFree {
            void* a;
            void* b;
            ~DFree() {
                if (a) (void)hipFree(a);
                if (b) (void)hipFree(b);
            }
        } dguard{d_pieces, d_spans};
        HIPCHECK(hipMemcpyAsync(d_pieces, pieces.data(), pieces.size() * sizeof(TwoBitPiece), hipMemcpyHostToDevice, st));
        if (!spans.empty())
            HIPCHECK(hipMemcpyAsync(d_spans, spans.data(), spans.size() * sizeof(NSpan), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(twobit_decode_kernel, dim3((uint32_t)pieces.size()), dim3(256), 0, st, fd.d_text, d_pieces,
                           d_spans, (uint8_t*)d_out);
        HIPCHECK(hipGetLastError());
        g.packed.resize(g.n_bytes + 256, 0);
        if (g.n_bytes) HIPCHECK(hipMemcpyAsync(g.packed.data(), d_out, g.n_bytes, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    } else {
        g.packed.resize(256, 0);
    }
    return true;
}

}  // namespace cof
