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

}  // namespace

bool read_fasta_device(const std::string& path, Genome& g, uint32_t* d_genome_alloc(uint64_t n_bytes, void* user),
                       void* user, hipStream_t st) {
    g = Genome();
    struct stat sb;
    if (stat(path.c_str(), &sb) != 0) throw Error(COF_E_IO, "cannot stat " + path);
    const uint64_t fsize = (uint64_t)sb.st_size;
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Error(COF_E_IO, "cannot open " + path);
    // The file streams through two pinned blocks: block k is indexed on the
    // host while block k-1 is copied to the device.
    constexpr uint64_t kBlock = 64ull << 20;
    uint8_t* stage[2] = {nullptr, nullptr};
    uint8_t* d_text = nullptr;
    hipEvent_t copied[2] = {nullptr, nullptr};
    struct Guard {
        FILE* f;
        uint8_t** stage;
        uint8_t** d_text;
        hipEvent_t* ev;
        hipStream_t st;
        ~Guard() {
            (void)hipStreamSynchronize(st);
            for (int k = 0; k < 2; k++) {
                if (stage[k]) (void)hipHostFree(stage[k]);
                if (ev[k]) (void)hipEventDestroy(ev[k]);
            }
            if (*d_text) (void)hipFree(*d_text);
            if (f) fclose(f);
        }
    } guard{f, stage, &d_text, copied, st};
    for (int k = 0; k < 2; k++) {
        HIPCHECK(hipHostMalloc((void**)&stage[k], kBlock, hipHostMallocDefault));
        HIPCHECK(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
    }
    HIPCHECK(hipMalloc(&d_text, std::max<uint64_t>(fsize, 1)));

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
        const uint8_t* b = stage[k];
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
        off += n;
    }
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
        hipLaunchKernelGGL(fasta_encode_kernel, dim3((uint32_t)ic.size()), dim3(256), 0, st, d_text, d_ic, d_codes,
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

}  // namespace cof
