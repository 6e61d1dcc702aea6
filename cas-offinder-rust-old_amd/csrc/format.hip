// Device-side TSV formatting of the final hits: the rows of the reference's
// writer (cas-offinder-cli/src/main.rs:130-247; plain branch :158-222 with the
// case rules of :198-217, bulge branch :145-157) are produced on the GPU from
// the hit records and the resident genome, then streamed to the file.
//
// Per batch of rows: one pass computes each row's length and a per-workgroup
// sum, one workgroup scans the sums, and a second pass recomputes the lengths,
// scans them inside the workgroup and writes every row at its offset.  The
// host copies a batch's bytes back and writes them while the device formats
// the next batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "cof_internal.hpp"
#include "engine.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

constexpr uint32_t kFmtThreads = 256;

struct FmtPat {
    uint32_t blob_off;     // pattern nibbles (pattern orientation) in the blob
    uint16_t L;
    uint8_t forward;
    uint8_t bulge;
    uint32_t shift_base;   // sat(pam_len - 1): reverse-strand position shift before the DNA-bulge term
    uint32_t pad;
};

struct FmtDec {
    char c[16];            // bit4 -> uppercase IUPAC letter, 0 -> NUL
};

__device__ __forceinline__ uint32_t comp4(uint32_t v) {
    return ((v & 1u) << 2) | ((v >> 2) & 1u) | ((v & 2u) << 2) | ((v >> 2) & 2u);
}

__device__ __forceinline__ bool single(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }

__device__ __forceinline__ uint32_t nib_at(const uint8_t* g, uint64_t i) {
    return (g[i >> 1] >> ((uint32_t)(i & 1) * 4)) & 0xFu;
}

__device__ __forceinline__ uint32_t ndigits(uint64_t v) {
    uint32_t n = 1;
    if (v >> 32) {
        while (v >= 10) { v /= 10; n++; }
        return n;
    }
    uint32_t w = (uint32_t)v;
    while (w >= 10) { w /= 10; n++; }
    return n;
}

__device__ __forceinline__ uint8_t* put_dec(uint8_t* o, uint64_t v, uint32_t nd) {
    if (v >> 32) {
        for (uint32_t k = nd; k-- > 0;) { o[k] = (uint8_t)('0' + v % 10); v /= 10; }
    } else {
        uint32_t w = (uint32_t)v;
        for (uint32_t k = nd; k-- > 0;) { o[k] = (uint8_t)('0' + w % 10); w /= 10; }
    }
    return o + nd;
}

// Everything of a row but its characters: length and the numbers it shows.
struct RowShape {
    uint32_t len, lr, type_len, nd_pos, nd_mm, nd_bs, bsize;
    uint64_t pos;
};

__device__ __forceinline__ RowShape row_shape(const cof_hit& h, const FmtPat& p, const uint32_t* name_len) {
    RowShape s;
    const uint32_t db = h.dna_bulge, rb = h.rna_bulge;
    s.lr = p.bulge ? h.n_ops : p.L;
    s.bsize = p.bulge ? db + rb : 0;
    s.type_len = (!p.bulge || s.bsize == 0) ? 1 : 3;
    const uint64_t shift = p.forward ? 0 : (p.shift_base > db ? p.shift_base - db : 0);
    s.pos = h.chrom_pos + shift;
    s.nd_pos = ndigits(s.pos);
    s.nd_mm = ndigits(h.mismatches);
    s.nd_bs = ndigits(s.bsize);
    s.len = s.type_len + 2 * s.lr + name_len[h.chrom] + s.nd_pos + s.nd_mm + s.nd_bs + 9;
    return s;
}

__global__ __launch_bounds__(kFmtThreads) void fmt_len_kernel(const cof_hit* __restrict__ hits, uint32_t n,
                                                              const FmtPat* __restrict__ pats,
                                                              const uint32_t* __restrict__ name_len,
                                                              unsigned long long* __restrict__ block_sum) {
    __shared__ unsigned long long part[kFmtThreads / 64];
    const uint32_t i = blockIdx.x * kFmtThreads + threadIdx.x;
    unsigned long long len = 0;
    if (i < n) {
        const cof_hit h = hits[i];
        len = row_shape(h, pats[h.pattern], name_len).len;
    }
    for (int o = 32; o > 0; o >>= 1) len += __shfl_down(len, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = len;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < kFmtThreads / 64; w++) t += part[w];
        block_sum[blockIdx.x] = t;
    }
}

// Exclusive prefix of the per-workgroup sums (one workgroup), total at [nb].
__global__ __launch_bounds__(1024) void fmt_scan_kernel(const unsigned long long* __restrict__ block_sum, uint32_t nb,
                                                        unsigned long long* __restrict__ block_off) {
    __shared__ unsigned long long buf[1024];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < nb ? block_sum[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t d = 1; d < 1024; d <<= 1) {
            const unsigned long long add = threadIdx.x >= d ? buf[threadIdx.x - d] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        const unsigned long long c = carry;
        if (i < nb) block_off[i] = c + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_off[nb] = carry;
}

__global__ __launch_bounds__(kFmtThreads) void fmt_write_kernel(
    const cof_hit* __restrict__ hits, const HitAux* __restrict__ aux, uint32_t n, const FmtPat* __restrict__ pats,
    const uint8_t* __restrict__ blob, const uint8_t* __restrict__ names, const uint64_t* __restrict__ name_off,
    const uint32_t* __restrict__ name_len, const unsigned long long* __restrict__ block_off,
    const uint8_t* __restrict__ genome, uint64_t g_nibs, FmtDec dec, uint8_t* __restrict__ out) {
    __shared__ uint32_t wsum[kFmtThreads / 64];
    const uint32_t i = blockIdx.x * kFmtThreads + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cof_hit h;
    FmtPat p;
    RowShape s;
    uint32_t len = 0;
    if (i < n) {
        h = hits[i];
        p = pats[h.pattern];
        s = row_shape(h, p, name_len);
        len = s.len;
    }
    // exclusive prefix of the lengths inside the workgroup
    uint32_t incl = len;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++) before += wsum[w];
    if (i >= n) return;
    uint8_t* o = out + block_off[blockIdx.x] + before + incl - len;

    // type
    if (s.type_len == 1) *o++ = 'X';
    else if (h.dna_bulge > 0) { o[0] = 'D'; o[1] = 'N'; o[2] = 'A'; o += 3; }
    else { o[0] = 'R'; o[1] = 'N'; o[2] = 'A'; o += 3; }
    *o++ = '\t';
    uint8_t* rna = o;
    uint8_t* dna = o + s.lr + 1;
    const uint8_t* pb = blob + p.blob_off;
    const uint64_t start = aux[i].nib;
    if (!p.bulge) {
        // plain: pattern window at `start`, N-masked, zero past the batch view
        const uint64_t view_end = min(aux[i].view_end, g_nibs);
        const uint32_t L = p.L;
        for (uint32_t j = 0; j < L; j++) {
            const uint32_t k = p.forward ? j : L - 1 - j;
            const uint64_t gi = start + k;
            uint32_t v = gi < view_end ? nib_at(genome, gi) : 0;
            if (v == 0xF) v = 0;
            const uint32_t pc = pb[k];
            const char r = dec.c[p.forward ? pc : comp4(pc)];
            char d = dec.c[p.forward ? v : comp4(v)];
            if (v == 0) d = pc == 0xF ? 'N' : 'n';
            else if (pc == 0xF) {}
            else if (!(single(v) && (v & pc))) d = (char)(d | 0x20);
            rna[j] = (uint8_t)r;
            dna[j] = (uint8_t)d;
        }
    } else {
        // bulge: aligned strings from the ops (alignment order), reversed and
        // complemented for reverse patterns
        // (reads stay inside the pattern and the genome even for an
        // inconsistent record)
        const uint32_t nops = h.n_ops;
        uint32_t pi = 0;
        uint64_t t = start;
        for (uint32_t k = 0; k < nops; k++) {
            const uint32_t op = (uint32_t)(h.ops[k >> 5] >> (2 * (k & 31))) & 3u;
            uint32_t rc = 0x10, dc = 0x10;      // 0x10: gap
            if (op != 2) { rc = pi < p.L ? pb[pi] : 0; pi++; }
            if (op != 3) { dc = t < g_nibs ? nib_at(genome, t) : 0; t++; }
            const uint32_t j = p.forward ? k : nops - 1 - k;
            rna[j] = rc == 0x10 ? (uint8_t)'-' : (uint8_t)dec.c[p.forward ? rc : comp4(rc)];
            dna[j] = dc == 0x10 ? (uint8_t)'-' : (uint8_t)dec.c[p.forward ? dc : comp4(dc)];
        }
    }
    rna[s.lr] = '\t';
    o = dna + s.lr;
    *o++ = '\t';
    const uint8_t* nm = names + name_off[h.chrom];
    for (uint32_t k = 0, e = name_len[h.chrom]; k < e; k++) *o++ = nm[k];
    *o++ = '\t';
    o = put_dec(o, s.pos, s.nd_pos);
    *o++ = '\t';
    *o++ = p.forward ? '+' : '-';
    *o++ = '\t';
    o = put_dec(o, h.mismatches, s.nd_mm);
    *o++ = '\t';
    o = put_dec(o, s.bsize, s.nd_bs);
    *o = '\n';
}

template <typename T>
void ensure_dev(T** p, size_t* cap, size_t n) {
    if (n <= *cap && *p) return;
    if (*p) HIPCHECK(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    const size_t want = std::max<size_t>(n, 1);
    HIPCHECK(hipMalloc((void**)p, want * sizeof(T)));
    *cap = want;
}

}  // namespace

// write_tsv_device waits for its last copy before returning, so no device
// work uses these buffers when the formatter goes away.
TsvFormatter::~TsvFormatter() {
    for (void* p : {(void*)d_hits, (void*)d_aux, (void*)d_bsum, (void*)d_boff, (void*)d_out, (void*)d_pats,
                    (void*)d_blob, (void*)d_names, (void*)d_name_off, (void*)d_name_len})
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; k++) {
        if (h_out[k]) (void)hipHostFree(h_out[k]);
        if (ev_out[k]) (void)hipEventDestroy(ev_out[k]);
    }
    if (h_total) (void)hipHostFree(h_total);
    if (ev_total) (void)hipEventDestroy(ev_total);
}

uint64_t write_tsv_device(const Genome& g, const Query& q, const HitVec& hits, const AuxVec& aux,
                          const uint32_t* d_genome, FILE* f, hipStream_t st, TsvFormatter& fm) {
    fputs("#Bulge type\tcrRNA\tDNA\tChromosome\tPosition\tDirection\tMismatches\tBulge Size\n", f);
    const uint64_t n = hits.size();
    if (n == 0) return 0;
    // per-query and per-genome tables
    std::vector<FmtPat> pats(q.pats.size());
    std::vector<uint8_t> blob;
    uint32_t max_lr = 1;
    for (size_t k = 0; k < q.pats.size(); k++) {
        const Pattern& p = q.pats[k];
        pats[k] = FmtPat{(uint32_t)blob.size(), (uint16_t)p.L, (uint8_t)p.forward, (uint8_t)p.bulge,
                         p.pam_len_out > 0 ? p.pam_len_out - 1 : 0, 0};
        blob.insert(blob.end(), p.b4.begin(), p.b4.end());
        max_lr = std::max(max_lr, p.bulge ? p.L + p.D + p.R : p.L);
    }
    std::vector<uint8_t> names;
    std::vector<uint64_t> name_off;
    std::vector<uint32_t> name_len;
    uint32_t max_name = 0;
    for (const auto& s : g.chrom_names) {
        name_off.push_back(names.size());
        name_len.push_back((uint32_t)s.size());
        names.insert(names.end(), s.begin(), s.end());
        max_name = std::max(max_name, (uint32_t)s.size());
    }
    FmtDec dec;
    memcpy(dec.c, kCodec.dec, 16);
    // row bound: type 3, two strings, name, 20-digit position, 5-digit counts, 7 tabs, dir, newline
    const uint64_t max_row = 3 + 2ull * max_lr + max_name + 20 + 5 + 5 + 9;
    constexpr uint64_t kBatchBytes = 64ull << 20;
    uint64_t B = std::min<uint64_t>(1u << 18, std::max<uint64_t>(kFmtThreads, kBatchBytes / max_row));
    B = B / kFmtThreads * kFmtThreads;
    const uint64_t out_cap = B * max_row;
    const uint32_t max_blocks = (uint32_t)(B / kFmtThreads);

    ensure_dev(&fm.d_hits, &fm.hits_cap, B);
    ensure_dev(&fm.d_aux, &fm.aux_cap, B);
    ensure_dev(&fm.d_bsum, &fm.bsum_cap, max_blocks);
    ensure_dev(&fm.d_boff, &fm.boff_cap, max_blocks + 1);
    ensure_dev(&fm.d_out, &fm.out_cap, out_cap);
    ensure_dev(&fm.d_pats, &fm.pats_cap, pats.size() * sizeof(FmtPat));
    ensure_dev(&fm.d_blob, &fm.blob_cap, blob.size());
    ensure_dev(&fm.d_names, &fm.names_cap, names.size());
    ensure_dev(&fm.d_name_off, &fm.name_off_cap, name_off.size());
    ensure_dev(&fm.d_name_len, &fm.name_len_cap, name_len.size());
    if (fm.h_cap < out_cap) {
        for (int k = 0; k < 2; k++) {
            if (fm.h_out[k]) HIPCHECK(hipHostFree(fm.h_out[k]));
            fm.h_out[k] = nullptr;
        }
        fm.h_cap = 0;
        for (int k = 0; k < 2; k++) HIPCHECK(hipHostMalloc((void**)&fm.h_out[k], out_cap, hipHostMallocDefault));
        fm.h_cap = out_cap;
    }
    if (!fm.h_total) HIPCHECK(hipHostMalloc((void**)&fm.h_total, 2 * sizeof(unsigned long long), hipHostMallocDefault));
    for (int k = 0; k < 2; k++)
        if (!fm.ev_out[k]) HIPCHECK(hipEventCreateWithFlags(&fm.ev_out[k], hipEventDisableTiming));
    if (!fm.ev_total) HIPCHECK(hipEventCreateWithFlags(&fm.ev_total, hipEventDisableTiming));

    HIPCHECK(hipMemcpyAsync(fm.d_pats, pats.data(), pats.size() * sizeof(FmtPat), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(fm.d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice, st));
    if (!names.empty()) HIPCHECK(hipMemcpyAsync(fm.d_names, names.data(), names.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(fm.d_name_off, name_off.data(), name_off.size() * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(fm.d_name_len, name_len.data(), name_len.size() * 4, hipMemcpyHostToDevice, st));

    const FmtPat* d_pats = (const FmtPat*)fm.d_pats;
    uint64_t prev_bytes = 0;
    bool io_ok = true;
    int k = 0;
    for (uint64_t r0 = 0; r0 < n; r0 += B, k ^= 1) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(B, n - r0);
        const uint32_t nb = (rows + kFmtThreads - 1) / kFmtThreads;
        // the row bound above holds for every record the device will read
        for (uint64_t r = r0; r < r0 + rows; r++) {
            const cof_hit& h = hits[r];
            if (h.pattern >= pats.size() || h.chrom >= name_len.size() ||
                (q.pats[h.pattern].bulge && h.n_ops > max_lr) || (uint32_t)h.dna_bulge + h.rna_bulge > 99999)
                throw Error(COF_E_INVALID, "tsv formatter: hit record out of range");
        }
        HIPCHECK(hipMemcpyAsync(fm.d_hits, hits.data() + r0, rows * sizeof(cof_hit), hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(fm.d_aux, aux.data() + r0, rows * sizeof(HitAux), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(fmt_len_kernel, dim3(nb), dim3(kFmtThreads), 0, st, fm.d_hits, rows, d_pats,
                           fm.d_name_len, fm.d_bsum);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(fmt_scan_kernel, dim3(1), dim3(1024), 0, st, fm.d_bsum, nb, fm.d_boff);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(fm.h_total + k, fm.d_boff + nb, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(fm.ev_total, st));
        hipLaunchKernelGGL(fmt_write_kernel, dim3(nb), dim3(kFmtThreads), 0, st, fm.d_hits, fm.d_aux, rows,
                           d_pats, fm.d_blob, fm.d_names, fm.d_name_off, fm.d_name_len, fm.d_boff,
                           (const uint8_t*)d_genome, 2 * g.n_bytes, dec, fm.d_out);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventSynchronize(fm.ev_total));
        const uint64_t bytes = fm.h_total[k];
        if (bytes > out_cap) throw Error(COF_E_HIP, "tsv formatter: batch exceeds its row bound");
        HIPCHECK(hipMemcpyAsync(fm.h_out[k], fm.d_out, bytes, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(fm.ev_out[k], st));
        if (r0 > 0) {
            // the previous batch's bytes are on the host: write them while
            // this batch is formatted and copied
            HIPCHECK(hipEventSynchronize(fm.ev_out[k ^ 1]));
            if (fwrite(fm.h_out[k ^ 1], 1, prev_bytes, f) != prev_bytes) io_ok = false;
        }
        prev_bytes = bytes;
    }
    HIPCHECK(hipEventSynchronize(fm.ev_out[k ^ 1]));
    if (fwrite(fm.h_out[k ^ 1], 1, prev_bytes, f) != prev_bytes) io_ok = false;
    if (!io_ok) throw Error(COF_E_IO, "short write of tsv rows");
    return n;
}

}  // namespace cof
