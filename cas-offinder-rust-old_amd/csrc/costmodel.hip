// Learned cost model for the bulge path's work order (no reference
// counterpart: the reference walks its batches in genome order,
// search.rs:1525-1572).
//
// Per genome chunk (<= 64 Knt, the unit of the packed layout) a feature kernel
// computes 12 FP32 features; an MLP 12 -> 64 -> 64 -> 1 (ReLU, 5,057
// parameters) predicts log(1 + traceback work) of the chunk.  The bulge
// scheduler uses the prediction only to order its filter tiles (heavy first)
// and to cut rounds of equal predicted work; the hit set never depends on it.
//
// Features (f[i], all FP32), over the chunk's n nucleotides:
//   0 GC fraction of the A/C/G/T bases          6 PAM hit fraction (mean over the query's PAMs)
//   1 N fraction                                7 positions whose 8-mer is a guide 8-mer (/ n)
//   2 8-mer Shannon entropy / 16 bits           8 positions whose 8-mer is a PAM-proximal guide 8-mer (/ n)
//   3 fraction of 8-mers seen >= 4 times        9 log2(1 + count of 7) / 16
//   4 CpG dinucleotides / n                    10 IUPAC (non-ACGTN) fraction
//   5 homopolymer 6-mer starts / n             11 n / 65536
//
// MLP kernels: one wave per workgroup, lane k owns hidden unit k of both
// layers (its W1 row and W2 row in registers, W2 also in LDS for the
// transposed product of the backward pass).  The backward kernel writes one
// partial gradient per workgroup; a reduction kernel sums them in a fixed
// order, so training is deterministic.  Adam runs on the device; one training
// step (gradient, reduction, Adam, step counter) is captured once into a
// hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "cof_internal.hpp"
#include "costmodel.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

constexpr int kFeatThreads = 1024;           // 64 positions per thread
constexpr int kKmerWords = 32768;            // 65536 16-bit 8-mer counters, two per word
constexpr int kMaxPams = 16;
constexpr int kMaxPamCells = 8;

struct FeatChunk {
    uint64_t nib_begin;
    uint32_t size;
    uint32_t pad;
};

struct FeatPam {
    uint32_t n_cells;
    uint32_t pos[kMaxPamCells];    // relative to the first cell
    uint32_t code[kMaxPamCells];
};

struct FeatQuery {
    uint32_t n_pams;
    FeatPam pams[kMaxPams];
};

__device__ __forceinline__ uint32_t gnib(const uint32_t* __restrict__ g, uint64_t i) {
    return (g[i >> 3] >> (4u * (uint32_t)(i & 7u))) & 0xFu;
}

// 2-bit code of a single base (T=1 C=2 A=4 G=8 -> 0..3), 4 otherwise
__device__ __forceinline__ uint32_t code2(uint32_t v) {
    switch (v) {
        case 1: return 0;
        case 2: return 1;
        case 4: return 2;
        case 8: return 3;
        default: return 4;
    }
}

__device__ __forceinline__ float block_sum(float v, float* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kFeatThreads / 64; w++) s += red[w];
    return s;   // valid in thread 0
}

__global__ __launch_bounds__(kFeatThreads) void chunk_features_kernel(
    const uint32_t* __restrict__ genome, const FeatChunk* __restrict__ chunks, const FeatQuery* __restrict__ fq,
    const uint32_t* __restrict__ seed_any, const uint32_t* __restrict__ seed_pam, float* __restrict__ out) {
    __shared__ uint32_t cnt[kKmerWords];
    __shared__ uint32_t s_any[2048];
    __shared__ uint32_t s_pam[2048];
    __shared__ float red[kFeatThreads / 64];

    const FeatChunk C = chunks[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < (uint32_t)kKmerWords; i += kFeatThreads) cnt[i] = 0;
    for (uint32_t i = tid; i < 2048u; i += kFeatThreads) {
        s_any[i] = seed_any[i];
        s_pam[i] = seed_pam[i];
    }
    __syncthreads();

    const uint32_t n = C.size;
    const uint32_t p0 = tid * 64u;
    float n_acgt = 0, n_gc = 0, n_n = 0, n_iupac = 0, n_cpg = 0, n_homo = 0, n_kv = 0, n_any = 0, n_pamseed = 0,
          n_pam = 0;
    // per-position counts over this thread's 64 starts
    for (uint32_t p = p0; p < p0 + 64u && p < n; p++) {
        const uint32_t v = gnib(genome, C.nib_begin + p);
        const uint32_t c = code2(v);
        n_n += v == 0xFu;
        n_iupac += (c == 4u && v != 0xFu && v != 0u);
        if (c < 4u) {
            n_acgt += 1;
            n_gc += (c == 1u || c == 3u);
        }
        if (p + 1 < n) n_cpg += (v == 2u && gnib(genome, C.nib_begin + p + 1) == 8u);
        if (p + 6 <= n && c < 4u) {
            bool h = true;
            for (uint32_t k = 1; k < 6 && h; k++) h = gnib(genome, C.nib_begin + p + k) == v;
            n_homo += h;
        }
        for (uint32_t a = 0; a < fq->n_pams; a++) {
            const FeatPam& P = fq->pams[a];
            if (p + P.pos[P.n_cells - 1] >= n) continue;
            bool ok = true;
            for (uint32_t k = 0; k < P.n_cells && ok; k++) {
                const uint32_t g = gnib(genome, C.nib_begin + p + P.pos[k]);
                ok = g != 0u && g != 0xFu && (g & P.code[k]) != 0u;
            }
            n_pam += ok;
        }
    }
    // 8-mers starting in this thread's range: rolling 2-bit code
    {
        uint32_t kmer = 0, run = 0;
        const uint32_t q_end = min(p0 + 64u + 7u, n);
        for (uint32_t q = p0; q < q_end; q++) {
            const uint32_t c = code2(gnib(genome, C.nib_begin + q));
            if (c < 4u) {
                kmer = ((kmer << 2) | c) & 0xFFFFu;
                run++;
            } else {
                run = 0;
            }
            if (run >= 8 && q >= p0 + 7u) {
                n_kv += 1;
                atomicAdd(&cnt[kmer >> 1], 1u << (16u * (kmer & 1u)));
                n_any += (s_any[kmer >> 5] >> (kmer & 31u)) & 1u;
                n_pamseed += (s_pam[kmer >> 5] >> (kmer & 31u)) & 1u;
            }
        }
    }
    __syncthreads();
    // repeats: 8-mers whose chunk count is >= 4
    float n_rep = 0;
    {
        uint32_t kmer = 0, run = 0;
        const uint32_t q_end = min(p0 + 64u + 7u, n);
        for (uint32_t q = p0; q < q_end; q++) {
            const uint32_t c = code2(gnib(genome, C.nib_begin + q));
            if (c < 4u) {
                kmer = ((kmer << 2) | c) & 0xFFFFu;
                run++;
            } else {
                run = 0;
            }
            if (run >= 8 && q >= p0 + 7u) n_rep += ((cnt[kmer >> 1] >> (16u * (kmer & 1u))) & 0xFFFFu) >= 4u;
        }
    }
    const float total_kv = block_sum(n_kv, red);
    __syncthreads();
    __shared__ float s_total;
    if (tid == 0) s_total = total_kv;
    __syncthreads();
    // entropy of the 8-mer distribution
    float h = 0;
    if (s_total > 0) {
        const float inv = 1.0f / s_total;
        for (uint32_t b = tid * 32u; b < tid * 32u + 32u; b++) {
            const uint32_t w = cnt[b];
            const float c0 = (float)(w & 0xFFFFu), c1 = (float)(w >> 16);
            if (c0 > 0) h -= c0 * inv * log2f(c0 * inv);
            if (c1 > 0) h -= c1 * inv * log2f(c1 * inv);
        }
    }
    float v[11] = {n_acgt, n_gc, n_n, n_iupac, n_cpg, n_homo, n_any, n_pamseed, n_pam, n_rep, h};
    for (int i = 0; i < 11; i++) v[i] = block_sum(v[i], red);
    if (tid == 0) {
        const float fn = (float)max(n, 1u);
        float* f = out + (size_t)blockIdx.x * kMlpIn;
        f[0] = v[0] > 0 ? v[1] / v[0] : 0.f;
        f[1] = v[2] / fn;
        f[2] = v[10] / 16.f;
        f[3] = s_total > 0 ? v[9] / s_total : 0.f;
        f[4] = v[4] / fn;
        f[5] = v[5] / fn;
        f[6] = fq->n_pams ? v[8] / (fn * (float)fq->n_pams) : 0.f;
        f[7] = v[6] / fn;
        f[8] = v[7] / fn;
        f[9] = log2f(1.f + v[6]) / 16.f;
        f[10] = v[3] / fn;
        f[11] = (float)n / 65536.f;
    }
}

// ---------------------------------------------------------------------------
// MLP 12 -> 64 -> 64 -> 1, parameter layout (floats):
//   W1[64][12] @0, b1[64] @768, W2[64][64] @832, b2[64] @4928, w3[64] @4992, b3 @5056
// ---------------------------------------------------------------------------
constexpr int kW1 = 0, kB1 = 768, kW2 = 832, kB2 = 4928, kW3 = 4992, kB3 = 5056;
static_assert(kB3 + 1 == kMlpParams, "parameter layout");

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One wave per workgroup, rows blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                         uint32_t n, float* __restrict__ Y) {
    __shared__ float h1[kMlpHidden];
    __shared__ float x[kMlpIn];
    const uint32_t k = threadIdx.x;
    float w1[kMlpIn], w2[kMlpHidden];
    for (int i = 0; i < kMlpIn; i++) w1[i] = P[kW1 + k * kMlpIn + i];
    for (int j = 0; j < kMlpHidden; j++) w2[j] = P[kW2 + k * kMlpHidden + j];
    const float b1 = P[kB1 + k], b2 = P[kB2 + k], w3 = P[kW3 + k], b3 = P[kB3];
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        if (k < (uint32_t)kMlpIn) x[k] = X[(size_t)r * kMlpIn + k];
        __syncthreads();
        float z1 = b1;
        for (int i = 0; i < kMlpIn; i++) z1 += w1[i] * x[i];
        h1[k] = fmaxf(z1, 0.f);
        __syncthreads();
        float z2 = b2;
        for (int j = 0; j < kMlpHidden; j++) z2 += w2[j] * h1[j];
        const float y = wave_sum(w3 * fmaxf(z2, 0.f)) + b3;
        if (k == 0) Y[r] = y;
        __syncthreads();
    }
}

// Loss sum((y - t)^2) / n and its gradient; partial sums per workgroup into
// part[blockIdx.x][kMlpParams + 1] (last slot: loss).
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                      const float* __restrict__ T, uint32_t n,
                                                      float* __restrict__ part) {
    __shared__ float W2s[kMlpHidden * kMlpHidden];
    __shared__ float h1[kMlpHidden];
    __shared__ float dz2s[kMlpHidden];
    __shared__ float x[kMlpIn];
    const uint32_t k = threadIdx.x;
    float w1[kMlpIn], w2[kMlpHidden];
    for (int i = 0; i < kMlpIn; i++) w1[i] = P[kW1 + k * kMlpIn + i];
    for (int j = 0; j < kMlpHidden; j++) {
        w2[j] = P[kW2 + k * kMlpHidden + j];
        W2s[k * kMlpHidden + j] = w2[j];
    }
    const float b1 = P[kB1 + k], b2 = P[kB2 + k], w3 = P[kW3 + k], b3 = P[kB3];
    float gw1[kMlpIn], gw2[kMlpHidden];
    for (int i = 0; i < kMlpIn; i++) gw1[i] = 0;
    for (int j = 0; j < kMlpHidden; j++) gw2[j] = 0;
    float gb1 = 0, gb2 = 0, gw3 = 0, gb3 = 0, loss = 0;
    const float inv_n = 1.0f / (float)n;
    __syncthreads();
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        if (k < (uint32_t)kMlpIn) x[k] = X[(size_t)r * kMlpIn + k];
        __syncthreads();
        float z1 = b1;
        for (int i = 0; i < kMlpIn; i++) z1 += w1[i] * x[i];
        const float a1 = fmaxf(z1, 0.f);
        h1[k] = a1;
        __syncthreads();
        float z2 = b2;
        for (int j = 0; j < kMlpHidden; j++) z2 += w2[j] * h1[j];
        const float a2 = fmaxf(z2, 0.f);
        const float y = wave_sum(w3 * a2) + b3;
        const float e = y - T[r];
        loss += e * e * inv_n;
        const float dy = 2.f * e * inv_n;
        gw3 += dy * a2;
        gb3 += dy;
        const float dz2 = z2 > 0.f ? dy * w3 : 0.f;
        gb2 += dz2;
        for (int j = 0; j < kMlpHidden; j++) gw2[j] += dz2 * h1[j];
        dz2s[k] = dz2;
        __syncthreads();
        float dh1 = 0;
        for (int m = 0; m < kMlpHidden; m++) dh1 += W2s[m * kMlpHidden + k] * dz2s[m];
        const float dz1 = z1 > 0.f ? dh1 : 0.f;
        gb1 += dz1;
        for (int i = 0; i < kMlpIn; i++) gw1[i] += dz1 * x[i];
        __syncthreads();
    }
    float* o = part + (size_t)blockIdx.x * (kMlpParams + 1);
    for (int i = 0; i < kMlpIn; i++) o[kW1 + k * kMlpIn + i] = gw1[i];
    for (int j = 0; j < kMlpHidden; j++) o[kW2 + k * kMlpHidden + j] = gw2[j];
    o[kB1 + k] = gb1;
    o[kB2 + k] = gb2;
    o[kW3 + k] = gw3;
    if (k == 0) {
        o[kB3] = gb3;
        o[kMlpParams] = loss;
    }
}

// grad[i] = sum over workgroups b (in order) of part[b][i], i <= kMlpParams
__global__ void mlp_reduce_kernel(const float* __restrict__ part, uint32_t n_parts, float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > (uint32_t)kMlpParams) return;
    float s = 0;
    for (uint32_t b = 0; b < n_parts; b++) s += part[(size_t)b * (kMlpParams + 1) + i];
    grad[i] = s;
}

// Adam (Kingma & Ba 2015) with bias correction; *step counts completed steps.
__global__ void mlp_adam_kernel(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ grad, const uint32_t* __restrict__ step, float lr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kMlpParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float t = (float)(*step + 1);
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, t));
    const float vh = vi / (1.f - powf(b2, t));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void mlp_step_kernel(uint32_t* step) { *step += 1; }

__global__ void zero_f32_kernel(float* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

constexpr uint32_t kGradBlocks = 256;

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct FeaturePlan {
    FeatQuery* d_q = nullptr;
    uint32_t* d_any = nullptr;
    uint32_t* d_pam = nullptr;
};

FeaturePlan* feature_plan_create(const Query& q) {
    FeatQuery fq{};
    std::vector<uint32_t> any(2048, 0), pam(2048, 0);
    std::map<std::vector<uint32_t>, int> seen;
    bool use_bulge = q.any_bulge;
    for (const Pattern& p : q.pats) {
        if (use_bulge && !p.bulge) continue;
        // PAM cells: non-N positions of the effective filter, relative to the first
        std::vector<uint32_t> key;
        uint32_t k0 = 0;
        bool first = true;
        for (uint32_t k = 0; k < p.L; k++) {
            if (is_n_char(p.eff_filter[k])) continue;
            if (first) { k0 = k; first = false; }
            if (key.size() / 2 >= (size_t)kMaxPamCells) break;
            key.push_back(k - k0);
            key.push_back(kCodec.enc_mixed[(uint8_t)p.eff_filter[k]]);
        }
        if (!key.empty() && !seen.count(key) && fq.n_pams < (uint32_t)kMaxPams) {
            seen[key] = 1;
            FeatPam& P = fq.pams[fq.n_pams++];
            P.n_cells = (uint32_t)key.size() / 2;
            for (uint32_t c = 0; c < P.n_cells; c++) {
                P.pos[c] = key[2 * c];
                P.code[c] = key[2 * c + 1];
            }
        }
        // guide 8-mers (single bases only), and the 8-mer next to the PAM's N block
        auto code = [&](uint32_t s, uint32_t& out) {
            out = 0;
            for (uint32_t k = s; k < s + 8; k++) {
                uint8_t v = p.b4[k];
                uint32_t c = v == 1 ? 0 : v == 2 ? 1 : v == 4 ? 2 : v == 8 ? 3 : 4;
                if (c == 4) return false;
                out = (out << 2) | c;
            }
            return true;
        };
        uint32_t km;
        for (uint32_t s = 0; s + 8 <= p.L; s++)
            if (code(s, km)) any[km >> 5] |= 1u << (km & 31);
        uint32_t e = p.L;
        while (e > 0 && p.b4[e - 1] == 0xF) e--;
        uint32_t b = 0;
        while (b < p.L && p.b4[b] == 0xF) b++;
        if (e < p.L && e >= 8) {
            if (code(e - 8, km)) pam[km >> 5] |= 1u << (km & 31);
        } else if (b > 0 && b + 8 <= p.L) {
            if (code(b, km)) pam[km >> 5] |= 1u << (km & 31);
        }
    }
    auto* fp = new FeaturePlan();
    HIPCHECK(hipMalloc(&fp->d_q, sizeof(FeatQuery)));
    HIPCHECK(hipMemcpy(fp->d_q, &fq, sizeof fq, hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&fp->d_any, 2048 * 4));
    HIPCHECK(hipMemcpy(fp->d_any, any.data(), 2048 * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&fp->d_pam, 2048 * 4));
    HIPCHECK(hipMemcpy(fp->d_pam, pam.data(), 2048 * 4, hipMemcpyHostToDevice));
    return fp;
}

void feature_plan_destroy(FeaturePlan* fp) {
    if (!fp) return;
    (void)hipFree(fp->d_q);
    (void)hipFree(fp->d_any);
    (void)hipFree(fp->d_pam);
    delete fp;
}

void chunk_features(const FeaturePlan* fp, const Genome& g, const uint32_t* d_genome, hipStream_t st,
                    DevBuf& chunks_buf, float* d_out) {
    const size_t nc = g.chunks.size();
    if (!nc) return;
    std::vector<FeatChunk> fc(nc);
    for (size_t c = 0; c < nc; c++) fc[c] = FeatChunk{g.chunks[c].nib_begin(), (uint32_t)g.chunks[c].size(), 0u};
    chunks_buf.ensure(nc * sizeof(FeatChunk));
    HIPCHECK(hipMemcpyAsync(chunks_buf.p, fc.data(), nc * sizeof(FeatChunk), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(chunk_features_kernel, dim3((uint32_t)nc), dim3(kFeatThreads), 0, st, d_genome,
                       (const FeatChunk*)chunks_buf.p, fp->d_q, fp->d_any, fp->d_pam, d_out);
    HIPCHECK(hipGetLastError());
}

void DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    HIPCHECK(hipMalloc(&p, bytes));
    cap = bytes;
}

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

Mlp::Mlp(int dev, uint64_t seed) : device(dev) {
    HIPCHECK(hipSetDevice(dev));
    HIPCHECK(hipStreamCreate(&stream));
    HIPCHECK(hipMalloc(&d_params, kMlpParams * sizeof(float)));
    HIPCHECK(hipMalloc(&d_m, kMlpParams * sizeof(float)));
    HIPCHECK(hipMalloc(&d_v, kMlpParams * sizeof(float)));
    HIPCHECK(hipMalloc(&d_grad, (kMlpParams + 1) * sizeof(float)));
    HIPCHECK(hipMalloc(&d_step, sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&d_part, (size_t)kGradBlocks * (kMlpParams + 1) * sizeof(float)));
    // He-uniform weights from a seeded splitmix64 stream, zero biases
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t s = seed;
    auto uni = [&](float lim) { return (float)(((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * lim; };
    for (int i = 0; i < kMlpHidden * kMlpIn; i++) p[kW1 + i] = uni(std::sqrt(6.f / kMlpIn));
    for (int i = 0; i < kMlpHidden * kMlpHidden; i++) p[kW2 + i] = uni(std::sqrt(6.f / kMlpHidden));
    for (int i = 0; i < kMlpHidden; i++) p[kW3 + i] = uni(std::sqrt(6.f / kMlpHidden));
    set_params(p.data());
}

Mlp::~Mlp() {
    (void)hipSetDevice(device);
    if (train_exec) (void)hipGraphExecDestroy(train_exec);
    for (void* q : {(void*)d_params, (void*)d_m, (void*)d_v, (void*)d_grad, (void*)d_step, (void*)d_part})
        if (q) (void)hipFree(q);
    if (stream) (void)hipStreamDestroy(stream);
}

void Mlp::set_params(const float* p) {
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipMemcpy(d_params, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    reset_optimizer();
}

void Mlp::get_params(float* p) const {
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipMemcpy(p, d_params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
}

void Mlp::reset_optimizer() {
    hipLaunchKernelGGL(zero_f32_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, d_m, (uint32_t)kMlpParams);
    hipLaunchKernelGGL(zero_f32_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, d_v, (uint32_t)kMlpParams);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(1), 0, stream, (float*)d_step, 1u);
    HIPCHECK(hipStreamSynchronize(stream));
}

void mlp_forward_dev(const float* d_params, const float* d_x, uint32_t n, float* d_y, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(std::min<uint32_t>(n, 1024)), dim3(64), 0, st, d_params, d_x, n, d_y);
    HIPCHECK(hipGetLastError());
}

void Mlp::forward(const float* x, uint64_t n, float* y) {
    HIPCHECK(hipSetDevice(device));
    if (!n) return;
    DevBuf dx, dy;
    dx.ensure(n * kMlpIn * sizeof(float));
    dy.ensure(n * sizeof(float));
    HIPCHECK(hipMemcpyAsync(dx.p, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, stream));
    mlp_forward_dev(d_params, (const float*)dx.p, (uint32_t)n, (float*)dy.p, stream);
    HIPCHECK(hipMemcpyAsync(y, dy.p, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
}

void Mlp::launch_grad(const float* d_x, const float* d_t, uint32_t n) {
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(kGradBlocks), dim3(64), 0, stream, d_params, d_x, d_t, n, d_part);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, stream, d_part,
                       kGradBlocks, d_grad);
}

void Mlp::loss_grad(const float* x, const float* t, uint64_t n, float* loss, float* grad) {
    HIPCHECK(hipSetDevice(device));
    if (!n) throw Error(COF_E_INVALID, "no rows");
    DevBuf dx, dt;
    dx.ensure(n * kMlpIn * sizeof(float));
    dt.ensure(n * sizeof(float));
    HIPCHECK(hipMemcpyAsync(dx.p, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dt.p, t, n * sizeof(float), hipMemcpyHostToDevice, stream));
    launch_grad((const float*)dx.p, (const float*)dt.p, (uint32_t)n);
    HIPCHECK(hipGetLastError());
    std::vector<float> g(kMlpParams + 1);
    HIPCHECK(hipMemcpyAsync(g.data(), d_grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (grad) std::copy(g.begin(), g.begin() + kMlpParams, grad);
    if (loss) *loss = g[kMlpParams];
}

void Mlp::train(const float* x, const float* t, uint64_t n, uint32_t steps, float lr, bool eager, float* loss) {
    HIPCHECK(hipSetDevice(device));
    if (!n) throw Error(COF_E_INVALID, "no rows");
    DevBuf dx, dt;
    dx.ensure(n * kMlpIn * sizeof(float));
    dt.ensure(n * sizeof(float));
    HIPCHECK(hipMemcpyAsync(dx.p, x, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dt.p, t, n * sizeof(float), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    auto step = [&]() {
        launch_grad((const float*)dx.p, (const float*)dt.p, (uint32_t)n);
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, d_params, d_m, d_v,
                           d_grad, d_step, lr);
        hipLaunchKernelGGL(mlp_step_kernel, dim3(1), dim3(1), 0, stream, d_step);
    };
    if (train_exec) {
        (void)hipGraphExecDestroy(train_exec);
        train_exec = nullptr;
    }
    if (eager) {
        for (uint32_t s = 0; s < steps; s++) step();
        HIPCHECK(hipGetLastError());
    } else {
        // one full-batch Adam step captured once, replayed `steps` times
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        step();
        HIPCHECK(hipStreamEndCapture(stream, &graph));
        HIPCHECK(hipGraphInstantiate(&train_exec, graph, nullptr, nullptr, 0));
        HIPCHECK(hipGraphDestroy(graph));
        for (uint32_t s = 0; s < steps; s++) HIPCHECK(hipGraphLaunch(train_exec, stream));
    }
    // loss after training
    launch_grad((const float*)dx.p, (const float*)dt.p, (uint32_t)n);
    float l = 0;
    HIPCHECK(hipMemcpyAsync(&l, d_grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (train_exec) HIPCHECK(hipGraphExecDestroy(train_exec));
    train_exec = nullptr;
    if (loss) *loss = l;
}

}  // namespace cof
