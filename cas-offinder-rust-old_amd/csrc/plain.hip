// Plain mismatch scan on gfx950 (bulge = 0 path).
//
// Semantics: the hit set of the reference's plain path, i.e. the kernel
// predicate of search_chunk_cpu (cas-offinder-lib/src/search.rs:378-463:
// kmm = L - sum popc(pattern & N-masked genome) <= M, signed) together with
// the writer's post-hoc checks (cas-offinder-cli/src/main.rs:158-194: every
// non-N filter position holds a plain base of the filter set, and the
// recounted mismatches <= M).  Rather than emitting every kmm hit and
// filtering on the host as the reference does, the device evaluates the whole
// predicate, so only final hits leave the GPU.
//
// Design (one workgroup = one tile of T consecutive start positions):
//   1. the tile's genome words (+ halo) are loaded coalesced into LDS with the
//      N-mask (0xF -> 0, search.rs:512-531) applied on the fly; a SWAR test
//      flags tiles holding IUPAC genome letters (R, Y, ...);
//   2. per pattern group (patterns sharing one effective filter), every lane
//      tests the PAM of one start position with a 16-bit nibble lookup per
//      filter position, and passing starts are compacted into an LDS list with
//      a wave64 ballot + mbcnt prefix;
//   3. each lane takes one compacted start, extracts its compare window with
//      v_alignbit funnel shifts, and loops over the group's patterns (uniform
//      scalar loads) doing AND + v_bcnt per 32-bit word.
// Without IUPAC letters in the window, genome single bases make the writer's
// recount equal kmm minus the genome-N cells under pattern N, so the single
// popcount threshold is exact; IUPAC tiles take an exact per-nibble path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>

#include "cof_internal.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

constexpr int kTile = 16384;                 // start positions per workgroup
constexpr int kThreads = 256;
constexpr int kMaxL = 64;                    // plain pattern length supported
constexpr int kHaloWords = (kMaxL + 8) / 8 + 2;
constexpr int kTileWords = kTile / 8 + kHaloWords;
constexpr int kMaxPam = 64;

struct GroupDev {
    uint32_t pat_begin, pat_count;
    uint32_t L, k_lo, n_pam, W;
    uint32_t pam_pos[kMaxPam];
    uint32_t pam_mask[kMaxPam];
};

// 4-bit nibble k (0..7) of word w
__device__ __forceinline__ uint32_t nib_of(uint32_t w, uint32_t k) { return (w >> (4 * k)) & 0xFu; }

template <int W>
__global__ __launch_bounds__(kThreads) void plain_scan_kernel(
    const uint32_t* __restrict__ genome, uint64_t scan_end, uint64_t first_tile_nib,
    const GroupDev* __restrict__ groups, uint32_t n_groups,
    const uint32_t* __restrict__ cmp_words,   // [pattern][8]
    const int32_t* __restrict__ thr,          // [pattern]
    const uint32_t* __restrict__ full_words,  // [pattern][8] full pattern nibbles (IUPAC path)
    const uint32_t* __restrict__ max_mm,      // [pattern]
    const uint32_t* __restrict__ pat_id,      // [pattern] -> query pattern index
    PlainRaw* __restrict__ out, uint32_t* __restrict__ out_count, uint32_t out_cap) {
    __shared__ uint32_t gw[kTileWords];
    __shared__ uint16_t cand[kTile];
    __shared__ uint32_t n_cand;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint64_t s0 = first_tile_nib + (uint64_t)blockIdx.x * kTile;
    const uint64_t w0 = s0 >> 3;

    // 1. genome tile -> LDS (N-masked), IUPAC detection
    int iupac = 0;
    for (uint32_t i = tid; i < (uint32_t)kTileWords; i += kThreads) {
        uint32_t w = genome[w0 + i];
        uint32_t f = w & (w >> 1) & (w >> 2) & (w >> 3) & 0x11111111u;   // nibble == 0xF
        uint32_t x = w & (w >> 1) & 0x55555555u;
        uint32_t o = (w | (w >> 1)) & 0x55555555u;
        uint32_t multi = (x | (x >> 2) | (o & (o >> 2))) & 0x11111111u;  // >= 2 bits set
        iupac |= (multi & ~f) != 0;
        gw[i] = w & ~(f * 0xFu);
    }
    iupac = __syncthreads_or(iupac);

    for (uint32_t gi = 0; gi < n_groups; gi++) {
        const GroupDev& G = groups[gi];
        if (tid == 0) n_cand = 0;
        __syncthreads();
        // 2. PAM test + compaction
        for (uint32_t base = 0; base < (uint32_t)kTile; base += kThreads) {
            uint32_t i = base + tid;
            bool ok = s0 + i < scan_end;
            for (uint32_t k = 0; k < G.n_pam && ok; k++) {
                uint32_t q = i + G.pam_pos[k];
                uint32_t v = nib_of(gw[q >> 3], q & 7u);
                ok = (G.pam_mask[k] >> v) & 1u;
            }
            uint64_t bal = __ballot(ok);
            uint32_t cnt = (uint32_t)__popcll(bal);
            uint32_t wbase = 0;
            if (lane == 0 && cnt) wbase = atomicAdd(&n_cand, cnt);
            wbase = __shfl(wbase, 0);
            if (ok) {
                uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                cand[wbase + rank] = (uint16_t)i;
            }
        }
        __syncthreads();
        const uint32_t nc = n_cand;
        // 3. compare windows against the group's patterns
        for (uint32_t c0 = wave * 64; c0 < nc; c0 += kThreads) {
            uint32_t c = c0 + lane;
            bool active = c < nc;
            uint32_t i = active ? cand[c] : 0;
            if (!iupac) {
                uint32_t q = i + G.k_lo;
                uint32_t wi = q >> 3, sh = (q & 7u) * 4u;
                uint32_t win[W];
#pragma unroll
                for (int w = 0; w < W; w++) win[w] = __builtin_amdgcn_alignbit(gw[wi + w + 1], gw[wi + w], sh);
                for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++) {
                    const uint32_t* cw = cmp_words + (size_t)p * 8;
                    uint32_t sum = 0;
#pragma unroll
                    for (int w = 0; w < W; w++) sum += __popc(win[w] & cw[w]);
                    if (active && (int32_t)sum >= thr[p]) {
                        uint32_t idx = atomicAdd(out_count, 1u);
                        if (idx < out_cap) out[idx] = PlainRaw{s0 + i, pat_id[p], 0u};
                    }
                }
            } else {
                for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++) {
                    const uint32_t* fw = full_words + (size_t)p * 8;
                    int32_t pop = 0;
                    uint32_t amm = 0;
                    for (uint32_t k = 0; k < G.L; k++) {
                        uint32_t q = i + k;
                        uint32_t g = nib_of(gw[q >> 3], q & 7u);
                        uint32_t pb = nib_of(fw[k >> 3], k & 7u);
                        pop += __popc(pb & g);
                        bool single = g != 0 && (g & (g - 1)) == 0;
                        amm += (pb != 0xFu && !(single && (g & pb))) ? 1u : 0u;
                    }
                    uint32_t M = max_mm[p];
                    if (active && (int32_t)G.L - pop <= (int32_t)M && amm <= M) {
                        uint32_t idx = atomicAdd(out_count, 1u);
                        if (idx < out_cap) out[idx] = PlainRaw{s0 + i, pat_id[p], 1u};
                    }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

struct PlainPlan {
    struct Launch {
        int W;
        uint32_t group_begin, n_groups;
    };
    std::vector<GroupDev> groups;
    std::vector<Launch> launches;
    GroupDev* d_groups = nullptr;
    uint32_t *d_cmp = nullptr, *d_full = nullptr, *d_mm = nullptr, *d_pid = nullptr;
    int32_t* d_thr = nullptr;
    PlainRaw* d_out = nullptr;
    uint32_t* d_count = nullptr;
    uint32_t out_cap = 0;
    uint32_t max_L = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

PlainPlan* plain_plan_create(const Query& q) {
    auto* pp = new PlainPlan();
    // group plain patterns by (L, effective filter)
    std::map<std::pair<uint32_t, std::string>, std::vector<uint32_t>> by_key;
    for (uint32_t i = 0; i < q.pats.size(); i++) {
        const Pattern& p = q.pats[i];
        if (p.bulge) continue;
        if (p.L > (uint32_t)kMaxL) throw Error(COF_E_UNSUPPORTED, "plain search supports patterns up to 64 nt");
        std::string f = p.eff_filter;
        for (auto& c : f) c = (char)toupper((unsigned char)c);
        by_key[{p.L, f}].push_back(i);
    }
    std::vector<uint32_t> cmp, full, mm, pid;
    std::vector<int32_t> thr;
    std::map<int, std::vector<GroupDev>> by_w;
    uint32_t pat_index = 0;
    for (auto& kv : by_key) {
        uint32_t L = kv.first.first;
        const std::string& f = kv.first.second;
        pp->max_L = std::max(pp->max_L, L);
        GroupDev G{};
        G.L = L;
        std::vector<bool> excl(L, false);
        for (uint32_t k = 0; k < L; k++) {
            if (is_n_char(f[k])) continue;
            if (G.n_pam >= (uint32_t)kMaxPam) throw Error(COF_E_UNSUPPORTED, "too many PAM positions");
            uint32_t mask = 0;
            uint8_t fv = kCodec.enc_mixed[(uint8_t)f[k]];
            for (uint32_t v = 0; v < 16; v++)
                if (single_base((uint8_t)v) && (v & fv)) mask |= 1u << v;
            G.pam_pos[G.n_pam] = k;
            G.pam_mask[G.n_pam] = mask;
            G.n_pam++;
            bool all_n = true;
            for (uint32_t pi : kv.second) all_n &= q.pats[pi].b4[k] == 0xF;
            excl[k] = all_n;
        }
        uint32_t k_lo = L, k_hi = 0, n_excl = 0;
        for (uint32_t k = 0; k < L; k++) {
            if (excl[k]) { n_excl++; continue; }
            k_lo = std::min(k_lo, k);
            k_hi = std::max(k_hi, k + 1);
        }
        if (k_lo >= k_hi) { k_lo = 0; k_hi = 1; }
        G.k_lo = k_lo;
        G.W = (k_hi - k_lo + 7) / 8;
        G.pat_begin = pat_index;
        G.pat_count = (uint32_t)kv.second.size();
        for (uint32_t pi : kv.second) {
            const Pattern& p = q.pats[pi];
            uint32_t cw[8] = {0}, fw[8] = {0};
            for (uint32_t k = k_lo; k < k_hi; k++) {
                if (excl[k]) continue;
                uint32_t r = k - k_lo;
                cw[r / 8] |= (uint32_t)p.b4[k] << (4 * (r % 8));
            }
            for (uint32_t k = 0; k < L; k++) fw[k / 8] |= (uint32_t)p.b4[k] << (4 * (k % 8));
            cmp.insert(cmp.end(), cw, cw + 8);
            full.insert(full.end(), fw, fw + 8);
            thr.push_back((int32_t)L - (int32_t)n_excl - (int32_t)p.M);
            mm.push_back(p.M);
            pid.push_back(pi);
            pat_index++;
        }
        by_w[(int)G.W].push_back(G);
    }
    for (auto& kv : by_w) {
        PlainPlan::Launch l{kv.first, (uint32_t)pp->groups.size(), (uint32_t)kv.second.size()};
        pp->groups.insert(pp->groups.end(), kv.second.begin(), kv.second.end());
        pp->launches.push_back(l);
    }
    if (pat_index) {
        HIPCHECK(hipMalloc(&pp->d_groups, pp->groups.size() * sizeof(GroupDev)));
        HIPCHECK(hipMemcpy(pp->d_groups, pp->groups.data(), pp->groups.size() * sizeof(GroupDev), hipMemcpyHostToDevice));
        auto up = [](auto*& dptr, const auto& v) {
            HIPCHECK(hipMalloc(&dptr, v.size() * sizeof(v[0])));
            HIPCHECK(hipMemcpy(dptr, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
        };
        up(pp->d_cmp, cmp);
        up(pp->d_full, full);
        up(pp->d_thr, thr);
        up(pp->d_mm, mm);
        up(pp->d_pid, pid);
        pp->out_cap = 1u << 20;
        HIPCHECK(hipMalloc(&pp->d_out, (size_t)pp->out_cap * sizeof(PlainRaw)));
        HIPCHECK(hipMalloc(&pp->d_count, sizeof(uint32_t)));
        HIPCHECK(hipEventCreate(&pp->ev0));
        HIPCHECK(hipEventCreate(&pp->ev1));
    }
    return pp;
}

void plain_plan_destroy(PlainPlan* pp) {
    if (!pp) return;
    (void)hipFree(pp->d_groups);
    (void)hipFree(pp->d_cmp);
    (void)hipFree(pp->d_full);
    (void)hipFree(pp->d_thr);
    (void)hipFree(pp->d_mm);
    (void)hipFree(pp->d_pid);
    (void)hipFree(pp->d_out);
    (void)hipFree(pp->d_count);
    if (pp->ev0) (void)hipEventDestroy(pp->ev0);
    if (pp->ev1) (void)hipEventDestroy(pp->ev1);
    delete pp;
}

bool plain_plan_empty(const PlainPlan* pp) { return !pp || pp->launches.empty(); }

uint64_t plain_tile_nucl() { return kTile; }
uint64_t plain_pad_words() { return kTileWords + 16; }

// Scans start positions [0, scan_end) of the resident genome.  Device words
// past the genome must be zero up to plain_pad_words() beyond the last tile.
void plain_run(PlainPlan* pp, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
               std::vector<PlainRaw>& out, double* kernel_ms, uint64_t* n_launches) {
    out.clear();
    if (plain_plan_empty(pp) || scan_end == 0) return;
    uint64_t n_tiles = (scan_end + kTile - 1) / kTile;
    for (;;) {
        HIPCHECK(hipMemsetAsync(pp->d_count, 0, sizeof(uint32_t), st));
        HIPCHECK(hipEventRecord(pp->ev0, st));
        for (const auto& l : pp->launches) {
            // grids of at most 2^20 tiles keep the tile index in 32 bits
            for (uint64_t t0 = 0; t0 < n_tiles; t0 += (1u << 20)) {
                uint32_t nb = (uint32_t)std::min<uint64_t>(n_tiles - t0, 1u << 20);
                uint64_t first = t0 * kTile;
                const GroupDev* g = pp->d_groups + l.group_begin;
#define COF_PLAIN_LAUNCH(WW)                                                                              \
    case WW:                                                                                              \
        hipLaunchKernelGGL(plain_scan_kernel<WW>, dim3(nb), dim3(kThreads), 0, st, d_genome, scan_end, first, g, \
                           l.n_groups, pp->d_cmp, pp->d_thr, pp->d_full, pp->d_mm, pp->d_pid, pp->d_out,  \
                           pp->d_count, pp->out_cap);                                                    \
        break;
                switch (l.W) {
                    COF_PLAIN_LAUNCH(1) COF_PLAIN_LAUNCH(2) COF_PLAIN_LAUNCH(3) COF_PLAIN_LAUNCH(4)
                    COF_PLAIN_LAUNCH(5) COF_PLAIN_LAUNCH(6) COF_PLAIN_LAUNCH(7) COF_PLAIN_LAUNCH(8)
                    default: throw Error(COF_E_UNSUPPORTED, "compare window too wide");
                }
#undef COF_PLAIN_LAUNCH
                HIPCHECK(hipGetLastError());
                if (n_launches) (*n_launches)++;
            }
        }
        HIPCHECK(hipEventRecord(pp->ev1, st));
        uint32_t cnt = 0;
        HIPCHECK(hipMemcpyAsync(&cnt, pp->d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, pp->ev0, pp->ev1));
        if (kernel_ms) *kernel_ms += ms;
        if (cnt > pp->out_cap) {
            // too many hits for the record buffer: grow it and scan again
            (void)hipFree(pp->d_out);
            pp->out_cap = (uint32_t)std::min<uint64_t>((uint64_t)cnt + cnt / 4 + 1024, 0xFFFFFFF0ull);
            HIPCHECK(hipMalloc(&pp->d_out, (size_t)pp->out_cap * sizeof(PlainRaw)));
            continue;
        }
        out.resize(cnt);
        if (cnt) HIPCHECK(hipMemcpyAsync(out.data(), pp->d_out, (size_t)cnt * sizeof(PlainRaw), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return;
    }
}

}  // namespace cof
