// Bulge search (DNA/RNA bulges) on gfx950: PAM quick check -> Myers filter
// -> all-alignment DP traceback, on the resident genome.
//
// Semantics follow the reference's CPU bulge path: search_chunk_myers
// (cas-offinder-lib/src/search.rs:779-897) per virtual batch, with
// check_pam_quick (:703-755), the windowed Myers score (:841-877),
// traceback_all / enumerate (traceback.rs:269-469) and the full PAM check of
// classify_myers_candidate (search.rs:634-671).
//
// Two kernels and a candidate queue between them:
//   bulge_filter_kernel — one workgroup per tile of end positions inside one
//     virtual batch (its window floor and PAM limit come with the tile).  Lanes
//     run the exact quick check on their end position, survivors are compacted
//     with a wave ballot, and each lane then runs a bit-parallel Myers for every
//     pattern of the group.  The bit-vector spans the text window (<= 64 nt) and
//     pattern characters are streamed, so the five per-window match masks are
//     built once per candidate and reused by all patterns, and the score falls
//     out of two popcounts at the end.  (j, pattern) pairs with score <= E go
//     to the queue.
//   bulge_verify_kernel — one lane per queued candidate.  The Myers columns
//     of the window are kept in LDS; column i gives D[i][*] by a masked
//     popcount, a lower bound of the traceback DP (unit costs, padding
//     counted as a mismatch), which prunes an iterative DFS over the op
//     strings in enumerate()'s order and rules.  Leaves pass the full PAM
//     check and are written as records (or folded into a count + checksum).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>

#include "engine.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

constexpr int kFilterThreads = 256;
constexpr int kTileEnds = 4096;          // end positions per filter workgroup
constexpr int kMaxW = 64;                // text window (L + D) supported
constexpr int kMaxR = 16;
constexpr int kWinWords = (kTileEnds + kMaxW + kMaxR + 64) / 8 + 4;

struct BGroup {
    uint32_t pat_begin, pat_count;
    uint32_t L, D, R, W;
    uint32_t pam_off, pam_len;
    uint32_t pam_filter[64];
};

struct BPat {
    uint8_t cls[64];          // Myers class per position: 0 A, 1 C, 2 G, 3 T, 4 wildcard (myers.rs:20-50)
    uint64_t pa, pc, pg, pt;  // pattern bit4 planes over positions
    uint64_t fa, fc, fg, ft;  // effective-filter bit4 planes over positions
    uint64_t is_n;            // ASCII N / n positions
    uint32_t L, M, D, R, E, pid;
};

struct BTile {
    uint64_t j0;              // first end position (global nibble)
    uint64_t floor;           // window floor: start of the virtual batch buffer
    uint64_t limit;           // PAM limit: end of the batch's last chunk
    uint32_t n;               // end positions in the tile
    uint32_t pad;
};

struct BCand {
    uint64_t j;
    uint64_t floor;
    uint32_t pat;             // index into the BPat table
    uint32_t pad;
};

__device__ __forceinline__ uint32_t lds_nib(const uint32_t* w, uint32_t q) { return (w[q >> 3] >> ((q & 7u) * 4u)) & 0xFu; }
__device__ __forceinline__ uint32_t glob_nib(const uint32_t* __restrict__ g, uint64_t q) {
    return (g[q >> 3] >> ((uint32_t)(q & 7u) * 4u)) & 0xFu;
}

// pam_cell_ok of the reference (search.rs:737-743)
__device__ __forceinline__ bool pam_cell_ok(uint32_t g, uint32_t f) {
    if (f == 0xFu) return g != 0;
    if (g == 0 || g == 0xFu) return false;
    return (g & f) != 0;
}

// cmp_loose_bit4 (traceback.rs:129-134)
__device__ __forceinline__ bool cmp_loose(uint32_t p, uint32_t t) {
    if (t == 0) return false;
    if (p == 0xFu) return true;
    if (t == 0xFu) return false;
    return (p & t) != 0;
}

// Five match masks of a text window (bit a = window position a), Myers rule
// (search.rs:846-858): A/C/G/T need a non-N base with that bit, wildcards
// match any non-padding nibble.
template <typename WORD>
struct Masks {
    WORD m[5];
};

template <typename WORD>
__device__ __forceinline__ uint32_t popc_w(WORD x) {
    if constexpr (sizeof(WORD) == 8) return (uint32_t)__popcll((unsigned long long)x);
    else return (uint32_t)__popc((uint32_t)x);
}

template <typename WORD>
__global__ __launch_bounds__(kFilterThreads) void bulge_filter_kernel(
    const uint32_t* __restrict__ genome, const BTile* __restrict__ tiles, const BGroup* __restrict__ groups,
    uint32_t n_groups, const BPat* __restrict__ pats, BCand* __restrict__ queue, uint32_t* __restrict__ q_count,
    uint32_t q_cap) {
    __shared__ uint32_t gw[kWinWords];
    __shared__ uint16_t cand[kTileEnds];
    __shared__ uint32_t n_cand;

    const BTile T = tiles[blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    // LDS window: from the lowest text start any end in the tile can use
    const uint64_t lo = T.j0 > T.floor + (kMaxW + 8) ? T.j0 - (kMaxW + 8) : T.floor;
    const uint64_t wlo = lo >> 3;
    const uint64_t base = wlo * 8;
    for (uint32_t i = tid; i < (uint32_t)kWinWords; i += kFilterThreads) gw[i] = genome[wlo + i];
    __syncthreads();

    for (uint32_t gi = 0; gi < n_groups; gi++) {
        const BGroup& G = groups[gi];
        if (tid == 0) n_cand = 0;
        __syncthreads();
        // quick PAM check per end position (check_pam_quick, search.rs:703-755)
        const uint32_t dbr = G.pam_off == 0 ? G.D + 1 : 1, rbr = G.pam_off == 0 ? G.R + 1 : 1;
        for (uint32_t b0 = 0; b0 < T.n; b0 += kFilterThreads) {
            uint32_t e = b0 + tid;
            bool ok = false;
            if (e < T.n) {
                uint64_t j = T.j0 + e;
                uint64_t jl = j - T.floor;              // batch-local end position
                if (jl + 1 >= G.L) {
                    for (uint32_t bd = 0; bd < dbr && !ok; bd++) {
                        for (uint32_t br = 0; br < rbr && !ok; br++) {
                            int64_t span = (int64_t)G.L + bd - br;
                            if (span <= 0 || span > (int64_t)jl + 1) continue;
                            uint64_t as = j + 1 - (uint64_t)span;
                            bool good = true;
                            for (uint32_t k = 0; k < G.pam_len && good; k++) {
                                uint64_t gp = as + G.pam_off + k;
                                if (gp >= T.limit) { good = false; break; }
                                good = pam_cell_ok(lds_nib(gw, (uint32_t)(gp - base)), G.pam_filter[k]);
                            }
                            ok = good;
                        }
                    }
                }
            }
            uint64_t bal = __ballot(ok);
            uint32_t cnt = (uint32_t)__popcll(bal);
            uint32_t wb = 0;
            if (lane == 0 && cnt) wb = atomicAdd(&n_cand, cnt);
            wb = __shfl(wb, 0);
            if (ok) cand[wb + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))] = (uint16_t)e;
        }
        __syncthreads();
        const uint32_t nc = n_cand;
        for (uint32_t c0 = (tid & ~63u); c0 < nc; c0 += kFilterThreads) {
            uint32_t c = c0 + lane;
            bool active = c < nc;
            uint32_t e = active ? cand[c] : 0;
            uint64_t j = T.j0 + e;
            uint64_t ts = j + 1 >= T.floor + G.W ? j + 1 - G.W : T.floor;
            uint32_t n = (uint32_t)(j - ts + 1);
            Masks<WORD> mk;
#pragma unroll
            for (int x = 0; x < 5; x++) mk.m[x] = 0;
            for (uint32_t a = 0; a < n; a++) {
                uint32_t t = lds_nib(gw, (uint32_t)(ts + a - base));
                WORD bit = (WORD)1 << a;
                bool real = t != 0 && t != 0xFu;
                mk.m[0] |= (real && (t & 4u)) ? bit : (WORD)0;
                mk.m[1] |= (real && (t & 2u)) ? bit : (WORD)0;
                mk.m[2] |= (real && (t & 8u)) ? bit : (WORD)0;
                mk.m[3] |= (real && (t & 1u)) ? bit : (WORD)0;
                mk.m[4] |= t != 0 ? bit : (WORD)0;
            }
            WORD nmask = n >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << n) - 1);
            for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++) {
                const BPat& P = pats[p];
                WORD VP = 0, VN = 0;
                for (uint32_t i = 0; i < P.L; i++) {
                    WORD eq = mk.m[P.cls[i]];
                    WORD Xv = eq | VN;
                    WORD Xh = (((eq & VP) + VP) ^ VP) | eq;
                    WORD Ph = VN | ~(Xh | VP);
                    WORD Mh = VP & Xh;
                    Ph = (Ph << 1) | (WORD)1;
                    Mh = Mh << 1;
                    VP = Mh | ~(Xv | Ph);
                    VN = Ph & Xv;
                }
                int32_t score = (int32_t)P.L + (int32_t)popc_w<WORD>(VP & nmask) - (int32_t)popc_w<WORD>(VN & nmask);
                if (active && score <= (int32_t)P.E) {
                    uint32_t idx = atomicAdd(q_count, 1u);
                    if (idx < q_cap) queue[idx] = BCand{j, T.floor, p, 0u};
                }
            }
        }
        __syncthreads();
    }
}

// Hit sink: records and/or an order-independent checksum of the records.
struct Sink {
    BulgeRaw* out;
    uint32_t* count;
    uint32_t cap;
    unsigned long long* checksum;   // sum of mix64(record) (may be null)
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}

template <typename WORD>
__global__ __launch_bounds__(64) void bulge_verify_kernel(const uint32_t* __restrict__ genome,
                                                          const BCand* __restrict__ queue, uint32_t n_queue,
                                                          const BPat* __restrict__ pats, Sink sink) {
    // Myers columns of each lane's window: col i at [i][lane] (L <= window bits)
    __shared__ WORD cvp[sizeof(WORD) * 8 + 1][64];
    __shared__ WORD cvn[sizeof(WORD) * 8 + 1][64];
    const uint32_t lane = threadIdx.x;
    const uint32_t qi = blockIdx.x * 64u + lane;
    if (qi >= n_queue) return;
    const BCand C = queue[qi];
    const BPat& P = pats[C.pat];
    const uint32_t L = P.L, W = L + P.D;
    const uint64_t ts = C.j + 1 >= C.floor + W ? C.j + 1 - W : C.floor;
    const uint32_t n = (uint32_t)(C.j - ts + 1);

    // window planes: Myers masks, padding, genome N, and bit4 planes
    Masks<WORD> mk;
    WORD z = 0, tA = 0, tC = 0, tG = 0, tT = 0, tN = 0;
#pragma unroll
    for (int x = 0; x < 5; x++) mk.m[x] = 0;
    for (uint32_t a = 0; a < n; a++) {
        uint32_t t = glob_nib(genome, ts + a);
        WORD bit = (WORD)1 << a;
        bool real = t != 0 && t != 0xFu;
        mk.m[0] |= (real && (t & 4u)) ? bit : (WORD)0;
        mk.m[1] |= (real && (t & 2u)) ? bit : (WORD)0;
        mk.m[2] |= (real && (t & 8u)) ? bit : (WORD)0;
        mk.m[3] |= (real && (t & 1u)) ? bit : (WORD)0;
        mk.m[4] |= t != 0 ? bit : (WORD)0;
        z |= t == 0 ? bit : (WORD)0;
        tN |= t == 0xFu ? bit : (WORD)0;
        tA |= (t & 4u) ? bit : (WORD)0;
        tC |= (t & 2u) ? bit : (WORD)0;
        tG |= (t & 8u) ? bit : (WORD)0;
        tT |= (t & 1u) ? bit : (WORD)0;
    }
    cvp[0][lane] = 0;
    cvn[0][lane] = 0;
    {
        WORD VP = 0, VN = 0;
        for (uint32_t i = 0; i < L; i++) {
            WORD eq = mk.m[P.cls[i]];
            WORD Xv = eq | VN;
            WORD Xh = (((eq & VP) + VP) ^ VP) | eq;
            WORD Ph = VN | ~(Xh | VP);
            WORD Mh = VP & Xh;
            Ph = (Ph << 1) | (WORD)1;
            Mh = Mh << 1;
            VP = Mh | ~(Xv | Ph);
            VN = Ph & Xv;
            cvp[i + 1][lane] = VP;
            cvn[i + 1][lane] = VN;
        }
    }
    auto text_at = [&](uint32_t a) -> uint32_t {   // window nibble a
        uint32_t t = (uint32_t)((tT >> a) & 1) | (uint32_t)((tC >> a) & 1) << 1 | (uint32_t)((tA >> a) & 1) << 2 |
                     (uint32_t)((tG >> a) & 1) << 3;
        return t;
    };
    auto pat_at = [&](uint32_t k) -> uint32_t {
        return (uint32_t)((P.pt >> k) & 1) | (uint32_t)((P.pc >> k) & 1) << 1 | (uint32_t)((P.pa >> k) & 1) << 2 |
               (uint32_t)((P.pg >> k) & 1) << 3;
    };
    auto filt_at = [&](uint32_t k) -> uint32_t {
        return (uint32_t)((P.ft >> k) & 1) | (uint32_t)((P.fc >> k) & 1) << 1 | (uint32_t)((P.fa >> k) & 1) << 2 |
               (uint32_t)((P.fg >> k) & 1) << 3;
    };
    auto isn = [&](uint32_t k) -> bool { return (P.is_n >> k) & 1ull; };
    // lower bound D[i][jj] of the remaining cost from node (i, jj)
    auto lb = [&](uint32_t i, uint32_t jj) -> uint32_t {
        WORD m = jj >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << jj) - 1);
        return i + popc_w<WORD>(cvp[i][lane] & m) - popc_w<WORD>(cvn[i][lane] & m);
    };
    const uint32_t E = P.E, M = P.M, D = P.D, R = P.R;
    // node admission (enumerate's entry checks, traceback.rs:349-354; padding
    // cells are infinite in the reference DP, traceback.rs:297-300)
    auto admit = [&](uint32_t i, uint32_t jj, uint32_t cost, uint32_t mm, uint32_t db, uint32_t rb) -> bool {
        if (mm > M || db > D || rb > R) return false;
        if (i >= 1 && jj >= 1 && ((z >> (jj - 1)) & 1)) return false;
        return cost + lb(i, jj) <= E;
    };
    if (!admit(L, n, 0, 0, 0, 0)) return;

    // DFS state: ops / next-transition per depth, 2 bits each
    uint64_t ops[3] = {0, 0, 0}, trs[3] = {0, 0, 0};
    auto get2 = [](const uint64_t* a, uint32_t d) -> uint32_t {
        uint64_t w = d < 32 ? a[0] : d < 64 ? a[1] : a[2];
        return (uint32_t)(w >> (2 * (d & 31))) & 3u;
    };
    auto set2 = [](uint64_t* a, uint32_t d, uint32_t v) {
        uint64_t m = ~(3ull << (2 * (d & 31))), b = (uint64_t)v << (2 * (d & 31));
        if (d < 32) a[0] = (a[0] & m) | b;
        else if (d < 64) a[1] = (a[1] & m) | b;
        else a[2] = (a[2] & m) | b;
    };
    uint32_t i = L, jj = n, d = 0, cost = 0, mm = 0, db = 0, rb = 0;
    set2(trs, 0, 0);
    for (;;) {
        if (i == 0) {
            // leaf: full PAM check along the alignment (search.rs:638-671)
            bool ok = true;
            uint32_t pi = 0, a = jj;
            for (uint32_t k = 0; k < d && ok; k++) {
                uint32_t op = get2(ops, d - 1 - k);
                if (op <= 1) {
                    uint32_t gv = (z >> a) & 1 ? 0u : ((tN >> a) & 1) ? 0xFu : text_at(a);
                    ok = pam_cell_ok(gv, filt_at(pi));
                    pi++;
                    a++;
                } else if (op == 2) {
                    a++;
                } else {
                    pi++;
                }
            }
            if (ok) {
                BulgeRaw r;
                r.start = ts + jj;
                r.pattern = P.pid;
                r.mm = (uint8_t)mm;
                r.db = (uint8_t)db;
                r.rb = (uint8_t)rb;
                r.n_ops = (uint8_t)d;
                r.ops[0] = r.ops[1] = r.ops[2] = 0;
                for (uint32_t k = 0; k < d; k++) r.ops[k / 32] |= (uint64_t)get2(ops, d - 1 - k) << (2 * (k % 32));
                uint32_t idx = atomicAdd(sink.count, 1u);
                if (sink.out && idx < sink.cap) sink.out[idx] = r;
                if (sink.checksum) {
                    uint64_t h = mix64(r.start * 0x9E3779B97F4A7C15ull ^ ((uint64_t)r.pattern << 40) ^
                                       ((uint64_t)r.mm << 24) ^ ((uint64_t)r.db << 16) ^ ((uint64_t)r.rb << 8) ^ r.n_ops);
                    h = mix64(h ^ r.ops[0]) + mix64(h ^ r.ops[1] ^ 0x1234567ull) + mix64(h ^ r.ops[2] ^ 0x89ABCDEFull);
                    atomicAdd(sink.checksum, (unsigned long long)h);
                }
            }
            goto pop;
        }
        {
            uint32_t t = get2(trs, d);
            if (t == 0) {
                set2(trs, d, 1);
                if (jj > 0) {                                   // diagonal
                    uint32_t tv = (z >> (jj - 1)) & 1 ? 0u : ((tN >> (jj - 1)) & 1) ? 0xFu : text_at(jj - 1);
                    uint32_t mc = cmp_loose(pat_at(i - 1), tv) ? 0u : 1u;
                    if (!(mc && isn(i - 1)) && admit(i - 1, jj - 1, cost + mc, mm + mc, db, rb)) {
                        set2(ops, d, mc);
                        i--; jj--; cost += mc; mm += mc; d++;
                        set2(trs, d, 0);
                    }
                }
                continue;
            }
            if (t == 1) {
                set2(trs, d, 2);
                if (rb < R && R > 0) {                          // RNA bulge
                    bool ok = !isn(i - 1) && !(i < L && isn(i)) && !(i >= 2 && isn(i - 2));
                    if (ok && admit(i - 1, jj, cost + 1, mm, db, rb + 1)) {
                        set2(ops, d, 3);
                        i--; cost++; rb++; d++;
                        set2(trs, d, 0);
                    }
                }
                continue;
            }
            if (t == 2) {
                set2(trs, d, 3);
                if (db < D && D > 0 && jj > 0 && i < L) {       // DNA bulge
                    bool ok = !(isn(i - 1) && isn(i));
                    if (ok && admit(i, jj - 1, cost + 1, mm, db + 1, rb)) {
                        set2(ops, d, 2);
                        jj--; cost++; db++; d++;
                        set2(trs, d, 0);
                    }
                }
                continue;
            }
        }
    pop:
        if (d == 0) break;
        d--;
        {
            uint32_t op = get2(ops, d);
            if (op <= 1) { i++; jj++; cost -= op; mm -= op; }
            else if (op == 2) { jj++; cost--; db--; }
            else { i++; cost--; rb--; }
        }
    }
}

}  // namespace

struct BulgePlan {
    std::vector<BGroup> groups32, groups64;
    std::vector<BPat> pats;
    BGroup *d_g32 = nullptr, *d_g64 = nullptr;
    BPat* d_pats = nullptr;
    BCand* d_queue = nullptr;
    uint32_t q_cap = 0;
    uint32_t* d_qcount = nullptr;
    BulgeRaw* d_out = nullptr;
    uint32_t out_cap = 0;
    uint32_t* d_ocount = nullptr;
    unsigned long long* d_checksum = nullptr;
    BTile* d_tiles = nullptr;
    size_t tiles_cap = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool records = true;
    uint64_t last_checksum = 0, last_count = 0;
};

BulgePlan* bulge_plan_create(const Query& q) {
    auto* bp = new BulgePlan();
    if (!q.any_bulge) return bp;
    // group bulge patterns sharing the quick PAM check
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, std::vector<uint8_t>>, std::vector<uint32_t>> by_key;
    for (uint32_t pi = 0; pi < q.pats.size(); pi++) {
        const Pattern& p = q.pats[pi];
        if (!p.bulge) continue;
        if (p.L + p.D > (uint32_t)kMaxW) throw Error(COF_E_UNSUPPORTED, "bulge search supports L + DNA bulges <= 64");
        if (p.R > (uint32_t)kMaxR) throw Error(COF_E_UNSUPPORTED, "RNA bulges above 16 are not supported");
        // precompute_pam_precheck (search.rs:761-777)
        uint32_t ps = p.L;
        for (uint32_t i = 0; i < p.L; i++) if (is_n_char(p.ascii[i])) { ps = i; break; }
        std::vector<uint8_t> pf;
        for (uint32_t i = ps; i < p.L; i++)
            if (is_n_char(p.ascii[i])) pf.push_back(kCodec.enc_mixed[(uint8_t)p.eff_filter[i]]);
        by_key[std::make_tuple(p.L, p.D, p.R, ps, pf)].push_back(pi);
    }
    for (auto& kv : by_key) {
        BGroup G{};
        G.L = std::get<0>(kv.first);
        G.D = std::get<1>(kv.first);
        G.R = std::get<2>(kv.first);
        G.W = G.L + G.D;
        G.pam_off = std::get<3>(kv.first);
        const auto& pf = std::get<4>(kv.first);
        G.pam_len = (uint32_t)pf.size();
        for (size_t k = 0; k < pf.size(); k++) G.pam_filter[k] = pf[k];
        G.pat_begin = (uint32_t)bp->pats.size();
        G.pat_count = (uint32_t)kv.second.size();
        for (uint32_t pi : kv.second) {
            const Pattern& p = q.pats[pi];
            BPat P{};
            for (uint32_t i = 0; i < p.L; i++) {
                char u = (char)toupper((unsigned char)p.ascii[i]);
                P.cls[i] = u == 'A' ? 0 : u == 'C' ? 1 : u == 'G' ? 2 : u == 'T' ? 3 : 4;
                uint8_t v = p.b4[i], f = kCodec.enc_mixed[(uint8_t)p.eff_filter[i]];
                P.pt |= (uint64_t)(v & 1) << i; P.pc |= (uint64_t)((v >> 1) & 1) << i;
                P.pa |= (uint64_t)((v >> 2) & 1) << i; P.pg |= (uint64_t)((v >> 3) & 1) << i;
                P.ft |= (uint64_t)(f & 1) << i; P.fc |= (uint64_t)((f >> 1) & 1) << i;
                P.fa |= (uint64_t)((f >> 2) & 1) << i; P.fg |= (uint64_t)((f >> 3) & 1) << i;
                if (is_n_char(p.ascii[i])) P.is_n |= 1ull << i;
            }
            P.L = p.L; P.M = p.M; P.D = p.D; P.R = p.R; P.E = p.M + p.D + p.R; P.pid = pi;
            bp->pats.push_back(P);
        }
        (G.W <= 32 ? bp->groups32 : bp->groups64).push_back(G);
    }
    auto up = [](auto*& dptr, const auto& v) {
        if (v.empty()) return;
        HIPCHECK(hipMalloc(&dptr, v.size() * sizeof(v[0])));
        HIPCHECK(hipMemcpy(dptr, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    };
    up(bp->d_g32, bp->groups32);
    up(bp->d_g64, bp->groups64);
    up(bp->d_pats, bp->pats);
    bp->q_cap = 1u << 22;
    HIPCHECK(hipMalloc(&bp->d_queue, (size_t)bp->q_cap * sizeof(BCand)));
    HIPCHECK(hipMalloc(&bp->d_qcount, sizeof(uint32_t)));
    bp->out_cap = 1u << 22;
    HIPCHECK(hipMalloc(&bp->d_out, (size_t)bp->out_cap * sizeof(BulgeRaw)));
    HIPCHECK(hipMalloc(&bp->d_ocount, sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&bp->d_checksum, sizeof(unsigned long long)));
    for (auto& e : bp->ev) HIPCHECK(hipEventCreate(&e));
    return bp;
}

void bulge_plan_destroy(BulgePlan* bp) {
    if (!bp) return;
    for (void* p : {(void*)bp->d_g32, (void*)bp->d_g64, (void*)bp->d_pats, (void*)bp->d_queue, (void*)bp->d_qcount,
                    (void*)bp->d_out, (void*)bp->d_ocount, (void*)bp->d_checksum, (void*)bp->d_tiles})
        if (p) (void)hipFree(p);
    for (auto& e : bp->ev) if (e) (void)hipEventDestroy(e);
    delete bp;
}

uint64_t bulge_pad_words() { return (uint64_t)kWinWords + 64; }

void bulge_set_sink(BulgePlan* bp, bool records) { if (bp) bp->records = records; }
void bulge_last_checksum(const BulgePlan* bp, uint64_t* count, uint64_t* checksum) {
    *count = bp ? bp->last_count : 0;
    *checksum = bp ? bp->last_checksum : 0;
}

// Runs the filter over every virtual batch's end range, then the traceback
// over the queued candidates.  Tile ranges are processed in rounds so the
// candidate queue stays bounded; a round whose queue overflows is split.
void bulge_run(BulgePlan* bp, const Genome& g, const std::vector<Batch>& batches, const uint32_t* d_genome,
               hipStream_t st, std::vector<BulgeRaw>& out, cof_stats* stats) {
    out.clear();
    bp->last_checksum = 0;
    bp->last_count = 0;
    if (!bp || bp->pats.empty()) return;
    // tiles of end positions per virtual batch (search.rs:1080-1092, 899-927)
    std::vector<BTile> tiles;
    for (const Batch& b : batches) {
        uint64_t floor = g.chunks[b.first].nib_begin();
        uint64_t a = g.chunks[b.first + (b.has_head ? 1 : 0)].nib_begin();
        uint64_t e = b.last < g.chunks.size() ? g.chunks[b.last].nib_begin() : g.n_nibbles();
        for (uint64_t j0 = a; j0 < e; j0 += kTileEnds)
            tiles.push_back(BTile{j0, floor, e, (uint32_t)std::min<uint64_t>(kTileEnds, e - j0), 0u});
    }
    if (tiles.empty()) return;
    if (tiles.size() > bp->tiles_cap) {
        if (bp->d_tiles) (void)hipFree(bp->d_tiles);
        bp->tiles_cap = tiles.size();
        HIPCHECK(hipMalloc(&bp->d_tiles, tiles.size() * sizeof(BTile)));
    }
    HIPCHECK(hipMemcpyAsync(bp->d_tiles, tiles.data(), tiles.size() * sizeof(BTile), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(bp->d_checksum, 0, sizeof(unsigned long long), st));
    uint64_t total_hits = 0;
    const size_t round_tiles = 1u << 14;
    std::vector<std::pair<size_t, size_t>> rounds;
    for (size_t t = 0; t < tiles.size(); t += round_tiles) rounds.push_back({t, std::min(tiles.size(), t + round_tiles)});
    std::reverse(rounds.begin(), rounds.end());
    while (!rounds.empty()) {
        auto [t0, t1] = rounds.back();
        rounds.pop_back();
        HIPCHECK(hipMemsetAsync(bp->d_qcount, 0, sizeof(uint32_t), st));
        HIPCHECK(hipEventRecord(bp->ev[0], st));
        uint32_t nt = (uint32_t)(t1 - t0);
        if (!bp->groups32.empty())
            hipLaunchKernelGGL(bulge_filter_kernel<uint32_t>, dim3(nt), dim3(kFilterThreads), 0, st, d_genome,
                               bp->d_tiles + t0, bp->d_g32, (uint32_t)bp->groups32.size(), bp->d_pats, bp->d_queue,
                               bp->d_qcount, bp->q_cap);
        if (!bp->groups64.empty())
            hipLaunchKernelGGL(bulge_filter_kernel<uint64_t>, dim3(nt), dim3(kFilterThreads), 0, st, d_genome,
                               bp->d_tiles + t0, bp->d_g64, (uint32_t)bp->groups64.size(), bp->d_pats, bp->d_queue,
                               bp->d_qcount, bp->q_cap);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(bp->ev[1], st));
        uint32_t nq = 0;
        HIPCHECK(hipMemcpyAsync(&nq, bp->d_qcount, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, bp->ev[0], bp->ev[1]));
        if (stats) stats->kernel_ms_filter += ms;
        if (nq > bp->q_cap) {
            if (t1 - t0 > 1) {           // split the round (search.rs:1140-1152)
                size_t mid = t0 + (t1 - t0) / 2;
                rounds.push_back({mid, t1});
                rounds.push_back({t0, mid});
                continue;
            }
            (void)hipFree(bp->d_queue);  // a single tile overflows: grow the queue
            bp->q_cap = nq + nq / 4;
            HIPCHECK(hipMalloc(&bp->d_queue, (size_t)bp->q_cap * sizeof(BCand)));
            rounds.push_back({t0, t1});
            continue;
        }
        if (stats) stats->n_candidates += nq;
        if (nq == 0) continue;
        // verify; rerun with a larger record buffer when it overflows
        for (;;) {
            HIPCHECK(hipMemsetAsync(bp->d_ocount, 0, sizeof(uint32_t), st));
            HIPCHECK(hipEventRecord(bp->ev[2], st));
            Sink sk{bp->records ? bp->d_out : nullptr, bp->d_ocount, bp->records ? bp->out_cap : 0u,
                    bp->records ? nullptr : bp->d_checksum};
            uint32_t blocks = (nq + 63) / 64;
            // candidates of W <= 32 and W > 32 groups share the queue; the
            // 64-bit kernel handles every window size
            if (bp->groups64.empty())
                hipLaunchKernelGGL(bulge_verify_kernel<uint32_t>, dim3(blocks), dim3(64), 0, st, d_genome, bp->d_queue,
                                   nq, bp->d_pats, sk);
            else
                hipLaunchKernelGGL(bulge_verify_kernel<uint64_t>, dim3(blocks), dim3(64), 0, st, d_genome, bp->d_queue,
                                   nq, bp->d_pats, sk);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipEventRecord(bp->ev[3], st));
            uint32_t no = 0;
            HIPCHECK(hipMemcpyAsync(&no, bp->d_ocount, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            HIPCHECK(hipEventElapsedTime(&ms, bp->ev[2], bp->ev[3]));
            if (stats) stats->kernel_ms_verify += ms;
            if (bp->records && no > bp->out_cap) {
                (void)hipFree(bp->d_out);
                bp->out_cap = no + no / 4;
                HIPCHECK(hipMalloc(&bp->d_out, (size_t)bp->out_cap * sizeof(BulgeRaw)));
                continue;
            }
            total_hits += no;
            if (bp->records && no) {
                size_t old = out.size();
                out.resize(old + no);
                HIPCHECK(hipMemcpyAsync(out.data() + old, bp->d_out, (size_t)no * sizeof(BulgeRaw), hipMemcpyDeviceToHost, st));
                HIPCHECK(hipStreamSynchronize(st));
            }
            break;
        }
    }
    if (!bp->records) {
        unsigned long long cs = 0;
        HIPCHECK(hipMemcpyAsync(&cs, bp->d_checksum, sizeof(cs), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        bp->last_checksum = cs;
    }
    bp->last_count = total_hits;
}

}  // namespace cof
