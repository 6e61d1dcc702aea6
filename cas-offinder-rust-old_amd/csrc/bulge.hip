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
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <type_traits>

#include "engine.hpp"
#include "hip_check.hpp"

namespace cof {

namespace {

constexpr int kFilterThreads = 256;
constexpr int kStage = 512;                  // staged screen survivors per wave
// Staged filter survivors per wave.  Each flush reserves queue space with one
// atomic on the round's counter, whose return the wave waits for, so fewer,
// larger flushes pay: 1,024 entries (8 KB more LDS, three workgroups per CU
// instead of four) took config-3 filter rounds from 2.03 to 1.94 s per search
// against 512 (256: 2.74 s; 768: 1.99 s; 1,536: 1.96 s; 2,048: 2.07 s).  The
// 64-bit filter, whose match tables take twice the LDS, keeps 512 and three
// workgroups per CU.
template <typename WORD>
constexpr int filter_stage() {
    return sizeof(WORD) == 4 ? 1024 : 512;
}
constexpr uint32_t kRunGroups = 8;           // groups a filter workgroup scans together
constexpr uint32_t kRunPats = 256;           // a run stops growing once it holds this many patterns
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
    uint64_t cls_bits[4];     // Myers class per position, 3 bits each, 21 per word:
                              // 0 A, 1 C, 2 G, 3 T, 4 wildcard (myers.rs:20-50)
    uint64_t pa, pc, pg, pt;  // pattern bit4 planes over positions
    uint64_t fa, fc, fg, ft;  // effective-filter bit4 planes over positions
    uint64_t is_n;            // ASCII N / n positions
    uint64_t hash_id;         // pattern_id_of (checksum key)
    uint64_t peq[5];          // Myers match masks per text base A, C, G, T and for N (myers.rs:20-50)
    uint32_t L, M, D, R, E, pid;
};

struct BTile {
    uint64_t j0;              // first end position (global nibble)
    uint64_t floor;           // window floor: start of the virtual batch buffer
    uint64_t limit;           // PAM limit: end of the batch's last chunk
    uint32_t n;               // end positions in the tile
    uint32_t last_chunk;      // one past the batch's last chunk
};

struct BCand {
    uint64_t j;
    uint64_t floor;
    uint32_t pat;             // index into the BPat table
    uint32_t last_chunk;
    uint32_t tile;            // filter tile (work accounting for the cost model)
    uint32_t pad;
};

// Chunk table for the device-side drop rule of convert_matches_myers
// (search.rs:1338-1350).
struct DChunk {
    uint64_t nib_begin;
    uint32_t size;
    uint32_t chrom_start;     // chunk starts a chromosome
};

// LDS exchange between the lanes of one wave: no hardware barrier is needed
// (a wave's LDS operations complete in issue order), only ordering at
// wavefront scope, which the fences give without emitting instructions.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// The same two tests as 0/1 values without branches (the traceback's hot
// path evaluates them for 64 lanes at once).
__device__ __forceinline__ uint32_t pam_cell_ok01(uint32_t g, uint32_t f) {
    const uint32_t any = g != 0u;
    const uint32_t spec = (g != 0u) & (g != 0xFu) & ((g & f) != 0u);
    return f == 0xFu ? any : spec;
}
__device__ __forceinline__ uint32_t cmp_loose01(uint32_t p, uint32_t t) {
    return (t != 0u) & ((p == 0xFu) | ((t != 0xFu) & ((p & t) != 0u)));
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

// bit 0 of nibble j -> bit j
__device__ __forceinline__ uint32_t gather_bit0(uint32_t m) {
    m &= 0x11111111u;
    m = (m | (m >> 3)) & 0x03030303u;
    m = (m | (m >> 6)) & 0x000F000Fu;
    return (m | (m >> 12)) & 0xFFu;
}

// Myers match masks of the n-nibble window whose words (aligned to the window
// start) are X[0..): per word, SWAR tests give 8 positions at once.
template <typename WORD>
__device__ __forceinline__ void window_masks(const uint32_t* X, uint32_t n, Masks<WORD>& mk, WORD* zero_mask,
                                             WORD* n_mask, WORD* bA, WORD* bC, WORD* bG, WORD* bT) {
#pragma unroll
    for (int c = 0; c < 5; c++) mk.m[c] = 0;
    WORD z = 0, nn = 0, a = 0, cc = 0, g = 0, t = 0;
#pragma unroll
    for (int k = 0; k < (int)sizeof(WORD) * 2; k++) {
        if ((uint32_t)k * 8 >= n) break;
        uint32_t x = X[k];
        uint32_t f = x & (x >> 1) & (x >> 2) & (x >> 3) & 0x11111111u;           // nibble == 0xF
        uint32_t nz = ((((x & 0x77777777u) + 0x77777777u) | x) >> 3) & 0x11111111u;  // nibble != 0
        uint32_t real = nz & ~f;
        uint32_t sh = (uint32_t)k * 8;
        mk.m[0] |= (WORD)gather_bit0((x >> 2) & real) << sh;
        mk.m[1] |= (WORD)gather_bit0((x >> 1) & real) << sh;
        mk.m[2] |= (WORD)gather_bit0((x >> 3) & real) << sh;
        mk.m[3] |= (WORD)gather_bit0(x & real) << sh;
        mk.m[4] |= (WORD)gather_bit0(nz) << sh;
        if (zero_mask) {
            z |= (WORD)gather_bit0(~nz) << sh;
            nn |= (WORD)gather_bit0(f) << sh;
            a |= (WORD)gather_bit0(x >> 2) << sh;
            cc |= (WORD)gather_bit0(x >> 1) << sh;
            g |= (WORD)gather_bit0(x >> 3) << sh;
            t |= (WORD)gather_bit0(x) << sh;
        }
    }
    WORD keep = n >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << n) - 1);
#pragma unroll
    for (int c = 0; c < 5; c++) mk.m[c] &= keep;
    if (zero_mask) {
        *zero_mask = z & keep; *n_mask = nn & keep;
        *bA = a & keep; *bC = cc & keep; *bG = g & keep; *bT = t & keep;
    }
}

// match mask of class c (A C G T wildcard): a 5-way select, no indexing
template <typename WORD>
__device__ __forceinline__ WORD class_mask(const Masks<WORD>& mk, uint32_t c) {
    WORD e01 = (c & 1u) ? mk.m[1] : mk.m[0];
    WORD e23 = (c & 1u) ? mk.m[3] : mk.m[2];
    WORD e = (c & 2u) ? e23 : e01;
    return (c & 4u) ? mk.m[4] : e;
}

template <typename WORD>
__device__ __forceinline__ void myers_step(WORD eq, WORD& VP, WORD& VN) {
    WORD Xv = eq | VN;
    WORD Xh = (((eq & VP) + VP) ^ VP) | eq;
    WORD Ph = VN | ~(Xh | VP);
    WORD Mh = VP & Xh;
    Ph = (Ph << 1) | (WORD)1;
    Mh = Mh << 1;
    VP = Mh | ~(Xv | Ph);
    VN = Ph & Xv;
}

// 0aaaa... -> 0a0a0a...: bit k of x to bit 2k
__device__ __forceinline__ uint64_t spread2(uint32_t x) {
    uint64_t v = x;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

// Ops of a diagonal chain of n steps pushed at depths [d, d + n): step k
// covers pattern position n - 1 - k and is a mismatch (op 1) where bit
// n - 1 - k of `mis` is set, else a match (op 0).  2 bits per depth, depth k
// at bits 2k of the u64 words of `ops`; depths past d + n are left as they are.
template <typename WORD, int kOpsW>
__device__ __forceinline__ void put_tail_ops(uint64_t (&ops)[kOpsW], uint32_t d, WORD mis, uint32_t n) {
    constexpr uint32_t kBits = sizeof(WORD) * 8;
    WORD rev;
    if constexpr (sizeof(WORD) == 8) rev = (WORD)__builtin_bitreverse64((uint64_t)mis);
    else rev = (WORD)__builtin_bitreverse32((uint32_t)mis);
    rev = n >= kBits ? rev : rev >> (kBits - n);               // bit k = bit n - 1 - k of mis
    const WORD keep = n >= kBits ? ~(WORD)0 : (((WORD)1 << n) - 1);
    // 2n-bit field of ops (s1:s0) and its mask (k1:k0), placed at bit 2d
    const uint64_t s0 = spread2((uint32_t)rev), k0 = spread2((uint32_t)keep) * 3ull;
    uint64_t s1 = 0, k1 = 0;
    if constexpr (sizeof(WORD) == 8) {
        s1 = spread2((uint32_t)((uint64_t)rev >> 32));
        k1 = spread2((uint32_t)((uint64_t)keep >> 32)) * 3ull;
    }
    const uint32_t b = 2u * d;
    if constexpr (kOpsW == 1) {
        // n + d <= 32: the field fits one word
        ops[0] = (ops[0] & ~(k0 << b)) | (s0 << b);
    } else {
        // 128-bit field shifted by b in [0, 128) over ops[0], ops[1]
        uint64_t m0, m1, v0, v1;
        if (b == 0) { m0 = k0; m1 = k1; v0 = s0; v1 = s1; }
        else if (b < 64) {
            m0 = k0 << b; m1 = (k1 << b) | (k0 >> (64 - b));
            v0 = s0 << b; v1 = (s1 << b) | (s0 >> (64 - b));
        } else {
            m0 = 0; v0 = 0;
            m1 = k0 << (b - 64); v1 = s0 << (b - 64);
        }
        ops[0] = (ops[0] & ~m0) | v0;
        ops[1] = (ops[1] & ~m1) | v1;
    }
}

template <typename WORD>
__device__ __forceinline__ uint32_t popc_w(WORD x) {
    if constexpr (sizeof(WORD) == 8) return (uint32_t)__popcll((unsigned long long)x);
    else return (uint32_t)__popc((uint32_t)x);
}

// Quick PAM check, bit-parallel.  A cell-pass bitmap marks, over the
// tile's LDS window, the positions whose nibble passes one PAM cell's filter
// (pam_cell_ok, search.rs:737-743), 32 positions per word from four genome
// words of SWAR tests; a group's quick bits for 32 ends are then the OR over
// its bulge shifts of the AND over its cells of the cell bitmaps at the
// cells' offsets, masked to the ends where the shift's window starts after
// the batch floor and its cells lie before the PAM limit.
constexpr int kPassWords = (kWinWords * 8 + 31) / 32;
constexpr uint32_t kMaxPamCells = 8;      // more cells: the per-end loop

__device__ __forceinline__ uint32_t cell_pass_word(const uint32_t* gw, uint32_t w, uint32_t f) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t gi = 4 * w + (uint32_t)k;
        const uint32_t x = gi < (uint32_t)kWinWords ? gw[gi] : 0u;
        const uint32_t nz_x = ((((x & 0x77777777u) + 0x77777777u) | x) >> 3) & 0x11111111u;
        uint32_t pass;
        if (f == 0xFu) {
            pass = nz_x;                                   // any nibble but padding
        } else {
            const uint32_t y = x & (f * 0x11111111u);
            const uint32_t nz_y = ((((y & 0x77777777u) + 0x77777777u) | y) >> 3) & 0x11111111u;
            const uint32_t isF = x & (x >> 1) & (x >> 2) & (x >> 3) & 0x11111111u;
            pass = nz_y & ~isF;                            // a shared base, not N
        }
        r |= gather_bit0(pass) << (8 * k);
    }
    return r;
}

// 32 bits of bitmap P starting at bit o (bits outside [0, 32 kPassWords) are 0)
__device__ __forceinline__ uint32_t bits_at(const uint32_t* P, int64_t o) {
    const int64_t i = o >> 5;
    const uint32_t r = (uint32_t)o & 31u;
    const uint32_t lo = (i >= 0 && i < kPassWords) ? P[i] : 0u;
    const uint32_t hi = (i + 1 >= 0 && i + 1 < kPassWords) ? P[i + 1] : 0u;
    return __builtin_amdgcn_alignbit(hi, lo, r);
}

// the ends e in [e0w, e0w + 32) with lo <= e < hi, as bits
__device__ __forceinline__ uint32_t range_bits(int64_t e0w, int64_t lo, int64_t hi) {
    uint32_t m = ~0u;
    if (lo > e0w) m = lo >= e0w + 32 ? 0u : m & (~0u << (uint32_t)(lo - e0w));
    if (hi < e0w + 32) m = hi <= e0w ? 0u : m & ((1u << (uint32_t)(hi - e0w)) - 1u);
    return m;
}

// One step of Myers' bit-parallel edit distance with the pattern in the
// bit-vector and the text streaming (semi-global: any text start); score is
// the distance of the whole pattern to the best text suffix so far.  Bits
// above L - 1 may hold garbage: carries and shifts only move upward, so they
// never reach the bits that are read.
template <typename WORD>
__device__ __forceinline__ void myers_scan_step(WORD eq, WORD& VP, WORD& VN, uint32_t& score, uint32_t last) {
    const WORD X = eq | VN;
    const WORD D0 = (((X & VP) + VP) ^ VP) | X;
    const WORD HN = VP & D0;
    const WORD HP = VN | ~(VP | D0);
    const WORD Xs = HP << 1;
    VN = Xs & D0;
    VP = (HN << 1) | ~(Xs | D0);
    score = score + (uint32_t)((HP >> last) & 1) - (uint32_t)((HN >> last) & 1);
}

// Filter of one tile of end positions for every group: the exact quick PAM
// check per end position (check_pam_quick, search.rs:703-755) into a bitmap,
// then for each pattern a Myers scan along the tile.  The reference scores
// each end position's own window of W = L + D characters
// (search.rs:841-877); a scan started W - 1 characters before its first end
// position scores each end over all text starts from there, which is never
// more than the window score, so it only pre-selects: a pre-selected end
// whose quick check passed is scored again over its exact window
// (transposed Myers, text in the bit-vector) before it is queued.  Lanes take
// (pattern, segment) tasks, each lane scanning its own segment of the tile's
// LDS copy of the genome with its own nibble -> match-mask table in LDS, so
// groups of any size keep the lanes busy.
template <typename WORD>
__global__ __launch_bounds__(kFilterThreads) void bulge_filter_kernel(
    const uint32_t* __restrict__ genome, const BTile* __restrict__ tiles, const BGroup* __restrict__ groups,
    uint32_t n_groups, const BPat* __restrict__ pats, BCand* __restrict__ queue, uint32_t* __restrict__ q_count,
    uint32_t q_cap, const uint32_t* __restrict__ tile_ids, uint32_t* __restrict__ tile_cand) {
    __shared__ uint32_t gw[kWinWords];
    __shared__ uint32_t qbits[kRunGroups][kTileEnds / 32];   // quick-check bitmap per group of a run
    __shared__ uint32_t pbits[kMaxPamCells][kPassWords];      // cell-pass bitmaps of the current group
    __shared__ WORD eqt[kFilterThreads / 64][16][64];
    // per-wave staging of pre-selected ends: (end offset << 20 | pattern -
    // group base); they are re-scored and queued a wave's batch at a time
    constexpr int kFStage = filter_stage<WORD>();
    __shared__ uint32_t stage[kFilterThreads / 64][kFStage];

    const uint32_t tile_id = tile_ids[blockIdx.x];
    const BTile T = tiles[tile_id];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t my_cand = 0;
    uint32_t* my_stage = stage[wave];
    uint32_t n_st = 0;   // staged entries of this wave (wave-uniform)
    // LDS window: from the lowest text start any end in the tile can use
    const uint64_t lo = T.j0 > T.floor + (kMaxW + 8) ? T.j0 - (kMaxW + 8) : T.floor;
    const uint64_t wlo = lo >> 3;
    const uint64_t base = wlo * 8;
    const uint32_t j_off = (uint32_t)(T.j0 - base);     // LDS nibble of end offset 0
    for (uint32_t i = tid; i < (uint32_t)kWinWords; i += kFilterThreads) gw[i] = genome[wlo + i];
    __syncthreads();

    // Exact window score of the staged ends; the survivors are compacted in
    // place and then queued with one reservation.  Once the queue count
    // exceeds q_cap it stops growing (the host redoes such a round), so the
    // 32-bit count cannot wrap.
    auto flush = [&](uint32_t P0) {
        wave_lds_sync();
        uint32_t n_pass = 0;    // survivors so far, at my_stage[0, n_pass) (<= entries read)
        for (uint32_t i0 = 0; i0 < n_st; i0 += 64) {
            const uint32_t i = i0 + lane;
            bool pass = false;
            uint32_t v = 0;
            if (i < n_st) {
                v = my_stage[i];
                const BPat& P = pats[P0 + (v & 0xFFFFFu)];
                const uint32_t W = P.L + P.D;
                const uint64_t j = T.j0 + (v >> 20);
                const uint64_t ts = j + 1 >= T.floor + W ? j + 1 - W : T.floor;
                const uint32_t n = (uint32_t)(j - ts + 1);
                const uint32_t q = (uint32_t)(ts - base);
                const uint32_t qw = q >> 3, sh = (q & 7u) * 4u;
                uint32_t X[sizeof(WORD) * 2];
#pragma unroll
                for (int k = 0; k < (int)sizeof(WORD) * 2; k++)
                    X[k] = __builtin_amdgcn_alignbit(gw[qw + k + 1], gw[qw + k], sh);
                Masks<WORD> mk;
                window_masks<WORD>(X, n, mk, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
                const WORD nmask = n >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << n) - 1);
                const uint32_t L = P.L;
                WORD VP = 0, VN = 0;
                for (uint32_t w = 0, ii = 0; ii < L; w++) {
                    uint64_t cw = P.cls_bits[w];
                    for (uint32_t k = 0; k < 21 && ii < L; k++, ii++, cw >>= 3)
                        myers_step<WORD>(class_mask<WORD>(mk, (uint32_t)(cw & 7u)), VP, VN);
                }
                const int32_t score = (int32_t)L + (int32_t)popc_w<WORD>(VP & nmask) - (int32_t)popc_w<WORD>(VN & nmask);
                pass = score <= (int32_t)P.E;
            }
            const uint64_t pb = __ballot(pass);
            const uint32_t off =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(pb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pb, 0u));
            wave_lds_sync();        // every lane has read its entry before any is overwritten
            if (pass) my_stage[n_pass + off] = v;
            n_pass += (uint32_t)__popcll(pb);
        }
        wave_lds_sync();
        if (n_pass) {
            uint32_t qb = q_cap;
            if (lane == 0) {
                if (__hip_atomic_load(q_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= q_cap)
                    qb = atomicAdd(q_count, n_pass);
                my_cand += n_pass;
            }
            qb = __builtin_amdgcn_readlane(qb, 0);
            for (uint32_t i = lane; i < n_pass; i += 64) {
                const uint32_t v = my_stage[i];
                if (qb + i < q_cap)
                    queue[qb + i] = BCand{T.j0 + (v >> 20), T.floor, P0 + (v & 0xFFFFFu), T.last_chunk, tile_id, 0u};
            }
        }
        n_st = 0;
        wave_lds_sync();
    };

    // Groups are taken in runs of consecutive groups (their patterns are
    // consecutive in the table), up to kRunGroups groups or until the run
    // holds kRunPats patterns, and a run's patterns share one pool of scan
    // tasks: small groups (a few guides of one length, PAM and bulge limit)
    // would otherwise need short segments, whose warm-up dominates.
    for (uint32_t g0 = 0; g0 < n_groups;) {
        uint32_t g1 = g0 + 1, np = groups[g0].pat_count, W = groups[g0].W;
        while (g1 < n_groups && g1 - g0 < kRunGroups && np < kRunPats && groups[g1].pat_begin == groups[g0].pat_begin + np) {
            np += groups[g1].pat_count;
            W = max(W, groups[g1].W);
            g1++;
        }
        const uint32_t P0 = groups[g0].pat_begin;
        // quick PAM check per end position and group (check_pam_quick, search.rs:703-755)
        for (uint32_t gr = 0; gr < g1 - g0; gr++) {
            const BGroup& G = groups[g0 + gr];
            const uint32_t dbr = G.pam_off == 0 ? G.D + 1 : 1, rbr = G.pam_off == 0 ? G.R + 1 : 1;
            if (G.pam_len <= kMaxPamCells) {
                for (uint32_t k = 0; k < G.pam_len; k++)
                    for (uint32_t w = tid; w < (uint32_t)kPassWords; w += kFilterThreads)
                        pbits[k][w] = cell_pass_word(gw, w, G.pam_filter[k]);
                __syncthreads();
                const int64_t jb = (int64_t)(T.j0 - base);    // LDS position of end offset 0
                for (uint32_t w = tid; w < (uint32_t)kTileEnds / 32u; w += kFilterThreads) {
                    const int64_t e0w = 32 * (int64_t)w;
                    uint32_t Q = 0;
                    for (uint32_t bd = 0; bd < dbr; bd++)
                        for (uint32_t br = 0; br < rbr; br++) {
                            const int64_t span = (int64_t)G.L + bd - br;
                            if (span <= 0) continue;
                            // window start after the batch floor; every cell before the PAM limit
                            const int64_t hi = G.pam_len ? (int64_t)T.limit + span - (int64_t)G.pam_off -
                                                               (int64_t)G.pam_len - (int64_t)T.j0
                                                         : (int64_t)1 << 40;
                            uint32_t A = range_bits(e0w, (int64_t)T.floor + span - 1 - (int64_t)T.j0, hi);
                            for (uint32_t k = 0; k < G.pam_len && A; k++)
                                A &= bits_at(pbits[k], jb + 1 - span + (int64_t)G.pam_off + k + e0w);
                            Q |= A;
                        }
                    qbits[gr][w] = Q & range_bits(e0w, (int64_t)T.floor + G.L - 1 - (int64_t)T.j0, (int64_t)T.n);
                }
                __syncthreads();      // the next group's cell bitmaps overwrite these
                continue;
            }
            for (uint32_t b0 = 0; b0 < (uint32_t)kTileEnds; b0 += kFilterThreads) {
                const uint32_t e = b0 + tid;
                bool ok = false;
                if (e < T.n) {
                    const uint64_t j = T.j0 + e;
                    const uint64_t jl = j - T.floor;              // batch-local end position
                    if (jl + 1 >= G.L) {
                        for (uint32_t bd = 0; bd < dbr && !ok; bd++) {
                            for (uint32_t br = 0; br < rbr && !ok; br++) {
                                const int64_t span = (int64_t)G.L + bd - br;
                                if (span <= 0 || span > (int64_t)jl + 1) continue;
                                const uint64_t as = j + 1 - (uint64_t)span;
                                bool good = true;
                                for (uint32_t k = 0; k < G.pam_len && good; k++) {
                                    const uint64_t gp = as + G.pam_off + k;
                                    if (gp >= T.limit) { good = false; break; }
                                    good = pam_cell_ok(lds_nib(gw, (uint32_t)(gp - base)), G.pam_filter[k]);
                                }
                                ok = good;
                            }
                        }
                    }
                }
                const uint64_t bal = __ballot(ok);
                if (lane == 0) {
                    qbits[gr][(b0 >> 5) + wave * 2] = (uint32_t)bal;
                    qbits[gr][(b0 >> 5) + wave * 2 + 1] = (uint32_t)(bal >> 32);
                }
            }
        }
        __syncthreads();
        // tasks (pattern of the run, segment of S end positions): segments
        // shrink until the run's tasks fill the workgroup a few times over,
        // down to 64 ends (128 for 64-bit windows, whose warm-up is longer)
        const uint32_t s_min = sizeof(WORD) == 8 ? 128u : 64u;
        uint32_t nseg = 1;
        while (nseg * np < 4u * kFilterThreads && (uint32_t)kTileEnds / (nseg * 2u) >= s_min) nseg *= 2;
        const uint32_t S = (uint32_t)kTileEnds / nseg;
        const uint32_t n_tasks = np * nseg;
        // every lane of a wave steps through the same number of characters:
        // the run's widest window - 1 warm-up characters (more than a
        // narrower pattern needs: harmless), then S scored end positions
        for (uint32_t t0 = wave * 64; t0 < n_tasks; t0 += kFilterThreads) {
            const uint32_t task = t0 + lane;
            const bool act = task < n_tasks;
            // consecutive lanes take consecutive patterns of one segment, so
            // their text reads are LDS broadcasts (segments lie a multiple of
            // 32 words apart: one bank)
            const uint32_t pi = act ? task % np : 0u;
            const uint32_t e0 = (act ? task / np : 0u) * S;
            const BPat& P = pats[P0 + pi];
            const uint32_t L = P.L, E = P.E;
            // the lane's group within the run (its quick-check bitmap)
            uint32_t gr = 0;
            for (uint32_t k = 1; k < g1 - g0; k++) gr += P0 + pi >= groups[g0 + k].pat_begin ? 1u : 0u;
            const uint32_t* qb_row = qbits[gr];
            // the lane's match mask for every text nibble (myers.rs:20-50)
            const WORD pA = (WORD)P.peq[0], pC = (WORD)P.peq[1], pG = (WORD)P.peq[2], pT = (WORD)P.peq[3],
                       pN = (WORD)P.peq[4];
#pragma unroll
            for (uint32_t c = 0; c < 16u; c++) {
                WORD m = ((c & 4u) ? pA : (WORD)0) | ((c & 2u) ? pC : (WORD)0) | ((c & 8u) ? pG : (WORD)0) |
                         ((c & 1u) ? pT : (WORD)0);
                eqt[wave][c][lane] = c == 0u ? (WORD)0 : c == 15u ? pN : m;
            }
            // The scan starts at least W - 1 characters before the segment's
            // first end, in blocks of 8 characters (one LDS word each, at
            // the same nibble offset for every lane: segment starts are
            // multiples of 64), or at the window's first character when
            // there is not that much text before it.
            const uint32_t qa = j_off + e0;                 // character of end offset e0
            const uint32_t sh = (qa & 7u) * 4u;
            const uint32_t nw = (W + 6u) & ~7u;            // >= W - 1, a multiple of 8
            const uint32_t last = L - 1;
            WORD VP = ~(WORD)0, VN = 0;
            uint32_t score = L;
            if (qa >= nw) {
                for (uint32_t wi = (qa - nw) >> 3; wi < qa >> 3; wi++) {
                    const uint32_t x = __builtin_amdgcn_alignbit(gw[wi + 1], gw[wi], sh);
                    WORD eqk[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) eqk[u] = eqt[wave][(x >> (4 * u)) & 15u][lane];
#pragma unroll
                    for (int u = 0; u < 8; u++) myers_scan_step<WORD>(eqk[u], VP, VN, score, last);
                }
            } else {
                for (uint32_t q = 0; q < qa; q++) myers_scan_step<WORD>(eqt[wave][lds_nib(gw, q)][lane], VP, VN, score, last);
            }
            const uint32_t e_end = act ? min(e0 + S, T.n) : 0u;
            for (uint32_t k0 = 0; k0 < S; k0 += 8) {
                const uint32_t e = e0 + k0;
                const uint32_t wi = (qa + k0) >> 3;
                const uint32_t x = __builtin_amdgcn_alignbit(gw[wi + 1], gw[wi], sh);
                WORD eqk[8];
#pragma unroll
                for (int u = 0; u < 8; u++) eqk[u] = eqt[wave][(x >> (4 * u)) & 15u][lane];
                // quick-check bits of ends e .. e + 7 (e is a multiple of 8)
                uint32_t hit = (qb_row[e >> 5] >> (e & 31u)) & 0xFFu;
                if (e + 8u > e_end) hit &= e_end > e ? ((1u << (e_end - e)) - 1u) : 0u;
                uint32_t low = 0;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    myers_scan_step<WORD>(eqk[u], VP, VN, score, last);
                    low |= (score <= E ? 1u : 0u) << u;
                }
                hit &= low;
                if (__ballot(hit != 0u)) {
                    for (uint32_t u = 0; u < 8u; u++) {
                        const bool sel = (hit >> u) & 1u;
                        const uint64_t pb = __ballot(sel);
                        if (!pb) continue;
                        const uint32_t off =
                            __builtin_amdgcn_mbcnt_hi((uint32_t)(pb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pb, 0u));
                        if (sel) my_stage[n_st + off] = ((e + u) << 20) | pi;
                        n_st += (uint32_t)__popcll(pb);
                        if (n_st > (uint32_t)kFStage - 64u) flush(P0);
                    }
                }
            }
        }
        if (n_st) flush(P0);
        __syncthreads();
        g0 = g1;
    }
    // candidates of this tile (one workgroup per tile)
    for (int off = 32; off > 0; off >>= 1) my_cand += __shfl_down(my_cand, off);
    if (lane == 0 && my_cand) atomicAdd(&tile_cand[tile_id], my_cand);
}

// Hit sink: records (count <= cap written), or, in count-only mode, the
// device-side drop rule plus a kept count and checksum per wave.
struct Sink {
    BulgeRaw* out;
    uint32_t* count;
    uint32_t cap;
    int count_only;
    const DChunk* chunks;
    uint32_t n_chunks;
    unsigned long long* kept;
    unsigned long long* checksum;
    unsigned long long* diag;       // [0] DFS transitions, [1] lane slots of the DFS loop, [2] busy slots
    uint32_t* error;                // set when a lane reaches an impossible state
    uint32_t* cand_steps;           // DFS transitions per queue entry (null: not recorded)
    const uint32_t* chunk_lut;      // [k] = chunk holding nibble k << 16 (count-only drop rule)
    uint32_t refill;                // refill a wave once this many lanes are idle
};

// tile_steps[queue[i].tile] += steps[i] over the round's queue; a wave whose
// 64 entries share one tile (the usual case: a filter workgroup writes one
// tile's survivors contiguously) adds them with one atomic.
__global__ void tile_steps_reduce_kernel(const BCand* __restrict__ queue, const uint32_t* __restrict__ q_count,
                                         uint32_t q_cap, const uint32_t* __restrict__ steps,
                                         unsigned long long* __restrict__ tile_steps) {
    const uint32_t n = *q_count;
    if (n > q_cap) return;   // skipped round: redone synchronously
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ((i & ~63u) >= n) return;
    const bool valid = i < n;
    const uint32_t tile = valid ? queue[i].tile : 0xFFFFFFFFu;
    const unsigned long long v = valid ? steps[i] : 0ull;
    const uint32_t t0 = __shfl(tile, 0);
    if (__ballot(valid && tile != t0) == 0) {
        unsigned long long sum = v;
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if ((threadIdx.x & 63u) == 0) atomicAdd(&tile_steps[t0], sum);
    } else if (valid) {
        atomicAdd(&tile_steps[tile], v);
    }
}

// Clears tile_cand of the tiles order[0, n) before they are filtered again.
__global__ void clear_tile_cand_kernel(const uint32_t* __restrict__ order, uint32_t n, uint32_t* __restrict__ tile_cand) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tile_cand[order[i]] = 0;
}

// Clears n 32-bit words (graph-safe replacement for a memset node).
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Feasibility table of one candidate for the traceback (patterns with D, R
// <= 2 and M <= 6): the budget screen's banded recurrence (see
// bulge_screen_kernel), mismatch budgets 0..5 as 5-bit fields of one 32-bit
// word per (DNA, RNA) budget pair, run over the pattern rows.  Row i of the
// lane's table holds, at bits 3 (3d' + r'), the fewest mismatches (6: at
// least 6, or none — exact for every budget below 6, a relaxation at 6)
// with which the prefix of length i aligns within d' DNA and r' RNA bulges
// ending at band position d' + R - r' — the one end a DFS node of row i with
// remaining budgets (d', r') can have.  ok_b / mis_b: the traceback's
// diagonal masks of band diagonal b (admissible steps, mismatches).
template <typename WORD>
__device__ __forceinline__ void exact_table(uint32_t* xt, uint32_t lane, uint32_t L, uint32_t D, uint32_t R,
                                            WORD ok0, WORD ok1, WORD ok2, WORD ok3, WORD ok4, WORD mis0, WORD mis1,
                                            WORD mis2, WORD mis3, WORD mis4, WORD pn) {
    constexpr int KM = 5;
    using PK = uint32_t;
    constexpr PK kRep = [] {
        PK r = 0;
        for (int k = 0; k <= KM; k++) r |= (PK)1 << (5 * k);
        return r;
    }();
    constexpr PK kAll = kRep * (PK)31, kNotB0 = kAll & ~kRep, kNotB4 = kAll & ~(kRep << 4);
    // match steps (no mismatch) and mismatch steps of each band diagonal
    const WORD mt0 = ok0 & ~mis0, mt1 = ok1 & ~mis1, mt2 = ok2 & ~mis2, mt3 = ok3 & ~mis3, mt4 = ok4 & ~mis4;
    const WORD ms0 = ok0 & mis0, ms1 = ok1 & mis1, ms2 = ok2 & mis2, ms3 = ok3 & mis3, ms4 = ok4 & mis4;
    PK S[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int r = 0; r < 3; r++) S[d][r] = kAll;
    xt[lane] = 0u;                          // the empty prefix: no mismatch needed
    for (uint32_t i = 1; i <= L; i++) {
        const uint32_t p = i - 1;
        const uint32_t bmt = (uint32_t)((mt0 >> p) & 1) | (uint32_t)((mt1 >> p) & 1) << 1 |
                             (uint32_t)((mt2 >> p) & 1) << 2 | (uint32_t)((mt3 >> p) & 1) << 3 |
                             (uint32_t)((mt4 >> p) & 1) << 4;
        const uint32_t bms = (uint32_t)((ms0 >> p) & 1) | (uint32_t)((ms1 >> p) & 1) << 1 |
                             (uint32_t)((ms2 >> p) & 1) << 2 | (uint32_t)((ms3 >> p) & 1) << 3 |
                             (uint32_t)((ms4 >> p) & 1) << 4;
        const PK mt = (PK)bmt * kRep, ms = (PK)bms * kRep;
        const uint32_t n_p = (uint32_t)(pn >> p) & 1u;
        const uint32_t n_next = i < L ? (uint32_t)(pn >> i) & 1u : 0u;
        const uint32_t n_prev = i >= 2 ? (uint32_t)(pn >> (i - 2)) & 1u : 0u;
        const PK rna = (n_p | n_next | n_prev) ? (PK)0 : kNotB4;
        const PK dna = (i < L && !(n_p & n_next)) ? kNotB0 : (PK)0;
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int r = 2; r >= 0; r--) {
                PK v = (S[d][r] & mt) | ((S[d][r] << 5) & ms);
                if (r > 0) v |= (S[d][r - 1] >> 1) & rna;
                if (d > 0) v |= (S[d - 1][r] << 1) & dna;
                S[d][r] = v & kAll;
            }
        uint32_t row = 0xFFFFFFFFu;
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                if ((uint32_t)d > D || (uint32_t)r > R) continue;
                const uint32_t b = (uint32_t)d + R - (uint32_t)r;
                const PK y = (S[d][r] >> b) & kRep;
                const uint32_t mk = y ? ((uint32_t)__builtin_ctz(y) * 13u) >> 6 : 6u;   // field index = bit / 5
                const uint32_t sh = 3u * (3u * (uint32_t)d + (uint32_t)r);
                row = (row & ~(7u << sh)) | (mk << sh);
            }
        xt[i * 64 + lane] = row;
    }
}

template <typename WORD>
__global__ __launch_bounds__(64) void bulge_verify_kernel(const uint32_t* __restrict__ genome,
                                                          const BCand* __restrict__ queue,
                                                          const uint32_t* __restrict__ q_count, uint32_t q_cap,
                                                          uint32_t* __restrict__ overflow,
                                                          uint32_t* __restrict__ q_next, const BPat* __restrict__ pats,
                                                          uint32_t n_cols, uint32_t has_myers, uint32_t has_exact,
                                                          Sink sink) {
    // the candidate count comes from device memory so rounds need no host
    // round trip; an overflowed queue is skipped and redone by the host
    const uint32_t n_queue = *q_count;
    if (n_queue > q_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(overflow, 1u);
        return;
    }
    if (blockIdx.x * 64u >= n_queue) return;
    // depth d <= L + D = window length <= bits of WORD.  Ops are stored for
    // depths 0..W-1 (2 bits each).  The unexplored children of the node at
    // each depth are three bit planes (bit d of cm0/cm1/cm2: diagonal, RNA,
    // DNA), so no access indexes a register array.  Depth W is always a
    // leaf (L + D ops consume the whole pattern), whose child mask is 0 and
    // needs no storage: the planes are words of WORD's width.
    constexpr int kOpsW = sizeof(WORD) / 4;        // u64 words of ops
    using CM = WORD;
    constexpr uint32_t kCmBits = sizeof(CM) * 8;
    extern __shared__ unsigned char verify_lds[];
    // Per lane, rows 0..L of one of two bounds: the Myers columns (cvp/cvn,
    // D[i][*] by masked popcounts: a unit-cost lower bound) or, for patterns
    // with D, R <= 2 and M <= 6, the feasibility table xt: row i packs,
    // for each remaining budget pair (d', r') at 3 bits from 3 (3d' + r'),
    // the fewest mismatches with which pattern prefix i aligns within d'
    // DNA and r' RNA bulges ending at the one text end a node of row i with
    // those budgets can have (band position d' + R - r' of the budget
    // screen's recurrence; 6: at least 6 or none).  A child is pushed only if its prefix
    // can still be completed within every cap, which the unit bound alone
    // (cost + D <= M + D + R) does not ensure.
    WORD* cvp = reinterpret_cast<WORD*>(verify_lds);
    WORD* cvn = cvp + (has_myers ? (size_t)n_cols * 64 : 0);
    uint32_t* xt = reinterpret_cast<uint32_t*>(cvn + (has_myers ? (size_t)n_cols * 64 : 0));
    // Diagonal masks: every diagonal step of the DFS lies on one of the
    // D + R + 1 diagonals jj - i = n - L - D + b (b = D - db + rb).  For
    // diagonal b: over pattern positions p, the steps on it that are
    // admissible apart from the mismatch cap (text position in the window,
    // no padding, filter cell passes, no mismatch at an N), and the steps
    // that are mismatches.  With D + R <= 4 they are registers dk0..dk4 /
    // dm0..dm4, otherwise LDS rows 2b / 2b + 1 after the Myers columns.
    // Once both bulge budgets are spent every node lies on diagonal R and
    // only diagonal steps remain: such a node resolves its whole chain from
    // that diagonal's masks at once.
    WORD* dgm = reinterpret_cast<WORD*>(xt + (has_exact ? (size_t)n_cols * 64 : 0));
    const uint32_t lane = threadIdx.x;

    // candidate state
    uint64_t ts = 0;
    WORD z = 0;                       // window padding plane
    WORD rna_pm = 0, dna_pm = 0;      // bulges' pattern-side N rules: bit i - 1 / bit i of a node at row i
    int dg_base = 0;                  // n - L - D
    bool exact = false;               // the candidate uses the feasibility table xt
    uint32_t L = 0, M = 0, D = 0, R = 0, E = 0, pid = 0, last_chunk = 0;
    uint64_t hash_id = 0;
    WORD dk0 = 0, dk1 = 0, dk2 = 0, dk3 = 0, dk4 = 0, dm0 = 0, dm1 = 0, dm2 = 0, dm3 = 0, dm4 = 0;
    auto pick5 = [](uint32_t b, WORD a0, WORD a1, WORD a2, WORD a3, WORD a4) -> WORD {
        WORD r = a0;
        r = b == 1u ? a1 : r;
        r = b == 2u ? a2 : r;
        r = b == 3u ? a3 : r;
        r = b == 4u ? a4 : r;
        return r;
    };
    // admissible steps (ok) and mismatches (mis) of diagonal b
    auto diag_of = [&](uint32_t b, WORD& ok, WORD& mis) {
        if (D + R <= 4u) {
            ok = pick5(b, dk0, dk1, dk2, dk3, dk4);
            mis = pick5(b, dm0, dm1, dm2, dm3, dm4);
        } else {
            ok = dgm[(2u * b) * 64u + lane];
            mis = dgm[(2u * b + 1u) * 64u + lane];
        }
    };
    // DFS state
    bool busy = false;
    uint32_t i = 0, jj = 0, d = 0, cost = 0, mm = 0, db = 0, rb = 0;
    uint32_t d_floor = 0;   // depth the lane's subtree starts at (> 0 for donated work)
    __shared__ uint32_t pairing[64];
    uint64_t ops[kOpsW];
    CM cm0 = 0, cm1 = 0, cm2 = 0, cm3 = 0;   // cm3: the diagonal child is a mismatch
#pragma unroll
    for (int k = 0; k < kOpsW; k++) ops[k] = 0;
    uint64_t my_kept = 0, my_sum = 0, my_steps = 0, my_slots = 0, my_busy = 0;
    uint32_t cand_steps = 0, cand_idx = 0, cand_chunk = 0;
    uint64_t c0_begin = 0;
    uint32_t c0_size = 0;
    bool c0_endc = false;
    bool drained = false;

    auto bit = [](WORD x, uint32_t k) -> uint32_t { return (uint32_t)(x >> k) & 1u; };
    auto set2 = [&](uint32_t dd, uint32_t v) {
        uint64_t m = ~(3ull << (2 * (dd & 31))), b = (uint64_t)v << (2 * (dd & 31));
        if constexpr (kOpsW > 1) {
            if (dd < 32) ops[0] = (ops[0] & m) | b; else ops[1] = (ops[1] & m) | b;
        } else {
            ops[0] = (ops[0] & m) | b;
        }
    };
    // bits 0-2: unexplored children (diagonal, RNA, DNA); bit 3: diagonal is a mismatch
    auto get4 = [&](uint32_t dd) -> uint32_t {
        if (dd >= kCmBits) return 0u;
        return ((uint32_t)(cm0 >> dd) & 1u) | (((uint32_t)(cm1 >> dd) & 1u) << 1) | (((uint32_t)(cm2 >> dd) & 1u) << 2) |
               (((uint32_t)(cm3 >> dd) & 1u) << 3);
    };
    auto set4 = [&](uint32_t dd, uint32_t val) {
        if (dd >= kCmBits) return;
        const CM m = (CM)1 << dd;
        cm0 = (val & 1u) ? (cm0 | m) : (cm0 & ~m);
        cm1 = (val & 2u) ? (cm1 | m) : (cm1 & ~m);
        cm2 = (val & 4u) ? (cm2 | m) : (cm2 & ~m);
        cm3 = (val & 8u) ? (cm3 | m) : (cm3 & ~m);
    };
    // counts of the mismatch, DNA and RNA ops at depths [k, dd) (2-bit op
    // codes 0 match, 1 mismatch, 2 DNA, 3 RNA, depth k at bits 2k)
    auto op_counts = [&](uint32_t k, uint32_t dd, uint32_t& nS, uint32_t& nD, uint32_t& nR) {
        nS = nD = nR = 0;
#pragma unroll
        for (int w = 0; w < kOpsW; w++) {
            const uint32_t lo = (uint32_t)w * 32u;
            uint64_t keep = ~0ull;
            if (dd <= lo) keep = 0;
            else if (dd < lo + 32) keep = (1ull << (2 * (dd - lo))) - 1ull;
            if (k > lo) keep = k >= lo + 32 ? 0ull : keep & ~((1ull << (2 * (k - lo))) - 1ull);
            const uint64_t x = ops[w] & keep;
            const uint64_t b0 = x & 0x5555555555555555ull, b1 = (x >> 1) & 0x5555555555555555ull;
            nS += (uint32_t)__popcll(b0 & ~b1);
            nD += (uint32_t)__popcll(b1 & ~b0);
            nR += (uint32_t)__popcll(b0 & b1);
        }
    };
    auto lbc = [&](uint32_t ii, WORD vp, WORD vn, uint32_t jx) -> uint32_t {
        WORD m = jx >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << jx) - 1);
        return ii + popc_w<WORD>(vp & m) - popc_w<WORD>(vn & m);
    };
    // Admissible children of node (i, jj) as a mask: 1 diagonal, 2 RNA bulge,
    // 4 DNA bulge (enumerate(), traceback.rs:398-468).  Each child must pass
    // enumerate()'s entry checks (caps, cost + DP bound <= E, no padding
    // cell, traceback.rs:297-300,349-354); a diagonal step must also pass the
    // filter cell of its pattern position, which is where the full PAM check
    // of classify_myers_candidate (search.rs:638-671) would reject the leaf.
    auto children = [&]() -> uint32_t {
        if (i == 0) return 0u;
        // every test below is a 0/1 value (straight-line code, no lane branches);
        // indices are clamped where a test does not apply and masked by it
        const uint32_t has_j = jj > 0u;
        const uint32_t j1 = jj - has_j;                     // jj - 1 (or 0)
        const uint32_t has_j2 = jj >= 2u;
        const uint32_t z1 = has_j & bit(z, j1);             // padding at jj - 1
        const uint32_t z2 = has_j2 & bit(z, has_j2 ? jj - 2u : 0u);   // padding at jj - 2
        const uint32_t i2 = i >= 2u;
        // diagonal (match or substitution): pattern i - 1 on text jj - 1
        WORD okw, misw;
        diag_of((uint32_t)((int)jj - (int)i - dg_base), okw, misw);
        const uint32_t mc = bit(misw, i - 1);
        // can each child's prefix still be completed?
        uint32_t f_diag, f_rna, f_dna;
        if (exact) {
            const uint32_t xr1 = xt[(i - 1) * 64 + lane], xr0 = xt[i * 64 + lane];
            const uint32_t q = 3u * (D - db) + (R - rb), rem = M - mm;
            f_diag = ((xr1 >> (3u * q)) & 7u) + mc <= rem;
            f_rna = ((xr1 >> (rb < R ? 3u * (q - 1u) : 0u)) & 7u) <= rem;
            f_dna = ((xr0 >> (db < D ? 3u * (q - 3u) : 0u)) & 7u) <= rem;
        } else {
            const WORD vp1 = cvp[(i - 1) * 64 + lane], vn1 = cvn[(i - 1) * 64 + lane];
            const WORD vp0 = cvp[i * 64 + lane], vn0 = cvn[i * 64 + lane];
            f_diag = cost + mc + lbc(i - 1, vp1, vn1, j1) <= E;
            f_rna = cost + 1u + lbc(i - 1, vp1, vn1, jj) <= E;
            f_dna = cost + 1u + lbc(i, vp0, vn0, j1) <= E;
        }
        const uint32_t diag = bit(okw, i - 1) & (mm + mc <= M) & (1u - (i2 & z2)) & f_diag;
        // RNA bulge (skip a pattern base)
        const uint32_t rna = (rb < R) & bit(rna_pm, i - 1) & (1u - (i2 & z1)) & f_rna;
        // DNA bulge (skip a text base)
        const uint32_t dna = (db < D) & has_j & bit(dna_pm, i) & (1u - z2) & (1u - z1) & f_dna;
        return diag | (rna << 1) | (dna << 2) | (mc << 3);
    };
    // leaf (i == 0): emit the alignment (ops stay in stack order; the host
    // reverses recorded ones)
    auto emit = [&]() {
        // ops of depths [0, d) only: deeper slots hold stale ops
        const uint64_t o0 = d >= 32 ? ops[0] : ops[0] & ((1ull << (2 * d)) - 1ull);
        uint64_t o1 = 0;
        if constexpr (kOpsW > 1) o1 = d <= 32 ? 0ull : d >= 64 ? ops[1] : ops[1] & ((1ull << (2 * (d - 32))) - 1ull);
        const uint64_t st = ts + jj;
        if (!sink.count_only) {
            BulgeRaw r;
            r.start = st; r.pattern = pid;
            r.mm = (uint8_t)mm; r.db = (uint8_t)db; r.rb = (uint8_t)rb; r.n_ops = (uint8_t)d;
            r.ops[0] = o0; r.ops[1] = o1; r.ops[2] = 0;
            // one reservation for the lanes emitting in this step
            const uint64_t em = __ballot(1);
            const uint32_t leader = (uint32_t)__builtin_ctzll(em);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(sink.count, (uint32_t)__popcll(em));
            base = __builtin_amdgcn_readfirstlane(base);   // the leader is the first active lane
            const uint32_t idx =
                base + __builtin_amdgcn_mbcnt_hi((uint32_t)(em >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)em, 0u));
            if (idx < sink.cap) sink.out[idx] = r;
        } else {
            // convert_matches_myers' drop rule (search.rs:1338-1350) on the
            // chunk of the alignment start: the candidate's first chunk
            // (fields cached at refill) or, past its end, a later one
            const uint32_t ext = L + db - rb;
            bool end_chrom, past_end;
            if (st < c0_begin + ((c0_size + 1u) & ~1u)) {
                end_chrom = c0_endc;
                past_end = st - c0_begin + ext > c0_size;
            } else {
                uint32_t c = cand_chunk;
                while (c + 1 < sink.n_chunks && sink.chunks[c + 1].nib_begin <= st) c++;
                end_chrom = c + 1 == last_chunk || sink.chunks[c + 1].chrom_start;
                past_end = st - sink.chunks[c].nib_begin + ext > sink.chunks[c].size;
            }
            if (!(end_chrom && past_end)) {
                my_kept++;
                my_sum += bulge_hit_hash(st, hash_id, mm, db, rb, d, o0, o1);
            }
        }
    };
    // chunk fields of the candidate's first chunk for the drop rule
    auto load_c0 = [&]() {
        const DChunk C0 = sink.chunks[cand_chunk];
        c0_begin = C0.nib_begin;
        c0_size = C0.size;
        c0_endc = cand_chunk + 1 == last_chunk || sink.chunks[cand_chunk + 1].chrom_start;
    };

    for (;;) {
        // refill idle lanes from the queue, a wave at a time
        uint64_t idle = __ballot(!busy);
        uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!drained && n_idle >= sink.refill) {
            uint32_t b0 = 0;
            if (lane == 0) b0 = atomicAdd(q_next, n_idle);
            b0 = __builtin_amdgcn_readlane(b0, 0);
            if (b0 + n_idle >= n_queue) drained = true;
            if (!busy) {
                uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                uint32_t qi = b0 + rank;
                if (qi < n_queue) {
                    const BCand C = queue[qi];
                    const BPat& P = pats[C.pat];
                    cand_idx = qi;
                    cand_steps = 0;
                    L = P.L; M = P.M; D = P.D; R = P.R; E = P.E; pid = P.pid; hash_id = P.hash_id;
                    last_chunk = C.last_chunk;
                    if (sink.count_only) {
                        uint32_t c = sink.chunk_lut[C.j >> 16];
                        const uint64_t t0 = C.j + 1 >= C.floor + L + D ? C.j + 1 - (L + D) : C.floor;
                        while (c > 0 && sink.chunks[c].nib_begin > t0) c--;
                        while (c + 1 < sink.n_chunks && sink.chunks[c + 1].nib_begin <= t0) c++;
                        cand_chunk = c;
                        load_c0();
                    }
                    uint32_t W = L + D;
                    ts = C.j + 1 >= C.floor + W ? C.j + 1 - W : C.floor;
                    uint32_t n = (uint32_t)(C.j - ts + 1);
                    uint64_t w0 = ts >> 3;
                    uint32_t sh = (uint32_t)(ts & 7u) * 4u;
                    uint32_t X[sizeof(WORD) * 2];
                    uint32_t prev = genome[w0];
#pragma unroll
                    for (int k = 0; k < (int)sizeof(WORD) * 2; k++) {
                        uint32_t nx = genome[w0 + k + 1];
                        X[k] = __builtin_amdgcn_alignbit(nx, prev, sh);
                        prev = nx;
                    }
                    Masks<WORD> mk;
                    WORD tN, tA, tC, tG, tT;
                    window_masks<WORD>(X, n, mk, &z, &tN, &tA, &tC, &tG, &tT);
                    exact = has_exact && D <= 2u && R <= 2u && M <= 6u;
                    WORD VP = 0, VN = 0;
                    if (!exact) {
                        cvp[lane] = 0;
                        cvn[lane] = 0;
                        for (uint32_t w = 0, ii = 0; ii < L; w++) {
                            uint64_t cw = P.cls_bits[w];
                            for (uint32_t k = 0; k < 21 && ii < L; k++, ii++, cw >>= 3) {
                                myers_step<WORD>(class_mask<WORD>(mk, (uint32_t)(cw & 7u)), VP, VN);
                                cvp[(ii + 1) * 64 + lane] = VP;
                                cvn[(ii + 1) * 64 + lane] = VN;
                            }
                        }
                    }
                    {
                        const WORD pa = (WORD)P.pa, pc = (WORD)P.pc, pg = (WORD)P.pg, pt = (WORD)P.pt;
                        const WORD fa = (WORD)P.fa, fc = (WORD)P.fc, fg = (WORD)P.fg, ft = (WORD)P.ft;
                        const WORD pn = (WORD)P.is_n, fF = fa & fc & fg & ft;
                        const WORD lm = L >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << L) - 1);
                        rna_pm = ~pn & ~(pn >> 1) & ~(pn << 1) & lm;
                        dna_pm = ~(pn & (pn << 1)) & lm;
                        dg_base = (int)n - (int)L - (int)D;
                        constexpr int kBits = (int)sizeof(WORD) * 8;
                        for (uint32_t bb = 0; bb <= D + R; bb++) {
                            const int delta = dg_base + (int)bb;
                            auto shift = [&](WORD x) -> WORD {
                                return delta >= 0 ? (delta >= kBits ? (WORD)0 : x >> delta)
                                                  : (-delta >= kBits ? (WORD)0 : x << -delta);
                            };
                            // pattern positions whose text position lies in [0, n)
                            const int hi = (int)n - delta;
                            WORD valid = hi >= kBits ? ~(WORD)0 : hi <= 0 ? (WORD)0 : (((WORD)1 << hi) - 1);
                            if (delta < 0) valid &= -delta >= kBits ? (WORD)0 : ~(((WORD)1 << -delta) - 1);
                            const WORD nz = ~shift(z) & valid, re = nz & ~shift(tN);
                            const WORD sA = shift(tA), sC = shift(tC), sG = shift(tG), sT = shift(tT);
                            const WORD match = (pn & nz) | (~pn & ((pa & sA) | (pc & sC) | (pg & sG) | (pt & sT)) & re);
                            const WORD cell = (fF & nz) | (~fF & ((fa & sA) | (fc & sC) | (fg & sG) | (ft & sT)) & re);
                            const WORD mis = ~match & cell, ok = cell & ~(mis & pn);
                            if (D + R <= 4u) {
                                dk0 = bb == 0u ? ok : dk0; dk1 = bb == 1u ? ok : dk1; dk2 = bb == 2u ? ok : dk2;
                                dk3 = bb == 3u ? ok : dk3; dk4 = bb == 4u ? ok : dk4;
                                dm0 = bb == 0u ? mis : dm0; dm1 = bb == 1u ? mis : dm1; dm2 = bb == 2u ? mis : dm2;
                                dm3 = bb == 3u ? mis : dm3; dm4 = bb == 4u ? mis : dm4;
                            } else {
                                dgm[(2u * bb) * 64u + lane] = ok;
                                dgm[(2u * bb + 1u) * 64u + lane] = mis;
                            }
                        }
                    }
                    if (exact) exact_table<WORD>(xt, lane, L, D, R, dk0, dk1, dk2, dk3, dk4, dm0, dm1, dm2, dm3, dm4,
                                                 (WORD)P.is_n);
                    i = L; jj = n; d = 0; d_floor = 0; cost = mm = db = rb = 0;
                    // root admission: padding at the window end, then the bound
                    busy = !bit(z, n - 1) &&
                           (exact ? ((xt[L * 64 + lane] >> (3u * (3u * D + R))) & 7u) <= M : lbc(L, VP, VN, n) <= E);
                    if (busy) set4(0, children());
                    else if (sink.cand_steps) sink.cand_steps[qi] = 0;
                }
            }
            continue;
        }
        if (drained && __ballot(busy) == 0) break;
        if (drained && !sink.cand_steps) {
            // Work donation once the queue is empty: a busy lane with an
            // unexplored sibling above its current node hands the siblings
            // of its shallowest such depth, with the path to them and the
            // candidate's context, to an idle lane of the wave.  The DFS of
            // one long candidate then spreads over the wave instead of
            // leaving the other lanes idle.
            const uint64_t idle = __ballot(!busy);
            if ((uint32_t)__popcll(idle) >= sink.refill) {
                uint32_t kd = 0xFFFFFFFFu;
                if (busy) {
                    // pending siblings at depths [d_floor, d)
                    CM v = cm0 | cm1 | cm2;
                    if (d < kCmBits) v &= ((CM)1 << d) - 1;
                    v &= ~(((CM)1 << d_floor) - 1);
                    if (v) kd = (uint32_t)__builtin_ctzll((uint64_t)v);
                }
                const bool can = kd != 0xFFFFFFFFu;
                const uint64_t donors = __ballot(can);
                const uint32_t n_pairs = min((uint32_t)__popcll(donors), (uint32_t)__popcll(idle));
                if (n_pairs) {
                    const uint32_t rd = __builtin_amdgcn_mbcnt_hi((uint32_t)(donors >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)donors, 0u));
                    const uint32_t ri = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                    const bool give = can && rd < n_pairs;
                    const bool recv = !busy && ri < n_pairs;
                    if (give) pairing[rd] = lane;
                    __syncthreads();
                    const int src = recv ? (int)pairing[ri] : (int)lane;
                    __syncthreads();
                    // donor: node state at depth kd (undo the ops of depths [kd, d))
                    uint32_t s_i = i, s_jj = jj, s_cost = cost, s_mm = mm, s_db = db, s_rb = rb, s_mask = 0;
                    if (give) {
                        uint32_t nS, nD, nR;
                        op_counts(kd, d, nS, nD, nR);
                        const uint32_t nM = d - kd - nS - nD - nR;
                        s_i = i + nM + nS + nR;
                        s_jj = jj + nM + nS + nD;
                        s_cost = cost - nS - nD - nR;
                        s_mm = mm - nS;
                        s_db = db - nD;
                        s_rb = rb - nR;
                        s_mask = get4(kd);          // children left at kd, with the diagonal's mismatch bit
                        set4(kd, 0u);
                    }
                    // every lane takes part in the shuffles; receivers keep the values
                    auto xfer32 = [&](uint32_t v) -> uint32_t { return (uint32_t)__shfl((int)v, src); };
                    auto xfer64 = [&](uint64_t v) -> uint64_t {
                        const uint32_t lo = xfer32((uint32_t)v), hi = xfer32((uint32_t)(v >> 32));
                        return ((uint64_t)hi << 32) | lo;
                    };
                    auto xferw = [&](WORD v) -> WORD {
                        if constexpr (sizeof(WORD) == 8) return (WORD)xfer64((uint64_t)v);
                        else return (WORD)xfer32((uint32_t)v);
                    };
                    const uint32_t r_i = xfer32(s_i), r_jj = xfer32(s_jj), r_cost = xfer32(s_cost), r_mm = xfer32(s_mm),
                                   r_db = xfer32(s_db), r_rb = xfer32(s_rb), r_mask = xfer32(s_mask),
                                   r_kd = xfer32(kd);
                    const uint64_t r_ts = xfer64(ts), r_hash = xfer64(hash_id);
                    const WORD r_z = xferw(z), r_rpm = xferw(rna_pm), r_dpm = xferw(dna_pm);
                    const uint32_t r_exact = xfer32(exact ? 1u : 0u);
                    const WORD r_dk0 = xferw(dk0), r_dk1 = xferw(dk1), r_dk2 = xferw(dk2), r_dk3 = xferw(dk3),
                               r_dk4 = xferw(dk4), r_dm0 = xferw(dm0), r_dm1 = xferw(dm1), r_dm2 = xferw(dm2),
                               r_dm3 = xferw(dm3), r_dm4 = xferw(dm4);
                    const int r_dgb = (int)xfer32((uint32_t)dg_base);
                    const uint32_t r_L = xfer32(L), r_M = xfer32(M), r_D = xfer32(D), r_R = xfer32(R),
                                   r_E = xfer32(E), r_pid = xfer32(pid), r_last = xfer32(last_chunk),
                                   r_chunk = xfer32(cand_chunk);
                    uint64_t r_ops[kOpsW];
#pragma unroll
                    for (int w = 0; w < kOpsW; w++) r_ops[w] = xfer64(ops[w]);
                    // the candidate's Myers columns (rows 0..L) and diagonal masks
                    // (rows 0..2(D + R) + 1) into this lane's slots
                    if (r_exact) {
                        for (uint32_t r = 0; r <= r_L; r++) {
                            const uint32_t v = xt[r * 64 + (uint32_t)src];
                            if (recv) xt[r * 64 + lane] = v;
                        }
                    } else {
                        for (uint32_t r = 0; r <= r_L; r++) {
                            const WORD vp = cvp[r * 64 + (uint32_t)src], vn = cvn[r * 64 + (uint32_t)src];
                            if (recv) {
                                cvp[r * 64 + lane] = vp;
                                cvn[r * 64 + lane] = vn;
                            }
                        }
                    }
                    for (uint32_t r = 0; r_D + r_R > 4u && r < 2u * (r_D + r_R + 1u); r++) {
                        const WORD v = dgm[r * 64 + (uint32_t)src];
                        if (recv) dgm[r * 64 + lane] = v;
                    }
                    if (recv) {
                        i = r_i; jj = r_jj; cost = r_cost; mm = r_mm; db = r_db; rb = r_rb;
                        d = r_kd; d_floor = r_kd;
                        ts = r_ts; hash_id = r_hash;
                        z = r_z;
                        exact = r_exact != 0u;
                        dk0 = r_dk0; dk1 = r_dk1; dk2 = r_dk2; dk3 = r_dk3; dk4 = r_dk4;
                        dm0 = r_dm0; dm1 = r_dm1; dm2 = r_dm2; dm3 = r_dm3; dm4 = r_dm4;
                        rna_pm = r_rpm;
                        dna_pm = r_dpm;
                        dg_base = r_dgb;
                        L = r_L; M = r_M; D = r_D; R = r_R; E = r_E; pid = r_pid; last_chunk = r_last;
                        cand_chunk = r_chunk;
                        if (sink.count_only) load_c0();
#pragma unroll
                        for (int w = 0; w < kOpsW; w++) ops[w] = r_ops[w];
                        cm0 = cm1 = cm2 = cm3 = 0;
                        set4(d, r_mask);
                        busy = true;
                    }
                }
            }
        }
        my_slots++;
        if (!busy) continue;
        my_busy++;
        // One iteration = pop exhausted nodes (cheap, may repeat) and then one
        // push with its child expansion (the costly part), so the expansion
        // runs with every busy lane of the wave at once.
        // Backtrack in one step: the deepest depth in [d_floor, d] that still
        // has unexplored children is the highest set bit of the child planes;
        // the state there follows from the counts of the ops above it.
        CM pend = cm0 | cm1 | cm2;
        if (d + 1 < kCmBits) pend &= ((CM)2 << d) - 1;
        pend &= ~(((CM)1 << d_floor) - 1);
        if (!pend) {
            // subtree exhausted: candidate (or donated part) done
            my_steps += d - d_floor + 1u;
            cand_steps += d - d_floor + 1u;
            busy = false;
            if (sink.cand_steps) sink.cand_steps[cand_idx] = cand_steps;
            continue;
        }
        const uint32_t k = sizeof(CM) == 8 ? 63u - (uint32_t)__builtin_clzll((uint64_t)pend)
                                           : 31u - (uint32_t)__builtin_clz((uint32_t)pend);
        my_steps += d - k + 1u;
        cand_steps += d - k + 1u;
        if (k < d) {
            uint32_t nS, nD, nR;
            op_counts(k, d, nS, nD, nR);
            const uint32_t nM = d - k - nS - nD - nR;
            i += nM + nS + nR;
            jj += nM + nS + nD;
            cost -= nS + nD + nR;
            mm -= nS;
            db -= nD;
            rb -= nR;
            d = k;
        }
        const uint32_t node = get4(d);
        const uint32_t viable = node & 7u;
        // push the next admissible child (diagonal, then RNA, then DNA) and expand it
        const uint32_t t = (uint32_t)__builtin_ctz(viable);
        const uint32_t mcb = node >> 3;
        set4(d, (viable & (viable - 1)) | (node & 8u));
        const uint32_t op = t == 0u ? mcb : (t == 1u ? 3u : 2u);
        set2(d, op);
        i -= (uint32_t)(t != 2u);
        jj -= (uint32_t)(t != 1u);
        cost += t == 0u ? mcb : 1u;
        mm += (uint32_t)(t == 0u) & mcb;
        rb += (uint32_t)(t == 1u);
        db += (uint32_t)(t == 2u);
        d++;
        if (d > (uint32_t)sizeof(WORD) * 8 || i > L || jj > (uint32_t)sizeof(WORD) * 8) {
            // impossible state: stop this lane and report instead of looping
            atomicOr(sink.error, 1u);
            busy = false;
            continue;
        }
        if (i == 0) {
            emit();
            set4(d, 0u);
        } else if (db == D && rb == R) {
            // only the diagonal chain to i = 0 remains, on diagonal b = R
            WORD tail_ok, tail_mm;
            diag_of(R, tail_ok, tail_mm);
            const WORD pm = i >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << i) - 1);
            const WORD tm = tail_mm & pm;
            const uint32_t tmm = popc_w<WORD>(tm);
            if ((tail_ok & pm) == pm && mm + tmm <= M) {
                // ops of depths d .. d+i-1: the step at depth d + k is pattern
                // position i - 1 - k, op 1 where it is a mismatch, else 0
                put_tail_ops(ops, d, tm, i);
                const uint32_t d0 = d, jj0 = jj, mm0 = mm;
                d += i;
                jj -= i;
                mm += tmm;
                cost += tmm;
                emit();
                d = d0;
                jj = jj0;
                mm = mm0;
                cost -= tmm;
            }
            set4(d, 0u);
        } else {
            set4(d, children());
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        my_kept += __shfl_down(my_kept, off);
        my_sum += __shfl_down(my_sum, off);
        my_steps += __shfl_down(my_steps, off);
        my_slots += __shfl_down(my_slots, off);
        my_busy += __shfl_down(my_busy, off);
    }
    if (lane == 0) {
        if (sink.count_only && my_kept) {
            atomicAdd(sink.kept, (unsigned long long)my_kept);
            atomicAdd(sink.checksum, (unsigned long long)my_sum);
        }
        atomicAdd(&sink.diag[0], (unsigned long long)my_steps);
        atomicAdd(&sink.diag[1], (unsigned long long)my_slots);
        atomicAdd(&sink.diag[2], (unsigned long long)my_busy);
    }
}

// Budget screen between the filter and the traceback.  A Myers survivor has
// unit-cost edit distance <= E = M + D + R, but the traceback only emits
// alignments with at most M mismatches, D DNA and R RNA bulges whose diagonal
// steps pass the filter cells, and most survivors (94% on a repeat-rich
// genome with M4 D2 R2) have none: their DFS explores the whole tree the
// E bound admits and emits nothing.  This kernel decides, per candidate and
// exactly up to the padding rules, whether any admissible alignment exists,
// and only candidates that have one enter the traceback queue.
//
// One lane per candidate.  S[d][r][k] is the set of text ends j at which the
// pattern prefix of the current row i is alignable to a text suffix ending
// at j with <= d DNA bulges, <= r RNA bulges and <= k mismatches, under
// enumerate()'s step rules (traceback.rs:398-468): a diagonal step needs a
// filter-cell pass (search.rs:634-671) and no mismatch at an N position; an
// RNA bulge none of the N neighbours; a DNA bulge i < L and not two N
// neighbours.  Only ends from which the rest of the pattern can still reach
// the window end n with the bulges allowed matter: j - (n - (L - i) - D) in
// [0, D + R], a band of at most 5 positions.  In band coordinates a diagonal
// step keeps the position, an RNA bulge (next row, same j) lowers it by one
// and a DNA bulge (same row, next j) raises it by one, so for one (d, r) all
// mismatch budgets k fit one word as 5-bit fields:
//   S[d][r] = S'[d][r] & match | (S'[d][r] << 5) & mismatch           (S': row i-1)
//           | (S'[d][r-1] >> 1) & rna_ok | (S[d-1][r] << 1) & dna_ok
// with match / mismatch the band's admissible diagonal steps of row i in
// every field.  Row 0 (the empty prefix) holds every band position, which
// also admits starts before the window (a relaxation, as is leaving out the
// traceback's padding tests): the screen keeps every candidate the traceback
// could emit from.  A candidate passes when bit D of field M of S[D][R] is
// set after row L.  Patterns with D or R above 2 or M above KM are passed
// unchecked.
template <typename WORD, int KM>
__global__ __launch_bounds__(256) void bulge_screen_kernel(const uint32_t* __restrict__ genome,
                                                          const BCand* __restrict__ queue,
                                                          const uint32_t* __restrict__ q_count, uint32_t q_cap,
                                                          const BPat* __restrict__ pats, BCand* __restrict__ out,
                                                          uint32_t* __restrict__ out_count,
                                                          unsigned long long* __restrict__ n_out) {
    constexpr int DM = 2, RM = 2;
    // (KM + 1) fields of 5 band bits
    using PK = typename std::conditional<5 * (KM + 1) <= 32, uint32_t, uint64_t>::type;
    constexpr PK kRep = [] {
        PK r = 0;
        for (int k = 0; k <= KM; k++) r |= (PK)1 << (5 * k);
        return r;
    }();
    constexpr PK kAll = kRep * (PK)31;         // every field, every band bit
    constexpr PK kNotB0 = kAll & ~kRep;        // a field's bit 0 cleared (target of << 1)
    constexpr PK kNotB4 = kAll & ~(kRep << 4); // a field's bit 4 cleared (target of >> 1)
    const uint32_t n_queue = *q_count;
    if (n_queue > q_cap) {
        // overflowed round: hand the count on, the traceback flags it
        if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = n_queue;
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    // passing queue indices are staged per wave and copied out with one
    // reservation of the output count per batch (a reservation per wave and
    // step contends on the one counter)
    __shared__ uint32_t stage[4][kStage];
    uint32_t* my_stage = stage[threadIdx.x >> 6];
    uint32_t n_st = 0;   // wave-uniform
    auto flush = [&]() {
        wave_lds_sync();
        uint32_t b0 = 0;
        if (lane == 0) b0 = atomicAdd(out_count, n_st);
        b0 = __builtin_amdgcn_readlane(b0, 0);
        for (uint32_t i = lane; i < n_st; i += 64) out[b0 + i] = queue[my_stage[i]];
        n_st = 0;
        wave_lds_sync();
    };
    uint32_t kept = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n_queue; base += stride) {
        const uint32_t qi = base + threadIdx.x;
        const bool active = qi < n_queue;
        bool pass = false;
        if (active) {
            const BCand C = queue[qi];
            const BPat& P = pats[C.pat];
            const uint32_t L = P.L, M = P.M, D = P.D, R = P.R;
            if (D > (uint32_t)DM || R > (uint32_t)RM || M > (uint32_t)KM) {
                pass = true;
            } else {
                const uint32_t W = L + D;
                const uint64_t ts = C.j + 1 >= C.floor + W ? C.j + 1 - W : C.floor;
                const uint32_t n = (uint32_t)(C.j - ts + 1);
                const uint64_t w0 = ts >> 3;
                const uint32_t sh = (uint32_t)(ts & 7u) * 4u;
                uint32_t X[sizeof(WORD) * 2];
                uint32_t prev = genome[w0];
#pragma unroll
                for (int k = 0; k < (int)sizeof(WORD) * 2; k++) {
                    const uint32_t nx = genome[w0 + k + 1];
                    X[k] = __builtin_amdgcn_alignbit(nx, prev, sh);
                    prev = nx;
                }
                Masks<WORD> mk;
                WORD z, tN, tA, tC, tG, tT;
                window_masks<WORD>(X, n, mk, &z, &tN, &tA, &tC, &tG, &tT);
                const WORD keep = n >= sizeof(WORD) * 8 ? ~(WORD)0 : (((WORD)1 << n) - 1);
                const WORD pa = (WORD)P.pa, pc = (WORD)P.pc, pg = (WORD)P.pg, pt = (WORD)P.pt;
                const WORD fa = (WORD)P.fa, fc = (WORD)P.fc, fg = (WORD)P.fg, ft = (WORD)P.ft, pn = (WORD)P.is_n;
                const WORD fF = fa & fc & fg & ft;
                // admissible diagonal steps over pattern positions p on the
                // band's five diagonals: band position b at row i = p + 1 is
                // text position n - L - D + b + p
                WORD dmt[5], dms[5];
                constexpr int kBits = (int)sizeof(WORD) * 8;
#pragma unroll
                for (int bb = 0; bb < 5; bb++) {
                    const int delta = (int)n - (int)L - (int)D + bb;
                    auto shift = [&](WORD x) -> WORD {
                        return delta >= 0 ? (delta >= kBits ? (WORD)0 : x >> delta)
                                          : (-delta >= kBits ? (WORD)0 : x << -delta);
                    };
                    const int hi = (int)n - delta;          // text position < n
                    WORD valid = hi >= kBits ? ~(WORD)0 : hi <= 0 ? (WORD)0 : (((WORD)1 << hi) - 1);
                    if (delta < 0) valid &= -delta >= kBits ? (WORD)0 : ~(((WORD)1 << -delta) - 1);
                    const WORD nz = ~shift(z & keep) & valid, re = nz & ~shift(tN);
                    const WORD sA = shift(tA), sC = shift(tC), sG = shift(tG), sT = shift(tT);
                    const WORD match = (pn & nz) | (~pn & ((pa & sA) | (pc & sC) | (pg & sG) | (pt & sT)) & re);
                    const WORD cell = (fF & nz) | (~fF & ((fa & sA) | (fc & sC) | (fg & sG) | (ft & sT)) & re);
                    dmt[bb] = match & cell;
                    dms[bb] = ~match & cell & ~pn;
                }
                PK S[DM + 1][RM + 1];
#pragma unroll
                for (int d = 0; d <= DM; d++)
#pragma unroll
                    for (int r = 0; r <= RM; r++) S[d][r] = kAll;   // row 0: the empty prefix
                for (uint32_t i = 1; i <= L; i++) {
                    const uint32_t p = i - 1;
                    uint32_t bmt = 0, bms = 0;
#pragma unroll
                    for (int bb = 0; bb < 5; bb++) {
                        bmt |= (uint32_t)((dmt[bb] >> p) & 1) << bb;
                        bms |= (uint32_t)((dms[bb] >> p) & 1) << bb;
                    }
                    const PK mt = (PK)bmt * kRep, ms = (PK)bms * kRep;
                    const uint32_t n_p = (uint32_t)(pn >> p) & 1u;
                    const uint32_t n_next = i < L ? (uint32_t)(pn >> i) & 1u : 0u;
                    const uint32_t n_prev = i >= 2 ? (uint32_t)(pn >> (i - 2)) & 1u : 0u;
                    const PK rna = (n_p | n_next | n_prev) ? (PK)0 : kNotB4;
                    const PK dna = (i < L && !(n_p & n_next)) ? kNotB0 : (PK)0;
#pragma unroll
                    for (int d = 0; d <= DM; d++)
#pragma unroll
                        for (int r = RM; r >= 0; r--) {
                            PK v = (S[d][r] & mt) | ((S[d][r] << 5) & ms);
                            if (r > 0) v |= (S[d][r - 1] >> 1) & rna;
                            if (d > 0) v |= (S[d - 1][r] << 1) & dna;
                            S[d][r] = v & kAll;
                        }
                }
                PK fin = 0;
#pragma unroll
                for (int d = 0; d <= DM; d++)
#pragma unroll
                    for (int r = 0; r <= RM; r++)
                        if ((uint32_t)d == D && (uint32_t)r == R) fin = S[d][r];
                pass = ((fin >> (5u * M + D)) & 1) != 0;
            }
        }
        const uint64_t pb = __ballot(pass);
        if (pb) {
            const uint32_t off = __builtin_amdgcn_mbcnt_hi((uint32_t)(pb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pb, 0u));
            if (pass) my_stage[n_st + off] = qi;
            n_st += (uint32_t)__popcll(pb);
            kept += pass ? 1u : 0u;
            if (n_st > (uint32_t)kStage - 64u) flush();
        }
    }
    if (n_st) flush();
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_down(kept, off);
    if (lane == 0 && kept) atomicAdd(n_out, (unsigned long long)kept);
}

}  // namespace

struct BulgePlan {
    std::vector<BGroup> groups32, groups64;
    std::vector<BPat> pats;
    BGroup *d_g32 = nullptr, *d_g64 = nullptr;
    BPat* d_pats = nullptr;
    bool count_only = false;
    uint32_t n_cols = 0;          // Myers columns per lane in the verify kernel (max L + 1)
    uint32_t screen_km = 0;       // mismatch states of the budget screen (2, 4, 6; 0: no pattern screened)
    uint32_t n_diag = 1;          // traceback diagonal-mask pairs per lane (max D + R + 1)
    bool any_exact = false;       // a pattern with D, R <= 2 and M <= 6 (traceback feasibility table)
    bool any_myers = false;       // a pattern outside that range (traceback Myers bound)
};

void exact_kinds(const std::vector<BPat>& pats, const std::vector<BGroup>& g32, const std::vector<BGroup>& g64,
                 bool& any_exact, bool& any_myers) {
    any_exact = any_myers = false;
    for (const auto* gs : {&g32, &g64})
        for (const BGroup& G : *gs)
            for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++) {
                const BPat& P = pats[p];
                (P.D <= 2 && P.R <= 2 && P.M <= 6 ? any_exact : any_myers) = true;
            }
}

uint32_t max_diagonals(const std::vector<BPat>& pats, const std::vector<BGroup>& g32, const std::vector<BGroup>& g64) {
    uint32_t m = 1;
    for (const auto* gs : {&g32, &g64})
        for (const BGroup& G : *gs)
            for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++)
                m = std::max(m, pats[p].D + pats[p].R + 1);
    return m;
}

// Mismatch states the budget screen needs for the patterns of the given
// groups: the smallest of 2, 4, 6 covering their largest M (patterns with
// D or R above 2 or M above 6 are passed unchecked); 0 when none is screened.
uint32_t screen_states(const std::vector<BPat>& pats, const std::vector<BGroup>& g32, const std::vector<BGroup>& g64) {
    uint32_t m = 0;
    bool any = false;
    for (const auto* gs : {&g32, &g64})
        for (const BGroup& G : *gs)
            for (uint32_t p = G.pat_begin; p < G.pat_begin + G.pat_count; p++) {
                const BPat& P = pats[p];
                if (P.D > 2 || P.R > 2 || P.M > 6) continue;
                any = true;
                m = std::max(m, P.M);
            }
    if (!any) return 0;
    return m <= 2 ? 2u : m <= 4 ? 4u : 6u;
}

BulgePlan* bulge_plan_create(const Query& q) {
    auto* bp = new BulgePlan();
    bp->count_only = (q.flags & COF_FLAG_COUNT_ONLY) != 0;
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
    // 32-bit window groups first, then 64-bit ones, so the patterns of
    // consecutive groups of one class are consecutive in the table (the
    // filter scans runs of such groups together)
    for (int wide = 0; wide < 2; wide++)
    for (auto& kv : by_key) {
        if ((std::get<0>(kv.first) + std::get<1>(kv.first) > 32) != (wide == 1)) continue;
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
        if (kv.second.size() >= (1u << 20)) throw Error(COF_E_UNSUPPORTED, "more than 2^20 patterns share one filter");
        G.pat_count = (uint32_t)kv.second.size();
        for (uint32_t pi : kv.second) {
            const Pattern& p = q.pats[pi];
            BPat P{};
            for (uint32_t i = 0; i < p.L; i++) {
                char u = (char)toupper((unsigned char)p.ascii[i]);
                uint64_t cls = u == 'A' ? 0 : u == 'C' ? 1 : u == 'G' ? 2 : u == 'T' ? 3 : 4;
                P.cls_bits[i / 21] |= cls << (3 * (i % 21));
                uint8_t v = p.b4[i], f = kCodec.enc_mixed[(uint8_t)p.eff_filter[i]];
                P.pt |= (uint64_t)(v & 1) << i; P.pc |= (uint64_t)((v >> 1) & 1) << i;
                P.pa |= (uint64_t)((v >> 2) & 1) << i; P.pg |= (uint64_t)((v >> 3) & 1) << i;
                P.ft |= (uint64_t)(f & 1) << i; P.fc |= (uint64_t)((f >> 1) & 1) << i;
                P.fa |= (uint64_t)((f >> 2) & 1) << i; P.fg |= (uint64_t)((f >> 3) & 1) << i;
                if (is_n_char(p.ascii[i])) P.is_n |= 1ull << i;
                const uint64_t bi = 1ull << i;
                if (cls == 0 || cls == 4) P.peq[0] |= bi;
                if (cls == 1 || cls == 4) P.peq[1] |= bi;
                if (cls == 2 || cls == 4) P.peq[2] |= bi;
                if (cls == 3 || cls == 4) P.peq[3] |= bi;
                if (cls == 4) P.peq[4] |= bi;
            }
            P.L = p.L; P.M = p.M; P.D = p.D; P.R = p.R; P.E = p.M + p.D + p.R; P.pid = pi;
            P.hash_id = pattern_id_of(p);
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
    for (const auto& P : bp->pats) bp->n_cols = std::max(bp->n_cols, P.L + 1);
    bp->screen_km = screen_states(bp->pats, bp->groups32, bp->groups64);
    bp->n_diag = max_diagonals(bp->pats, bp->groups32, bp->groups64);
    exact_kinds(bp->pats, bp->groups32, bp->groups64, bp->any_exact, bp->any_myers);
    return bp;
}

void bulge_plan_destroy(BulgePlan* bp) {
    if (!bp) return;
    for (void* p : {(void*)bp->d_g32, (void*)bp->d_g64, (void*)bp->d_pats})
        if (p) (void)hipFree(p);
    delete bp;
}

size_t bulge_plan_size(const BulgePlan* bp) { return bp ? bp->pats.size() : 0; }

uint64_t bulge_pad_words() { return (uint64_t)kWinWords + 64; }


namespace {

void fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

struct RoundCtx {
    BulgePlan* bp;
    Workspace& ws;
    const Genome& g;
    const uint32_t* d_genome;
    uint32_t n_tiles;
    uint32_t verify_blocks;
    Sink base_sink;
    bool screen;              // budget screen before the traceback
};

// filter of tiles order[t0, t1) into `queue`, count at *q_count
void launch_filter(RoundCtx& c, const uint32_t* d_order, size_t t0, size_t t1, BCand* queue, uint32_t* q_count,
                   uint32_t q_cap, hipStream_t st) {
    uint32_t nt = (uint32_t)(t1 - t0);
    const BTile* tiles = (const BTile*)c.ws.tiles;
    uint32_t* tile_cand = c.ws.tile_cand;
    if (!c.bp->groups32.empty())
        hipLaunchKernelGGL(bulge_filter_kernel<uint32_t>, dim3(nt), dim3(kFilterThreads), 0, st, c.d_genome, tiles,
                           c.bp->d_g32, (uint32_t)c.bp->groups32.size(), c.bp->d_pats, queue, q_count, q_cap,
                           d_order + t0, tile_cand);
    if (!c.bp->groups64.empty())
        hipLaunchKernelGGL(bulge_filter_kernel<uint64_t>, dim3(nt), dim3(kFilterThreads), 0, st, c.d_genome, tiles,
                           c.bp->d_g64, (uint32_t)c.bp->groups64.size(), c.bp->d_pats, queue, q_count, q_cap,
                           d_order + t0, tile_cand);
}

// sums[5] += *q_count (candidates traced without the screen)
__global__ void add_count_kernel(const uint32_t* __restrict__ q_count, uint32_t q_cap, unsigned long long* __restrict__ dst) {
    const uint32_t n = *q_count;
    if (threadIdx.x == 0 && n <= q_cap) atomicAdd(dst, (unsigned long long)n);
}

template <typename WORD>
void launch_screen_w(RoundCtx& c, uint32_t km, const BCand* queue, const uint32_t* q_count, uint32_t q_cap,
                     BCand* out, uint32_t* out_count, uint32_t blocks, hipStream_t st) {
    unsigned long long* n_out = c.ws.sums + 5;
    if (km <= 2)
        hipLaunchKernelGGL((bulge_screen_kernel<WORD, 2>), dim3(blocks), dim3(256), 0, st, c.d_genome, queue, q_count,
                           q_cap, c.bp->d_pats, out, out_count, n_out);
    else if (km <= 4)
        hipLaunchKernelGGL((bulge_screen_kernel<WORD, 4>), dim3(blocks), dim3(256), 0, st, c.d_genome, queue, q_count,
                           q_cap, c.bp->d_pats, out, out_count, n_out);
    else
        hipLaunchKernelGGL((bulge_screen_kernel<WORD, 6>), dim3(blocks), dim3(256), 0, st, c.d_genome, queue, q_count,
                           q_cap, c.bp->d_pats, out, out_count, n_out);
}

// Traceback of one queue; with `steps` (q_cap entries) it also records the
// DFS transitions of each entry and adds them to the workspace's tile_steps.
// With the budget screen on, the queue is first screened into sq[slot] (count
// in screen_ctr[slot]), which the traceback then reads; `n_hint` (0: unknown)
// bounds the screen's grid.
void launch_verify(RoundCtx& c, const BCand* queue, const uint32_t* q_count, uint32_t q_cap, uint32_t* overflow,
                   uint32_t* q_next, uint32_t blocks, hipStream_t st, uint32_t* steps, int slot, uint32_t n_hint) {
    if (c.screen && c.bp->screen_km) {
        BCand* sq = (BCand*)c.ws.sq[slot];
        uint32_t* sc = c.ws.screen_ctr + slot;
        fill_u32(sc, 1, 0, st);
        const uint32_t n = n_hint ? n_hint : q_cap;
        const uint32_t sb = std::max(1u, std::min((n + 255) / 256, (uint32_t)std::max(1, c.ws.cus) * 8u));
        if (c.bp->groups64.empty()) launch_screen_w<uint32_t>(c, c.bp->screen_km, queue, q_count, q_cap, sq, sc, sb, st);
        else launch_screen_w<uint64_t>(c, c.bp->screen_km, queue, q_count, q_cap, sq, sc, sb, st);
        queue = sq;
        q_count = sc;
    } else {
        hipLaunchKernelGGL(add_count_kernel, dim3(1), dim3(64), 0, st, q_count, q_cap, c.ws.sums + 5);
    }
    // per lane: Myers columns (2 x n_cols rows of the window word) when some
    // pattern uses the Myers bound, feasibility rows (n_cols u32) when some
    // uses the exact table, diagonal masks (2 x n_diag rows) when some has
    // D + R > 4; 64 lanes
    const size_t wrows = (c.bp->any_myers ? 2 * (size_t)c.bp->n_cols : 0) + (c.bp->n_diag > 5 ? 2 * (size_t)c.bp->n_diag : 0);
    const size_t xbytes = c.bp->any_exact ? (size_t)c.bp->n_cols * 64 * sizeof(uint32_t) : 0;
    size_t lds32 = wrows * 64 * sizeof(uint32_t) + xbytes;
    size_t lds64 = wrows * 64 * sizeof(uint64_t) + xbytes;
    const uint32_t hm = c.bp->any_myers ? 1u : 0u, hx = c.bp->any_exact ? 1u : 0u;
    Sink sink = c.base_sink;
    sink.cand_steps = steps;
    if (c.bp->groups64.empty())
        hipLaunchKernelGGL(bulge_verify_kernel<uint32_t>, dim3(blocks), dim3(64), lds32, st, c.d_genome, queue, q_count,
                           q_cap, overflow, q_next, c.bp->d_pats, c.bp->n_cols, hm, hx, sink);
    else
        hipLaunchKernelGGL(bulge_verify_kernel<uint64_t>, dim3(blocks), dim3(64), lds64, st, c.d_genome, queue, q_count,
                           q_cap, overflow, q_next, c.bp->d_pats, c.bp->n_cols, hm, hx, sink);
    if (steps)
        hipLaunchKernelGGL(tile_steps_reduce_kernel, dim3((q_cap + 255) / 256), dim3(256), 0, st, queue, q_count, q_cap,
                           (const uint32_t*)steps, c.ws.tile_steps);
}

}  // namespace

// Runs the filter over every virtual batch's end range (tiles of 4 Ki end
// positions, search.rs:1080-1092, 899-927), then the traceback over the
// queued candidates.
//
// Schedule: a first round of tiles runs synchronously and measures the
// candidate density; the remaining tiles are cut into rounds expected to fill
// half a queue and run as a pipeline — filter of round r+1 on one stream while
// round r verifies on another, two queues alternating — captured into one
// hipGraph unless disabled.  The verify kernel reads its count from device
// memory and skips a round whose queue overflowed; such rounds are redone
// synchronously in halves (the reference bisects the same way,
// search.rs:1140-1152).  `order` (optional) is the tile processing order.
namespace {

// Myers columns (L + 1) up to which a pattern goes to the short 32-bit pass.
constexpr uint32_t kShortCols = 24;

// Adds src to dst (n words): per-tile candidate counts of two passes.
__global__ void add_u32_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

// The search for the plan's groups, all of one window class (32- or 64-bit
// windows).  The first pass of a search clears the record count, sums and
// per-tile step counts; a later pass adds to them.
void bulge_pass(BulgePlan* bp, Workspace& ws, const Genome& g, const std::vector<Batch>& batches,
                const uint32_t* d_genome, hipStream_t st, cof_stats* stats, const BulgeTune& tune, bool first_pass,
                uint32_t window_class) {
    const uint32_t flags = tune.flags;
    std::vector<BTile> tiles;
    for (const Batch& b : batches) {
        uint64_t floor = g.chunks[b.first].nib_begin();
        uint64_t a = g.chunks[b.first + (b.has_head ? 1 : 0)].nib_begin();
        uint64_t e = b.last < g.chunks.size() ? g.chunks[b.last].nib_begin() : g.n_nibbles();
        for (uint64_t j0 = a; j0 < e; j0 += kTileEnds)
            tiles.push_back(BTile{j0, floor, e, (uint32_t)std::min<uint64_t>(kTileEnds, e - j0), (uint32_t)b.last});
    }
    if (tiles.empty()) return;
    const uint32_t n_tiles = (uint32_t)tiles.size();
    ws.tile_j0.clear();
    ws.tile_n.clear();
    ws.tile_chunk.clear();
    ws.last_tiles = n_tiles;
    ws.tile_j0.resize(n_tiles);
    ws.tile_n.resize(n_tiles);
    ws.tile_chunk.resize(n_tiles);
    for (uint32_t t = 0; t < n_tiles; t++) {
        ws.tile_j0[t] = tiles[t].j0;
        ws.tile_n[t] = tiles[t].n;
        ws.tile_chunk[t] = (uint32_t)g.chunk_of(tiles[t].j0);
    }
    // Work weight per tile: its end positions, or with a cost model the
    // chunk's predicted traceback work shared out by end positions (+1 per
    // end for the filter).  With a model, tiles run heavy-first.
    std::vector<double> wt(n_tiles);
    const std::vector<float>* cc = tune.chunk_cost;
    const bool model = cc && cc->size() == g.chunks.size();
    for (uint32_t t = 0; t < n_tiles; t++) {
        wt[t] = tiles[t].n;
        if (model) {
            const uint32_t c = ws.tile_chunk[t];
            wt[t] += (double)std::max(0.f, (*cc)[c]) * tiles[t].n / (double)std::max<uint64_t>(1, g.chunks[c].size());
        }
    }
    std::vector<uint32_t> ord(n_tiles);
    for (uint32_t t = 0; t < n_tiles; t++) ord[t] = t;
    if (model) std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return wt[a] > wt[b]; });
    // The first (synchronous) round measures candidates per unit of weight:
    // it takes tiles evenly spaced over the order, so the sample spans the
    // genome (and, with a model, the whole cost range); the rest keep order.
    const size_t first = std::min<size_t>(n_tiles, std::max<uint32_t>(1, tune.first_tiles));
    if (first < n_tiles) {
        std::vector<uint8_t> picked(n_tiles, 0);
        std::vector<uint32_t> o2;
        o2.reserve(n_tiles);
        for (size_t i = 0; i < first; i++) {
            const size_t k = i * (size_t)n_tiles / first;
            picked[k] = 1;
            o2.push_back(ord[k]);
        }
        for (size_t k = 0; k < n_tiles; k++)
            if (!picked[k]) o2.push_back(ord[k]);
        ord.swap(o2);
    }
    ws_grow(&ws.tiles, &ws.tiles_cap, n_tiles, sizeof(BTile));
    ws_grow((void**)&ws.tile_order, &ws.tile_order_cap, n_tiles, sizeof(uint32_t));
    ws_grow((void**)&ws.tile_cand, &ws.tile_cand_cap, n_tiles, sizeof(uint32_t));
    ws_grow((void**)&ws.tile_steps, &ws.tile_steps_cap, n_tiles, sizeof(unsigned long long));
    HIPCHECK(hipMemcpyAsync(ws.tiles, tiles.data(), n_tiles * sizeof(BTile), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(ws.tile_order, ord.data(), n_tiles * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemsetAsync(ws.tile_cand, 0, n_tiles * sizeof(uint32_t), st));
    if (first_pass) {
        HIPCHECK(hipMemsetAsync(ws.tile_steps, 0, n_tiles * sizeof(unsigned long long), st));
        HIPCHECK(hipMemsetAsync(ws.sums, 0, 8 * sizeof(unsigned long long), st));
        HIPCHECK(hipMemsetAsync(ws.counters, 0, 4 * sizeof(uint32_t), st));
    }
    if (bp->count_only) {
        std::vector<DChunk> dc(g.chunks.size());
        for (size_t c = 0; c < g.chunks.size(); c++)
            dc[c] = DChunk{g.chunks[c].nib_begin(), (uint32_t)g.chunks[c].size(), g.chunks[c].start == 0 ? 1u : 0u};
        ws_grow(&ws.chunks, &ws.chunks_cap, dc.size(), sizeof(DChunk));
        HIPCHECK(hipMemcpyAsync(ws.chunks, dc.data(), dc.size() * sizeof(DChunk), hipMemcpyHostToDevice, st));
        // chunk holding nibble k << 16, for the device drop rule
        std::vector<uint32_t> lut((g.n_nibbles() >> 16) + 2);
        uint32_t c = 0;
        for (size_t k = 0; k < lut.size(); k++) {
            while (c + 1 < g.chunks.size() && g.chunks[c + 1].nib_begin() <= ((uint64_t)k << 16)) c++;
            lut[k] = c;
        }
        ws_grow((void**)&ws.chunk_lut, &ws.chunk_lut_cap, lut.size(), sizeof(uint32_t));
        HIPCHECK(hipMemcpyAsync(ws.chunk_lut, lut.data(), lut.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    const uint32_t cap_max = std::max<uint32_t>(64, tune.queue_cap);
    auto ensure_queue = [&](int k, uint32_t cap) {
        size_t c = ws.qs_cap[k];
        ws_grow(&ws.qs[k], &c, cap, sizeof(BCand));
        ws.qs_cap[k] = c;
    };
    if (!ws.bulge_out && !bp->count_only) {
        size_t cap = 0;
        ws_grow(&ws.bulge_out, &cap, 1u << 20, sizeof(BulgeRaw));
        ws.bulge_cap = (uint32_t)cap;
    }
    if (!ws.st2) HIPCHECK(hipStreamCreate(&ws.st2));
    if (!ws.st3) HIPCHECK(hipStreamCreate(&ws.st3));
    RoundCtx rc{bp, ws, g, d_genome, n_tiles, (uint32_t)std::max(1, ws.cus) * std::max(1u, tune.verify_per_cu),
                Sink{}, tune.screen != 0};
    auto ensure_screen = [&](int k, uint32_t cap) {
        if (!rc.screen || !bp->screen_km) return;
        ws_grow(&ws.sq[k], &ws.sq_cap[k], cap, sizeof(BCand));
    };
    rc.base_sink = Sink{(BulgeRaw*)ws.bulge_out, ws.counters + 1, ws.bulge_cap, bp->count_only ? 1 : 0,
                        (const DChunk*)ws.chunks, (uint32_t)g.chunks.size(), ws.sums, ws.sums + 1, ws.sums + 2,
                        ws.counters + 3, nullptr, ws.chunk_lut, std::max(1u, std::min(64u, tune.refill))};
    const bool want_steps = (flags & COF_FLAG_TILE_STATS) != 0;
    auto ensure_steps = [&](int k, uint32_t cap) {
        if (!want_steps) return;
        size_t c = ws.qsteps_cap[k];
        ws_grow((void**)&ws.qsteps[k], &c, cap, sizeof(uint32_t));
        ws.qsteps_cap[k] = c;
    };
    const uint32_t* d_order = ws.tile_order;
    auto t_start = std::chrono::steady_clock::now();

    uint64_t cand_sync = 0;
    // synchronous round over order[t0, t1) with queue 0 (capacity q_cap):
    // filter, count, traceback, timed with events.  Returns the count; a
    // count above q_cap means the traceback was not run.
    auto sync_round = [&](size_t t0, size_t t1, uint32_t q_cap) -> uint32_t {
        uint32_t* cnt = ws.counters + 2;       // queue count of this round
        HIPCHECK(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st));
        HIPCHECK(hipMemsetAsync(ws.round_ctr, 0, 2 * sizeof(uint32_t), st));   // cursor, overflow
        hipLaunchKernelGGL(clear_tile_cand_kernel, dim3((uint32_t)((t1 - t0 + 255) / 256)), dim3(256), 0, st,
                           d_order + t0, (uint32_t)(t1 - t0), ws.tile_cand);
        HIPCHECK(hipEventRecord(ws.ev[0], st));
        launch_filter(rc, d_order, t0, t1, (BCand*)ws.qs[0], cnt, q_cap, st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ws.ev[1], st));
        uint32_t nq = 0;
        HIPCHECK(hipMemcpyAsync(&nq, cnt, sizeof nq, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]));
        if (stats) stats->kernel_ms_filter += ms;
        if (nq > q_cap) return nq;
        cand_sync += nq;
        if (stats) stats->n_candidates += nq;
        if (nq) {
            HIPCHECK(hipEventRecord(ws.ev[2], st));
            ensure_screen(0, q_cap);
            launch_verify(rc, (const BCand*)ws.qs[0], cnt, q_cap, ws.round_ctr + 1, ws.round_ctr,
                          std::min<uint32_t>((nq + 63) / 64, rc.verify_blocks), st, want_steps ? ws.qsteps[0] : nullptr, 0,
                          nq);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipEventRecord(ws.ev[3], st));
            HIPCHECK(hipEventSynchronize(ws.ev[3]));
            HIPCHECK(hipEventElapsedTime(&ms, ws.ev[2], ws.ev[3]));
            if (stats) stats->kernel_ms_verify += ms;
        }
        return nq;
    };
    // synchronous rounds, halving a range whose queue overflowed.  With
    // COF_LOG_ROUNDS set, each completed round is reported on stderr (tile
    // range, candidates, time), so a search of many minutes shows progress.
    static const bool log_rounds = getenv("COF_LOG_ROUNDS") != nullptr;
    auto run_sync = [&](size_t t0, size_t t1, uint32_t q_cap) {
        std::vector<std::pair<size_t, size_t>> stack{{t0, t1}};
        while (!stack.empty()) {
            auto [a0, a1] = stack.back();
            stack.pop_back();
            const auto r0 = std::chrono::steady_clock::now();
            uint32_t nq = sync_round(a0, a1, q_cap);
            if (log_rounds)
                fprintf(stderr, "cof: bulge round tiles [%zu, %zu) of %zu: %u candidates, %.1f ms\n", a0, a1, (size_t)n_tiles,
                        nq, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count());
            if (nq > q_cap) {
                if (a1 - a0 == 1) throw Error(COF_E_NOMEM, "one filter tile produced more candidates than the queue holds");
                if (stats) stats->n_redo_rounds++;
                size_t mid = a0 + (a1 - a0) / 2;
                stack.push_back({mid, a1});
                stack.push_back({a0, mid});
            }
        }
    };

    // 1. first round: its filter runs synchronously and measures the
    //    candidate density; its traceback becomes round 0 of the pipeline
    //    (own queue, so the pipeline's queues can be sized afterwards)
    const uint32_t cap0 = std::min<uint32_t>(cap_max, 1u << 24);
    const bool sync_rounds = (flags & COF_FLAG_SYNC_ROUNDS) != 0;
    bool have_r0 = false;
    if (sync_rounds) {
        ensure_queue(0, cap0);
        ensure_steps(0, cap0);
        run_sync(0, first, cap0);
    } else {
        ws_grow(&ws.q_first, &ws.q_first_cap, cap0, sizeof(BCand));
        if (want_steps) ws_grow((void**)&ws.q_first_steps, &ws.q_first_steps_cap, cap0, sizeof(uint32_t));
        uint32_t* cnt = ws.counters + 2;
        HIPCHECK(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st));
        HIPCHECK(hipEventRecord(ws.ev[0], st));
        launch_filter(rc, d_order, 0, first, (BCand*)ws.q_first, cnt, cap0, st);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(ws.ev[1], st));
        uint32_t nq = 0;
        HIPCHECK(hipMemcpyAsync(&nq, cnt, sizeof nq, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]));
        if (stats) stats->kernel_ms_filter += ms;
        if (nq > cap0) {
            // dense start: synchronous rounds in halves
            if (stats) stats->n_redo_rounds++;
            ensure_queue(0, cap0);
            ensure_steps(0, cap0);
            run_sync(0, first, cap0);
        } else {
            cand_sync = nq;
            have_r0 = nq > 0;
            if (stats) stats->n_candidates += nq;
        }
    }
    if (first < n_tiles || have_r0) {
        // 2. the rest in rounds expected to fill half a queue, at least ~4
        //    rounds (so filter and traceback overlap) unless rounds would fall
        //    below 4 Mi candidates; the first round calibrates candidates per
        //    unit of weight
        double w_first = 0, w_rest = 0;
        for (size_t k = 0; k < first; k++) w_first += wt[ord[k]];
        for (size_t k = first; k < n_tiles; k++) w_rest += wt[ord[k]];
        const double per_w = std::max(1.0, (double)cand_sync) / std::max(1.0, w_first);
        const double target = std::min(0.5 * cap_max, std::max(per_w * w_rest / 4.0, (double)(1u << 22)));
        uint32_t q_cap = cap_max;
        if (!tune.round_tiles) {
            uint64_t want = 1;
            while (want < (uint64_t)(2.0 * target)) want <<= 1;
            q_cap = (uint32_t)std::min<uint64_t>(cap_max, want);
        }
        std::vector<std::pair<size_t, size_t>> rounds;
        if (have_r0) rounds.push_back({0, first});      // filtered already, queue q_first
        if (tune.round_tiles) {
            for (size_t t0 = first; t0 < n_tiles; t0 += tune.round_tiles)
                rounds.push_back({t0, std::min<size_t>(n_tiles, t0 + tune.round_tiles)});
        } else {
            size_t t0 = first;
            double acc = 0;
            for (size_t k = first; k < n_tiles; k++) {
                acc += wt[ord[k]] * per_w;
                if (acc >= target || k + 1 == n_tiles) {
                    rounds.push_back({t0, k + 1});
                    t0 = k + 1;
                    acc = 0;
                }
            }
        }
        const uint32_t R = (uint32_t)rounds.size();
        if (!sync_rounds) {
            // screened queues are sized before any capture begins
            const uint32_t scap = std::max(q_cap, have_r0 ? cap0 : 0u);
            ensure_screen(0, scap);
            ensure_screen(1, scap);
        }
        if (R > (have_r0 ? 1u : 0u) || sync_rounds) {
            ensure_queue(0, q_cap);
            ensure_steps(0, q_cap);
            if (!sync_rounds) {
                ensure_queue(1, q_cap);
                ensure_steps(1, q_cap);
            }
        }
        if (sync_rounds) {
            for (auto& r : rounds) run_sync(r.first, r.second, q_cap);
        } else {
            // 3. pipeline: filter of round r on `st`, its traceback on
            //    st2/st3 alternately (consecutive tracebacks overlap, so one
            //    round's tail does not idle the GPU).  Round 0 may be the
            //    first round (queue q_first, count in counters[2]); the other
            //    rounds alternate between the two queues, a queue being
            //    refilled only after the traceback that read it has ended.
            ws_grow((void**)&ws.rounds, &ws.rounds_cap, 3 * (size_t)R, sizeof(uint32_t));
            ws.ensure_round_events(R);
            const bool use_graph = !(flags & COF_FLAG_NO_GRAPH);
            hipGraph_t graph = nullptr;
            if (use_graph) HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
            fill_u32(ws.rounds, 3 * R, 0, st);
            uint32_t k_buf = 0;                   // pipeline rounds using qs[]
            std::vector<int> buf_of(R, -1);
            for (uint32_t r = 0; r < R; r++) {
                uint32_t* rq = ws.rounds + 3 * r;          // count, cursor, overflow
                hipStream_t vs = (r & 1) ? ws.st3 : ws.st2;
                const BCand* qv;
                const uint32_t* qcount;
                uint32_t cap_r;
                uint32_t* steps;
                if (have_r0 && r == 0) {
                    qv = (const BCand*)ws.q_first;
                    qcount = ws.counters + 2;
                    cap_r = cap0;
                    steps = want_steps ? ws.q_first_steps : nullptr;
                    HIPCHECK(hipEventRecord(ws.rev[0], st));
                } else {
                    const int buf = (int)(k_buf & 1);
                    buf_of[r] = buf;
                    // the queue's previous user is two pipeline rounds back
                    for (int q = (int)r - 1; q >= 0; q--) {
                        if (buf_of[q] == buf) {
                            HIPCHECK(hipStreamWaitEvent(st, ws.rev[2 * q + 1], 0));
                            break;
                        }
                    }
                    launch_filter(rc, d_order, rounds[r].first, rounds[r].second, (BCand*)ws.qs[buf], rq, q_cap, st);
                    HIPCHECK(hipEventRecord(ws.rev[2 * r], st));
                    qv = (const BCand*)ws.qs[buf];
                    qcount = rq;
                    cap_r = q_cap;
                    steps = want_steps ? ws.qsteps[buf] : nullptr;
                    k_buf++;
                }
                HIPCHECK(hipStreamWaitEvent(vs, ws.rev[2 * r], 0));
                launch_verify(rc, qv, qcount, cap_r, rq + 2, rq + 1, rc.verify_blocks, vs, steps, (int)(r & 1), 0);
                HIPCHECK(hipEventRecord(ws.rev[2 * r + 1], vs));
            }
            for (uint32_t r = R >= 2 ? R - 2 : 0; r < R; r++) HIPCHECK(hipStreamWaitEvent(st, ws.rev[2 * r + 1], 0));
            if (use_graph) {
                // The window class's executable graph from the last search is
                // updated in place from this capture when the pipeline has the
                // same shape (kernel arguments may differ) and replayed;
                // otherwise it is instantiated again.
                HIPCHECK(hipStreamEndCapture(st, &graph));
                hipGraphExec_t& exec = ws.bulge_exec[window_class];
                bool updated = false;
                if (exec) {
                    hipGraphNode_t err_node = nullptr;
                    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
                    updated = hipGraphExecUpdate(exec, graph, &err_node, &res) == hipSuccess &&
                              res == hipGraphExecUpdateSuccess;
                    if (!updated) {
                        (void)hipGetLastError();
                        HIPCHECK(hipGraphExecDestroy(exec));
                        exec = nullptr;
                    }
                }
                if (!exec) HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                if (stats) (updated ? stats->graph_updates : stats->graph_instantiations)++;
                HIPCHECK(hipGraphLaunch(exec, st));
                HIPCHECK(hipStreamSynchronize(st));
                HIPCHECK(hipGraphDestroy(graph));
            }
            HIPCHECK(hipStreamSynchronize(st));
            HIPCHECK(hipGetLastError());
            std::vector<uint32_t> rs(3 * (size_t)R);
            HIPCHECK(hipMemcpy(rs.data(), ws.rounds, rs.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (uint32_t r = (have_r0 ? 1u : 0u); r < R; r++) {
                if (rs[3 * r] > q_cap) {
                    // the traceback skipped this round; redo it (in halves if needed)
                    if (stats) stats->n_redo_rounds++;
                    run_sync(rounds[r].first, rounds[r].second, q_cap);
                } else if (stats) {
                    stats->n_candidates += rs[3 * r];
                }
            }
            HIPCHECK(hipStreamSynchronize(st));
            if (stats) stats->n_pipeline_rounds += R;
        }
    }
    if (stats) {
        stats->bulge_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
}

// The plan restricted to the groups of one window class whose Myers columns
// (L + 1) lie in [min_cols, max_cols): the same pattern table, and the class's
// device group table from the first kept group on.  Groups are ordered by L
// first (bulge_plan_create), so the kept groups are one contiguous range of it.
BulgePlan groups_view(const BulgePlan& bp, bool wide, uint32_t min_cols, uint32_t max_cols) {
    BulgePlan v;
    v.d_pats = bp.d_pats;
    v.count_only = bp.count_only;
    const std::vector<BGroup>& all = wide ? bp.groups64 : bp.groups32;
    std::vector<BGroup>& kept = wide ? v.groups64 : v.groups32;
    size_t first = all.size();
    for (size_t k = 0; k < all.size(); k++) {
        const BGroup& G = all[k];
        if (G.L + 1 < min_cols || G.L + 1 > max_cols) continue;
        if (kept.empty()) first = k;
        else if (k != first + kept.size()) throw Error(COF_E_HIP, "bulge pattern classes are not contiguous");
        kept.push_back(G);
        v.n_cols = std::max(v.n_cols, G.L + 1);
    }
    if (!kept.empty()) (wide ? v.d_g64 : v.d_g32) = (wide ? bp.d_g64 : bp.d_g32) + first;
    v.screen_km = screen_states(bp.pats, v.groups32, v.groups64);
    v.n_diag = max_diagonals(bp.pats, v.groups32, v.groups64);
    exact_kinds(bp.pats, v.groups32, v.groups64, v.any_exact, v.any_myers);
    return v;
}

}  // namespace

// Runs the filter over every virtual batch's end range and the traceback over
// the queued candidates (see bulge_pass), once per class of patterns present:
// 64-bit windows (L + D > 32), 32-bit windows with more than kShortCols
// Myers columns, and the rest.  A query mixing them runs each class with its
// own word and with LDS for its own longest pattern, rather than every
// candidate with the widest word and the longest pattern's columns (which
// cuts the traceback's resident workgroups per CU).
void bulge_run(BulgePlan* bp, Workspace& ws, const Genome& g, const std::vector<Batch>& batches,
               const uint32_t* d_genome, hipStream_t st, RawView<BulgeRaw>& out, cof_stats* stats,
               const BulgeTune& tune) {
    out = RawView<BulgeRaw>{};
    ws.tile_j0.clear();
    ws.tile_n.clear();
    ws.tile_chunk.clear();
    if (!bp || bp->pats.empty()) return;
    ws.ensure_events();
    std::vector<BulgePlan> views;
    std::vector<uint32_t> view_class;
    uint32_t cls = 0;
    for (auto [wide, lo, hi] : {std::make_tuple(false, 0u, kShortCols), std::make_tuple(false, kShortCols + 1, 0xFFFFu),
                                std::make_tuple(true, 0u, 0xFFFFu)}) {
        BulgePlan v = groups_view(*bp, wide, lo, hi);
        if (!v.groups32.empty() || !v.groups64.empty()) {
            views.push_back(std::move(v));
            view_class.push_back(cls);
        }
        cls++;
    }
    if (views.size() == 1) {
        bulge_pass(bp, ws, g, batches, d_genome, st, stats, tune, true, view_class[0]);
    } else {
        // per-tile candidate counts add up over the passes (each pass clears
        // and refills them)
        for (size_t k = 0; k < views.size(); k++) {
            bulge_pass(&views[k], ws, g, batches, d_genome, st, stats, tune, k == 0, view_class[k]);
            const uint32_t nt = ws.last_tiles;
            if (!nt) continue;
            if (k == 0) {
                ws_grow((void**)&ws.tile_cand_prev, &ws.tile_cand_prev_cap, nt, sizeof(uint32_t));
                HIPCHECK(hipMemcpyAsync(ws.tile_cand_prev, ws.tile_cand, nt * sizeof(uint32_t),
                                        hipMemcpyDeviceToDevice, st));
            } else {
                hipLaunchKernelGGL(add_u32_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, ws.tile_cand_prev,
                                   (const uint32_t*)ws.tile_cand, nt);
                HIPCHECK(hipGetLastError());
            }
            if (k + 1 == views.size())
                HIPCHECK(hipMemcpyAsync(ws.tile_cand, ws.tile_cand_prev, nt * sizeof(uint32_t),
                                        hipMemcpyDeviceToDevice, st));
        }
        HIPCHECK(hipStreamSynchronize(st));
    }
    uint32_t err = 0, no = 0;
    HIPCHECK(hipMemcpy(&err, ws.counters + 3, sizeof err, hipMemcpyDeviceToHost));
    if (err) throw Error(COF_E_HIP, "bulge traceback reached an impossible state");
    if (!bp->count_only) {
        HIPCHECK(hipMemcpy(&no, ws.counters + 1, sizeof no, hipMemcpyDeviceToHost));
        if (no > ws.bulge_cap) {
            // record buffer too small: grow it and search again
            size_t cap = ws.bulge_cap;
            ws_grow(&ws.bulge_out, &cap, (size_t)no + no / 4, sizeof(BulgeRaw));
            ws.bulge_cap = (uint32_t)cap;
            if (stats) stats->n_candidates = 0;
            bulge_run(bp, ws, g, batches, d_genome, st, out, stats, tune);
            return;
        }
        ws.ensure_host(&ws.h_bulge, &ws.h_bulge_cap, (size_t)no * sizeof(BulgeRaw));
        if (no) HIPCHECK(hipMemcpy(ws.h_bulge, ws.bulge_out, (size_t)no * sizeof(BulgeRaw), hipMemcpyDeviceToHost));
        out = RawView<BulgeRaw>{(const BulgeRaw*)ws.h_bulge, no};
    }
    if (stats) {
        unsigned long long sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        HIPCHECK(hipMemcpy(sums, ws.sums, sizeof sums, hipMemcpyDeviceToHost));
        stats->verify_steps = sums[2];
        stats->verify_slots = sums[3];
        stats->verify_busy = sums[4];
        stats->n_traced = sums[5];
        if (bp->count_only) {
            stats->bulge_hits = sums[0];
            stats->bulge_checksum = sums[1];
        }
    }
}

}  // namespace cof
