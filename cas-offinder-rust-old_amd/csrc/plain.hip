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
// Design (one workgroup = one tile of 64 x THREADS consecutive start
// positions, optionally one slice of a group's patterns per grid row):
//   1. the genome's scan image (built once per loaded genome) gives the
//      tile's flags (IUPAC genome letters, zero nibbles) and its two
//      32-position bit planes (base bit 1, base bit 0), which are all that
//      tiles of plain bases need.  Flagged tiles (and every tile of a launch
//      with a group that compares bit4 words) also load their bit4 words
//      (+ halo) with 16-byte loads into LDS, N-masked (0xF -> 0,
//      search.rs:512-531) where the tile holds zero nibbles;
//   2. per pattern group (patterns sharing one effective filter), every thread
//      tests the PAM of its 64 starts and the passing starts are compacted
//      into an LDS list of half a tile by a block-wide scan (two passes over
//      the tile's halves when they do not fit).  Plain-base tiles test a PAM
//      cell on 32 starts with three v_bitop3 selects over the planes and an
//      AND; other tiles test 8 starts per word (SWAR over nibbles);
//   3. groups whose patterns are plain bases on every position that is not N
//      in all of them (windows of at most 32 positions) score plain-base
//      tiles' windows on the matrix cores: 16 windows x 16 patterns per fp4
//      v_mfma_f32_16x16x128_f8f6f4 built from the planes, an exact
//      mismatch count (see mfma_chunk); with pattern slices that are not
//      whole blocks they compare the planes (xor + bitop3 + v_bcnt per
//      pattern, 16 patterns per batch of scalar loads).  Other tiles and
//      groups extract each window with v_alignbit funnel shifts, AND the bit4
//      words and v_bcnt them (W words per pattern);
//   4. hits (with the writer's mismatch recount) go to one of 64 record
//      segments, each with its own counter.
// Without IUPAC letters in the window, genome single bases make the writer's
// recount equal kmm minus the genome-N cells under pattern N, so the single
// popcount threshold is exact; IUPAC tiles take an exact per-nibble path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include "cof_internal.hpp"
#include "engine.hpp"
#include "hip_check.hpp"

// Device scratch of the allocation-free plain search (cof_search_static): the
// 64 hit-segment counters (kSegStride u32 apart), 64 segments of
// kStaticSegCap records and a packed copy of them.  One copy per GPU, held
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (plain_static_init).
constexpr uint32_t kStaticSegCap = 8192;
constexpr size_t kStaticCountBytes = 64 * 32 * sizeof(uint32_t);
constexpr size_t kStaticSegBytes = (size_t)64 * kStaticSegCap * sizeof(cof::PlainRaw);
static __device__ __align__(256) uint8_t scratch[kStaticCountBytes + 2 * kStaticSegBytes];

namespace cof {

namespace {

constexpr int kMaxThreads = 512;             // largest workgroup (tile = 64 starts per thread)
constexpr int kMaxL = 64;                    // plain pattern length supported
constexpr int kHaloWords = (kMaxL + 8) / 8 + 2;
constexpr int kMaxTileWords = kMaxThreads * 64 / 8 + kHaloWords;
constexpr int kMaxPam = 64;
// Hit records go to 64 segments, one per workgroup index mod 64, each with
// its own counter on its own 128-byte line: a single counter serialises the
// atomics of hit-dense queries (about one per 7 ns for the whole GPU).
constexpr uint32_t kHitSegs = 64;
constexpr uint32_t kSegStride = 32;          // u32 between segment counters

struct GroupDev {
    uint32_t pat_begin, pat_count;
    uint32_t L, k_lo, n_pam, W;
    // 2-bit plane compare: window from pk_lo, wmask = positions that are not
    // N in every pattern
    uint32_t planes, pk_lo, wmask, prec_begin;   // prec_begin: the group's first plane record
    uint32_t prec_padded;                        // the group's plane records with fillers (multiple of 16)
    // matrix-core filter (0: off) and the group's first operand fragment
    // (units of 16 bytes)
    uint32_t mfma, frag_begin, pad2;
    uint32_t pam_pos[kMaxPam];
    uint32_t pam_mask[kMaxPam];   // filter code in every nibble
};

// One pattern as two bit planes over the group's plane window (base T C A G =
// planes (0,0) (0,1) (1,0) (1,1), positions outside the window 0) and
// K = 64 - M: a window hits when popcount(mismatch) + K <= 64, and that
// popcount is the writer's recount.  A group's records are padded to a
// multiple of 16 with records that never hit (K = 200), so the compare runs in
// whole batches of 16 or 8; `pid` is the query pattern index.
struct __attribute__((aligned(16))) PlaneRec {
    uint32_t ph, pl;
    uint32_t K;
    uint32_t pid;
};

// One pattern of a group: its compare window (the pattern's nibbles from
// k_lo, with filter-fixed N positions zeroed), the popcount threshold
// L - M - (zeroed positions), M and the query pattern index.  64 bytes, so a
// wave fetches a record with one scalar load.
struct __attribute__((aligned(64))) PatRec {
    int32_t thr;          // first, so thr + the W compare words are one scalar load
    uint32_t w[8];
    uint32_t M;
    uint32_t pid;
    uint32_t pad[5];
};

// 4-bit nibble k (0..7) of word w
__device__ __forceinline__ uint32_t nib_of(uint32_t w, uint32_t k) { return (w >> (4 * k)) & 0xFu; }

// bit 3 of each nibble set iff the nibble is nonzero
__device__ __forceinline__ uint32_t nz_nibbles(uint32_t y) { return (((y & 0x77777777u) + 0x77777777u) | y) & 0x88888888u; }
// bit 3 of each nibble set iff the nibble has two or more bits set
__device__ __forceinline__ uint32_t multi_nibbles(uint32_t w) {
    uint32_t x = w & (w >> 1) & 0x55555555u;
    uint32_t o = (w | (w >> 1)) & 0x55555555u;
    return ((x | (x >> 2) | (o & (o >> 2))) & 0x11111111u) << 3;
}
// a0, a1 &= [the base at offset sh of plane words q0 / q1 is in the filter
// set `code`], 32 starts per word (base (h, l) = T C A G -> code bit 0 1 2 3).
// Branch-free: the set's bits as uniform masks, selected by l then by h
// (bitop3 table 0xCA: first operand ? second : third).
__device__ __forceinline__ void pam_cell(uint32_t& a0, uint32_t& a1, const uint2* gp, uint32_t q0, uint32_t q1,
                                         uint32_t sh, uint32_t code) {
    const uint2 x0 = gp[q0], x1 = gp[q0 + 1], y0 = gp[q1], y1 = gp[q1 + 1];
    const uint32_t h0 = __builtin_amdgcn_alignbit(x1.x, x0.x, sh);
    const uint32_t l0 = __builtin_amdgcn_alignbit(x1.y, x0.y, sh);
    const uint32_t h1 = __builtin_amdgcn_alignbit(y1.x, y0.x, sh);
    const uint32_t l1 = __builtin_amdgcn_alignbit(y1.y, y0.y, sh);
    const uint32_t cT = (code & 1u) ? ~0u : 0u, cC = (code & 2u) ? ~0u : 0u;
    const uint32_t cA = (code & 4u) ? ~0u : 0u, cG = (code & 8u) ? ~0u : 0u;
    a0 &= __builtin_amdgcn_bitop3_b32(h0, __builtin_amdgcn_bitop3_b32(l0, cG, cA, 0xCA),
                                      __builtin_amdgcn_bitop3_b32(l0, cC, cT, 0xCA), 0xCA);
    a1 &= __builtin_amdgcn_bitop3_b32(h1, __builtin_amdgcn_bitop3_b32(l1, cG, cA, 0xCA),
                                      __builtin_amdgcn_bitop3_b32(l1, cC, cT, 0xCA), 0xCA);
}

// Workgroup barrier for LDS exchange.  A one-wave workgroup needs no
// hardware barrier (a wave's LDS operations complete in issue order), only
// memory ordering at wavefront scope: the release fence keeps this wave's
// LDS writes before the barrier and the acquire fence its later reads after
// it, so lanes see each other's writes whatever the compiler can prove about
// aliasing.  At wavefront scope the fences emit no instructions.
template <int THREADS>
__device__ __forceinline__ void tile_sync() {
    if constexpr (THREADS == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active): DPP
// row shifts by 1, 2, 4 and 8 sum within each row of 16 lanes (a lane without
// a source adds 0), then row_bcast:15 adds lane 15 of rows 0 and 2 to rows 1
// and 3, and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_prefix_sum(uint32_t v) {
    // (bound_ctrl: a lane without a source reads 0, so each shift step is
    // one v_add_u32 with a DPP operand)
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return v;
}

// bit 0 of nibble j -> bit j (other bits zero)
__device__ __forceinline__ uint32_t gather_nibble_bit0(uint32_t m) {
    m = (m | (m >> 3)) & 0x03030303u;
    m = (m | (m >> 6)) & 0x000F000Fu;
    return (m | (m >> 12)) & 0xFFu;
}
// bit 3 of nibble j -> bit j
__device__ __forceinline__ uint32_t gather_nibble_flags(uint32_t f) { return gather_nibble_bit0(f >> 3); }

// Exact writer recount (main.rs:184-187) of one window: pattern positions
// that are not N and do not hold a plain base of the pattern's set, computed
// 8 positions per word op: a position counts when the
// pattern nibble is not N and the genome nibble is not a single base sharing a
// bit with it (N is already masked to 0 in the tile).
__device__ __forceinline__ uint32_t recount_swar(const uint32_t* gw, uint32_t i, const uint32_t* __restrict__ fw,
                                                 uint32_t L) {
    const uint32_t qw = i >> 3, sh = (i & 7u) * 4u;
    uint32_t amm = 0;
    for (uint32_t k = 0; k * 8 < L; k++) {
        const uint32_t g = __builtin_amdgcn_alignbit(gw[qw + k + 1], gw[qw + k], sh);
        const uint32_t p = fw[k];
        const uint32_t pn = (p & (p >> 1) & (p >> 2) & (p >> 3) & 0x11111111u) << 3;   // pattern N
        const uint32_t single = nz_nibbles(g) & ~multi_nibbles(g);
        const uint32_t ok = single & nz_nibbles(g & p);
        uint32_t miss = ~pn & ~ok & 0x88888888u;
        const uint32_t rem = L - k * 8;
        if (rem < 8) miss &= (1u << (4 * rem)) - 1u;
        amm += (uint32_t)__popc(miss);
    }
    return amm;
}

// Matrix-core score of plain-base windows (groups whose patterns are plain
// bases on every position of a window of at most 32 positions; tiles of plain
// bases only).  Per window position with planes (h, l) (T C A G = (0,0)
// (0,1) (1,0) (1,1)) and pattern base b, [genome == b] is linear in the
// features h, l and h AND l: T = 1 - h - l + hl, C = l - hl, A = h - hl,
// G = hl.  Features are fp4 (e2m1) 0.5 where the bit is set (nibble 0x1),
// weights are -1, 0 or +1 (nibbles 0xA, 0x0, 0x2), so a window scores
// 0.5 (matches - #T) over the window, and 32 bias slots (feature 0.5, fp4
// weights chosen on the host) add 0.5 (#T - Lw + M) + 0.25: the score is
// 0.5 (M - mismatches) + 0.25, exact in fp32, never zero, positive iff the
// window hits, and the hit's mismatch count (the writer's recount on a tile
// of plain bases) is M + 0.5 - 2 score.
//
// One v_mfma_f32_16x16x128_f8f6f4 (fp4 x fp4) scores 16
// windows (rows) against 16 patterns (columns) over 128 K nibbles.  Lane l
// holds row/column l % 16 and K nibbles 32 (l / 16) + 8 d + j (dword d,
// nibble j) of both operands; output v of lane l is row 4 (l / 16) + v,
// column l % 16.  Lane group c = l / 16 takes window positions c + 4 j:
// dwords 0-2 are the h, l and hl features ((plane >> c) & 0x11111111: one
// AND each after one funnel shift), dword 3 the bias feature.  Pattern
// fragments (fillers: weights 0, bias -96) are built on the host, 16 bytes
// per lane per block of 16 patterns.
typedef int32_t mf_v4i __attribute__((ext_vector_type(4)));
typedef int32_t mf_v8i __attribute__((ext_vector_type(8)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf_v4f mfma_fp4(const mf_v8i& a, const mf_v8i& b) {
    // formats 4/4: fp4 e2m1 operands (4 dwords each).  With zero scale
    // operands the compiler emits the unscaled v_mfma_f32_16x16x128_f8f6f4
    // (no scale registers, and the i8 MFMA's result latency; the scaled form
    // with scales of 2^0 took ~12 cycles longer per dependent use).
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, mf_v4f{0.f, 0.f, 0.f, 0.f}, 4, 4, 0, 0, 0, 0);
}

// One chunk of NB pattern blocks (held in registers) against every window of
// the list: per 16 windows, NB independent MFMAs and an AND of their score
// signs; the hit path runs only when some score is positive.
template <int NB, int THREADS>
__device__ __forceinline__ void mfma_chunk(const uint2* gp, const uint16_t* cand,
                                           uint32_t nc, const GroupDev& G, uint32_t bc,
                                           const uint4* __restrict__ mfrag,
                                           const PlaneRec* __restrict__ prec,
                                           uint64_t s0, PlainRaw* out, uint32_t* out_count, uint32_t out_cap,
                                           uint32_t lane, uint32_t wave) {
    const uint32_t lg = lane >> 4, r = lane & 15u;
    const uint32_t off = G.pk_lo + lg;   // plane position of this lane group's first window position
    mf_v4i bv[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint4 v = mfrag[G.frag_begin + (bc + b) * 64u + lane];
        bv[b] = mf_v4i{(int32_t)v.x, (int32_t)v.y, (int32_t)v.z, (int32_t)v.w};
    }
    // The fragments arrive here (an empty asm that reads them): otherwise the
    // loop waits for them on every iteration, and so also for the hit records
    // stored by the previous one (stores count in vmcnt).
#pragma unroll
    for (int b = 0; b < NB; b++) asm volatile("" ::"v"(bv[b]));
    // the list is padded to whole blocks; rows past the list are dropped in
    // the hit path
    // Software pipelined over the blocks of 16 list entries: an iteration
    // reads the plane words of the next block and the list entry of the one
    // after it, so no LDS latency lies between the compares of two blocks.
    // Reads past the list take entry cb instead (their values go unused).
    constexpr uint32_t kStep = 16u * (THREADS / 64);
    uint32_t cb = wave * 16u;
    if (cb >= nc) return;
    uint32_t q1 = (uint32_t)cand[cb + r] + off;   // plane position of this block's window
    uint2 p0 = gp[q1 >> 5], p1 = gp[(q1 >> 5) + 1];   // (hi, lo) plane words
    uint32_t q2 = (uint32_t)cand[(cb + kStep < nc ? cb + kStep : cb) + r] + off;
    for (; cb < nc; cb += kStep) {
        const uint32_t wh = __builtin_amdgcn_alignbit(p1.x, p0.x, q1 & 31u);
        const uint32_t wl = __builtin_amdgcn_alignbit(p1.y, p0.y, q1 & 31u);
        q1 = q2;
        p0 = gp[q1 >> 5];
        p1 = gp[(q1 >> 5) + 1];
        q2 = (uint32_t)cand[(cb + 2u * kStep < nc ? cb + 2u * kStep : cb) + r] + off;
        const mf_v4i af = {(int32_t)(wh & 0x11111111u), (int32_t)(wl & 0x11111111u),
                           (int32_t)__builtin_amdgcn_bitop3_b32(wh, wl, 0x11111111u, 0x80), 0x11111111};
        const mf_v8i a = {af.x, af.y, af.z, af.w, 0, 0, 0, 0};
        // sign bit of the AND of all scores: clear iff some score is positive
        uint32_t red = 0xFFFFFFFFu;
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const mf_v8i bf = {bv[b].x, bv[b].y, bv[b].z, bv[b].w, 0, 0, 0, 0};
            const mf_v4f d = mfma_fp4(a, bf);
            red = __builtin_amdgcn_bitop3_b32(red, __float_as_uint(d[0]), __float_as_uint(d[1]), 0x80);
            red = __builtin_amdgcn_bitop3_b32(red, __float_as_uint(d[2]), __float_as_uint(d[3]), 0x80);
        }
        if (!__ballot((int32_t)red >= 0)) continue;
        // window row 4 lg + v against pattern column r of block bc + b (scores
        // computed again, so they need no registers in the common case; blocks
        // without a positive score are skipped as a whole).  The features pass
        // through an empty asm so the compiler cannot reuse the scores above
        // (and keep all of them live across the loop).
        mf_v4i af2 = af;
        asm volatile("" : "+v"(af2));
        const mf_v8i a2 = {af2.x, af2.y, af2.z, af2.w, 0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const mf_v8i bf = {bv[b].x, bv[b].y, bv[b].z, bv[b].w, 0, 0, 0, 0};
            const mf_v4f d = mfma_fp4(a2, bf);
            if (!__ballot(max(max(d[0], d[1]), max(d[2], d[3])) > 0.f)) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const uint32_t c = cb + 4u * lg + (uint32_t)v;
                if (!(d[v] > 0.f) || c >= nc) continue;
                const PlaneRec& rec = prec[G.prec_begin + (bc + b) * 16u + r];
                const uint32_t amm = (uint32_t)((float)(64 - (int32_t)rec.K) + 0.5f - 2.f * d[v]);
                const uint32_t idx = atomicAdd(out_count, 1u);
                if (idx < out_cap) out[idx] = PlainRaw{s0 + cand[c], rec.pid, amm};
            }
        }
    }
}

// The group's pattern blocks [pb, pe) in chunks of up to 8 blocks.
template <int THREADS>
__device__ __forceinline__ void mfma_compare(const uint2* gp, const uint16_t* cand,
                                             uint32_t nc, const GroupDev& G, uint32_t pb, uint32_t pe,
                                             const uint4* __restrict__ mfrag,
                                             const PlaneRec* __restrict__ prec,
                                             uint64_t s0, PlainRaw* out, uint32_t* out_count, uint32_t out_cap,
                                             uint32_t lane, uint32_t wave) {
    const uint32_t blk_lo = (pb - G.pat_begin) >> 4;
    const uint32_t blk_hi = (pe - G.pat_begin + 15u) >> 4;
    for (uint32_t bc = blk_lo; bc < blk_hi; bc += 8) {
        const uint32_t nb = min(8u, blk_hi - bc);   // uniform
#define COF_MFMA_CHUNK(N)                                                                                    \
    case N:                                                                                                  \
        mfma_chunk<N, THREADS>(gp, cand, nc, G, bc, mfrag, prec, s0, out, out_count, out_cap, lane, wave); \
        break;
        switch (nb) {
            COF_MFMA_CHUNK(1) COF_MFMA_CHUNK(2) COF_MFMA_CHUNK(3) COF_MFMA_CHUNK(4)
            COF_MFMA_CHUNK(5) COF_MFMA_CHUNK(6) COF_MFMA_CHUNK(7) COF_MFMA_CHUNK(8)
            default: break;
        }
#undef COF_MFMA_CHUNK
    }
}

// One pattern group over one tile held in LDS (plane words gp, bit4 words gw
// for flagged tiles), with a window list of half a tile in cand: the PAM
// test, the list, and the compare, by a workgroup of THREADS threads
// (tid < THREADS; wave = tid / 64).  kPlainTile: the tile holds plain bases
// only (its planes serve).
template <int W, int THREADS, bool kPlainTile>
__device__ __forceinline__ void scan_group(
    const GroupDev& G, uint32_t pat_chunk, uint32_t chunk_row, const uint2* gp, const uint32_t* gw, uint16_t* cand,
    uint32_t* wave_sum, uint32_t tid, uint32_t lane, uint32_t wave, uint64_t s0, uint64_t scan_end, bool tail,
    int iupac, const PatRec* __restrict__ pats, const PlaneRec* __restrict__ prec,
    const uint32_t* __restrict__ full_words, const uint4* __restrict__ mfrag, PlainRaw* out,
    uint32_t* out_count, uint32_t out_cap) {
    constexpr int kThreads = THREADS;
    constexpr int kTile = THREADS * 64;
    uint32_t pb = G.pat_begin, pe = G.pat_begin + G.pat_count;
    if (pat_chunk) {
        const uint32_t lo = chunk_row * pat_chunk;   // uniform over the workgroup
        if (lo >= G.pat_count) return;
        pb += lo;
        pe = min(pe, pb + pat_chunk);
    }
    // 2. PAM test.  m8[w]: pass bits of this thread's starts in word w.
    uint32_t m8[8];
    if constexpr (kPlainTile) {
        // planes: per PAM cell and 32 starts, two bitop3 selects by the l
        // plane, one by the h plane and an AND (pam_cell).  The tile holding
        // scan_end is a flagged tile (the bit4 branch masks its starts), so
        // no start here lies past it.
        uint32_t ok[2] = {0xFFFFFFFFu, 0xFFFFFFFFu};
        for (uint32_t k = 0; k < G.n_pam; k++) {
            const uint32_t pk = G.pam_pos[k], code = G.pam_mask[k] & 0xFu;   // uniform
            const uint32_t qo = pk >> 5;
            pam_cell(ok[0], ok[1], gp, tid + qo, kThreads + tid + qo, pk & 31u, code);
        }
        m8[0] = ok[0];
        m8[1] = ok[1];
#pragma unroll
        for (int w = 2; w < 8; w++) m8[w] = 0;
    } else {
        // bit4, 8 positions per word: nibble j of X_k is the genome at
        // start 8t+j+k; with N masked to 0 and no IUPAC letters, a nibble
        // is in the filter's plain-base set iff it overlaps the filter code.
        uint32_t ok[8];
#pragma unroll
        for (int w = 0; w < 8; w++) ok[w] = 0x88888888u;
        for (uint32_t k = 0; k < G.n_pam; k++) {
            const uint32_t pk = G.pam_pos[k], pmask = G.pam_mask[k];   // uniform
            const uint32_t qo = pk >> 3, sh = (pk & 7u) * 4u;
#pragma unroll
            for (int w = 0; w < 8; w++) {
                const uint32_t q = (uint32_t)w * kThreads + tid + qo;
                const uint32_t X = __builtin_amdgcn_alignbit(gw[q + 1], gw[q], sh);
                const uint32_t bad = iupac ? multi_nibbles(X) : 0u;
                ok[w] &= nz_nibbles(X & pmask) & ~bad;
            }
        }
#pragma unroll
        for (int w = 0; w < 8; w++) m8[w] = gather_nibble_flags(ok[w]);
        if (tail) {
#pragma unroll
            for (int w = 0; w < 8; w++) {
                const uint64_t first = s0 + ((uint64_t)w * kThreads + tid) * 8u;
                if (scan_end <= first) m8[w] = 0;
                else if (scan_end < first + 8) m8[w] &= (1u << (uint32_t)(scan_end - first)) - 1u;
            }
        }
    }
    // block-wide exclusive scan of per-thread values: a DPP prefix sum in
    // the wave, the wave totals through LDS; tot = the block total.  The
    // sums stay below 2^16, so a value may pack two counts.
    auto block_scan = [&](uint32_t v, uint32_t& tot, uint32_t& excl) {
        const uint32_t inc = wave_prefix_sum(v);
        const uint32_t wtot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
        uint32_t off = 0;
        tot = wtot;
        if (kThreads > 64) {
            if (lane == 0) wave_sum[wave] = wtot;
            __syncthreads();
            tot = 0;
#pragma unroll
            for (int x = 0; x < kThreads / 64; x++) {
                uint32_t ws = wave_sum[x];
                off += (uint32_t)x < wave ? ws : 0u;
                tot += ws;
            }
            __syncthreads();   // wave_sum is written again by the next scan
        }
        excl = off + inc - v;
    };
    // The window list holds half a tile: the tile's windows are compared
    // in one pass when they fit, else in two passes over the halves of
    // its starts (plane word h of each thread, bit4 words 4h .. 4h + 3).
    // Plain-base tiles list their windows in start order (plane word 0 of
    // every thread, then word 1), so 16 consecutive entries lie in a few
    // consecutive LDS words: one scan of both words' counts, packed.
    uint32_t nc_all, pos_all = 0, tot0 = 0, ex0 = 0, ex1 = 0;
    if constexpr (kPlainTile) {
        uint32_t tot, ex;
        block_scan((uint32_t)__popc(m8[0]) | ((uint32_t)__popc(m8[1]) << 16), tot, ex);
        tot0 = tot & 0xFFFFu;
        nc_all = tot0 + (tot >> 16);
        ex0 = ex & 0xFFFFu;
        ex1 = ex >> 16;
    } else {
        uint32_t cnt_all = 0;
#pragma unroll
        for (int w = 0; w < 8; w++) cnt_all += (uint32_t)__popc(m8[w]);
        block_scan(cnt_all, nc_all, pos_all);
    }
    const uint32_t npass = nc_all <= (uint32_t)kTile / 2 ? 1u : 2u;   // uniform
    for (uint32_t h = 0; h < npass; h++) {
        uint32_t nc;
        if constexpr (kPlainTile) {
            // word 0's windows from 0, word 1's after them (one pass) or
            // from 0 (second pass)
            nc = npass == 1 ? nc_all : h == 0 ? tot0 : nc_all - tot0;
            uint32_t m = (npass == 1 || h == 0) ? m8[0] : 0u;
            uint32_t pos = ex0;
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctz(m);
                cand[pos++] = (uint16_t)(tid * 32u + j);
                m &= m - 1u;
            }
            m = (npass == 1 || h == 1) ? m8[1] : 0u;
            pos = (npass == 1 ? tot0 : 0u) + ex1;
            while (m) {
                const uint32_t j = (uint32_t)__builtin_ctz(m);
                cand[pos++] = (uint16_t)(((uint32_t)kThreads + tid) * 32u + j);
                m &= m - 1u;
            }
        } else {
            uint32_t mh[8];
#pragma unroll
            for (int w = 0; w < 8; w++) mh[w] = (npass == 1 || ((uint32_t)w >> 2) == h) ? m8[w] : 0u;
            uint32_t pos = pos_all;
            nc = nc_all;
            if (npass == 2) {
                uint32_t cnt = 0;
#pragma unroll
                for (int w = 0; w < 8; w++) cnt += (uint32_t)__popc(mh[w]);
                block_scan(cnt, nc, pos);
            }
#pragma unroll
            for (int w = 0; w < 8; w++) {
                uint32_t m = mh[w];
                while (m) {
                    uint32_t j = (uint32_t)__builtin_ctz(m);
                    cand[pos++] = (uint16_t)(((uint32_t)w * kThreads + tid) * 8u + j);
                    m &= m - 1u;
                }
            }
        }
        const bool planes = kPlainTile && G.planes;   // uniform
        const bool mfma = planes && G.mfma && (pat_chunk & 15u) == 0;
        // the matrix-core compare reads whole blocks of 16 list entries: pad
        // the list with start 0 (outputs of rows past the list are dropped)
        if (mfma && tid < ((16u - (nc & 15u)) & 15u)) cand[nc + tid] = 0;
        tile_sync<kThreads>();
        // 3. compare windows against the group's patterns
        if (mfma) {
            mfma_compare<kThreads>(gp, cand, nc, G, pb, pe, mfrag, prec, s0, out, out_count, out_cap, lane, wave);
            tile_sync<kThreads>();
            continue;
        }
        for (uint32_t c0 = wave * 64; c0 < nc; c0 += kThreads) {
            uint32_t c = c0 + lane;
            bool active = c < nc;
            uint32_t i = active ? cand[c] : 0;
            // bit4 compare of the lanes with `act`: W AND + v_bcnt per pattern
            auto bit4_compare = [&](bool act) {
                uint32_t q = i + G.k_lo;
                uint32_t wi = q >> 3, sh = (q & 7u) * 4u;
                uint32_t win[W];
#pragma unroll
                for (int w = 0; w < W; w++) win[w] = __builtin_amdgcn_alignbit(gw[wi + w + 1], gw[wi + w], sh);
                // four patterns per iteration: their records are fetched with
                // one batch of scalar loads (pointer offsets, no per-pattern
                // address arithmetic), and one branch covers the four hits
                uint32_t p = pb;
                for (; p + 4 <= pe; p += 4) {
                    const PatRec* const rp = pats + p;   // uniform base, immediate offsets below
                    uint32_t hit = 0;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        uint32_t sum = 0;
#pragma unroll
                        for (int w = 0; w < W; w++) sum += __popc(win[w] & rp[u].w[w]);
                        hit |= ((int32_t)sum >= rp[u].thr ? 1u : 0u) << u;
                    }
                    if (act && hit) {
                        for (int u = 0; u < 4; u++) {
                            if (!((hit >> u) & 1u)) continue;
                            const PatRec& r = pats[p + u];
                            uint32_t amm = recount_swar(gw, i, full_words + (size_t)(p + u) * 8, G.L);
                            uint32_t idx = atomicAdd(out_count, 1u);
                            if (idx < out_cap) out[idx] = PlainRaw{s0 + i, r.pid, amm};
                        }
                    }
                }
                for (; p < pe; p++) {
                    const PatRec& r = pats[p];
                    uint32_t sum = 0;
#pragma unroll
                    for (int w = 0; w < W; w++) sum += __popc(win[w] & r.w[w]);
                    if (act && (int32_t)sum >= r.thr) {
                        uint32_t amm = recount_swar(gw, i, full_words + (size_t)p * 8, G.L);
                        uint32_t idx = atomicAdd(out_count, 1u);
                        if (idx < out_cap) out[idx] = PlainRaw{s0 + i, r.pid, amm};
                    }
                }
            };
            if (planes) {
                // plain bases only: kmm = window positions whose planes differ
                // (positions that are N in every pattern match any base)
                const uint32_t q = i + G.pk_lo;
                const uint2 w0 = gp[q >> 5], w1 = gp[(q >> 5) + 1];
                const uint32_t gh = __builtin_amdgcn_alignbit(w1.x, w0.x, q & 31u) & G.wmask;
                const uint32_t gl = __builtin_amdgcn_alignbit(w1.y, w0.y, q & 31u) & G.wmask;
                const bool act = active;
                const uint64_t act_mask = __ballot(act);
                // plane records of patterns [pb, pe); a whole group runs on
                // its padded record count
                uint32_t p = G.prec_begin + (pb - G.pat_begin);
                const uint32_t qe = pat_chunk ? G.prec_begin + (pe - G.pat_begin)
                                              : G.prec_begin + ((G.pat_count + 7u) & ~7u);
                // NB patterns per batch: per pattern xor, bitop3, v_bcnt
                // accumulating K and half a v_min3, one compare per batch; all
                // NB records are loaded before the batch computes
                auto batch = [&](auto nb_tag) {
                    constexpr int NB = decltype(nb_tag)::value;
                    const PlaneRec* const rp = prec + p;   // uniform base
                    uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
                    for (int u = 0; u < NB; u += 2) {
                        const uint32_t t0 = (uint32_t)__popc((gh ^ rp[u].ph) | (gl ^ rp[u].pl)) + rp[u].K;
                        const uint32_t t1 = (uint32_t)__popc((gh ^ rp[u + 1].ph) | (gl ^ rp[u + 1].pl)) + rp[u + 1].K;
                        mn = min(mn, min(t0, t1));
                    }
                    if (__ballot(mn <= 64u) & act_mask) {
                        for (int u = 0; u < NB; u++) {
                            const uint32_t amm = (uint32_t)__popc((gh ^ rp[u].ph) | (gl ^ rp[u].pl));
                            if (!(act && amm + rp[u].K <= 64u)) continue;
                            uint32_t idx = atomicAdd(out_count, 1u);
                            if (idx < out_cap) out[idx] = PlainRaw{s0 + i, rp[u].pid, amm};
                        }
                    }
                    p += NB;
                };
                // batches of 16 while more than 8 records remain (a whole
                // group may run into its fillers, a pattern slice may not),
                // then at most one batch of 8, then single records
                const uint32_t q16 = pat_chunk ? qe : min(G.prec_begin + G.prec_padded, qe + 7u);
                while (p + 16 <= q16) batch(std::integral_constant<int, 16>{});
                if (p + 8 <= qe) batch(std::integral_constant<int, 8>{});
                for (; p < qe; p++) {
                    const PlaneRec& r = prec[p];
                    const uint32_t amm = (uint32_t)__popc((gh ^ r.ph) | (gl ^ r.pl));
                    if (act && amm + r.K <= 64u) {
                        uint32_t idx = atomicAdd(out_count, 1u);
                        if (idx < out_cap) out[idx] = PlainRaw{s0 + i, r.pid, amm};
                    }
                }
            } else if (!iupac) {
                bit4_compare(active);
            } else {
                for (uint32_t p = pb; p < pe; p++) {
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
                    uint32_t M = pats[p].M;
                    if (active && (int32_t)G.L - pop <= (int32_t)M && amm <= M) {
                        uint32_t idx = atomicAdd(out_count, 1u);
                        if (idx < out_cap) out[idx] = PlainRaw{s0 + i, pats[p].pid, amm};
                    }
                }
            }
        }
        tile_sync<kThreads>();
    }
}

// Workgroups go to the 8 XCDs round-robin (workgroup b on XCD b mod 8), and
// each XCD has its own L2.  Tile order: within every run of 512 workgroups,
// XCD x takes 64 consecutive tiles, so a tile's halo (the next tile's first
// words and plane words) is normally in its own XCD's L2 already.  A
// bijection on whole runs; the last partial run keeps the identity order.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    constexpr uint32_t kRun = 64, kGroup = 8 * kRun;
    if (b >= n / kGroup * kGroup) return b;
    const uint32_t g = b / kGroup, r = b % kGroup;
    return g * kGroup + (r & 7u) * kRun + (r >> 3);
}

template <int W, int THREADS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(6, 8))) void plain_scan_kernel(
    const uint32_t* __restrict__ genome, uint64_t scan_end, uint64_t first_tile_nib,
    const GroupDev* __restrict__ groups, uint32_t n_groups,
    const PatRec* __restrict__ pats,          // [pattern] compare words + threshold
    const PlaneRec* __restrict__ prec,        // [pattern] 2-bit planes (groups with planes)
    const uint32_t* __restrict__ full_words,  // [pattern][8] full pattern nibbles
    const uint4* __restrict__ mfrag,          // matrix-core pattern fragments
    const uint32_t* __restrict__ img_hi, const uint32_t* __restrict__ img_lo,   // scan image planes
    const uint8_t* __restrict__ img_flags,    // scan image flags per 4096-start unit
    PlainRaw* __restrict__ out_all, uint32_t* __restrict__ counts, uint32_t seg_cap,
    uint32_t pat_chunk,                       // patterns per grid row (0: whole group)
    uint32_t bit4_all) {                      // a group compares bit4 words: every tile is a flagged tile
    constexpr int kTile = THREADS * 64;
    constexpr int kTileWords = kTile / 8 + kHaloWords;
    constexpr int kPlaneWords = (kTileWords + 3) / 4 + 1;
    constexpr int kThreads = THREADS;
    // one LDS block, the plane words at offset 0: a window's or PAM cell's
    // four plane words are one ds_read2_b64 with no base to add.  A tile of
    // plain bases reads only its planes and a flagged tile only its bit4
    // words, so the two share their LDS: 6.2 KB per one-wave workgroup, which
    // with at most 80 VGPRs (waves_per_eu 6; a 16-byte spill) keeps six waves
    // per SIMD resident instead of five (+1.4% on the 3.1 Gbp headline).
    struct alignas(16) TileLds {
        union {
            uint2 gp[kPlaneWords];     // (hi, lo) plane words (tiles of plain bases)
            uint32_t gw[kTileWords];   // bit4 words (flagged tiles)
        } w;
        uint16_t cand[kTile / 2];   // one half tile's windows
        uint32_t wave_sum[kThreads / 64];
    };
    __shared__ TileLds lds;
    uint32_t* const gw = lds.w.gw;
    uint2* const gp = lds.w.gp;
    uint16_t* const cand = lds.cand;
    uint32_t* const wave_sum = lds.wave_sum;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: scalar loops over waves
    const uint64_t s0 = first_tile_nib + (uint64_t)xcd_tile(blockIdx.x, gridDim.x) * kTile;
    const uint64_t w0 = s0 >> 3;
    // this workgroup's hit segment
    uint32_t* const out_count = counts + (blockIdx.x & (kHitSegs - 1)) * kSegStride;
    PlainRaw* const out = out_all + (size_t)(blockIdx.x & (kHitSegs - 1)) * seg_cap;
    const uint32_t out_cap = seg_cap;

    // 1. genome tile -> LDS.  The genome's scan image gives the tile's two
    //    32-position bit planes (base bit 1 = A/G, base bit 0 = C/G), which
    //    serve tiles of plain bases only, and its flags (IUPAC letters; zero
    //    nibbles: masked N, padding) as the OR of its 4096-start units.  The
    //    plane loads are issued before the flags arrive.  Flagged tiles
    //    (every tile when a group of the launch compares bit4 words:
    //    bit4_all) load their bit4 words instead, N-masked (0xF -> 0,
    //    search.rs:512-531) where the tile has zero nibbles; a tile without
    //    them holds no N.  Every load of a thread is issued before any is
    //    used.
    constexpr int kPlaneIt = (kPlaneWords + kThreads - 1) / kThreads;
    uint32_t ph[kPlaneIt], pl[kPlaneIt];
#pragma unroll
    for (int it = 0; it < kPlaneIt; it++) {
        const uint32_t j = tid + (uint32_t)it * kThreads;
        const bool in = j < (uint32_t)kPlaneWords;
        ph[it] = in ? img_hi[(w0 >> 2) + j] : 0u;
        pl[it] = in ? img_lo[(w0 >> 2) + j] : 0u;
    }
    // only the tile holding scan_end masks starts past it; it runs as a
    // flagged tile (bit 1: the bit4 branch), whose PAM test does the masking.
    // With bit4_all every tile is flagged: plane groups then compare on the
    // bit4 words too (same results), which frees the planes' LDS for them.
    const bool tail = s0 + (uint64_t)kTile > scan_end;   // uniform
    int flags = tail || bit4_all ? 2 : 0;
#pragma unroll
    for (int u = 0; u < kThreads / 64; u++) flags |= img_flags[(s0 >> 12) + u];   // uniform
    if (bit4_all || flags) {
        constexpr int kGroups4 = (kTileWords + 3) / 4;
        constexpr int kLoadIt = (kGroups4 + kThreads - 1) / kThreads;
        const uint4* const g4 = reinterpret_cast<const uint4*>(genome + w0);
        uint4 lv[kLoadIt];
#pragma unroll
        for (int it = 0; it < kLoadIt; it++) {
            const uint32_t j = tid + (uint32_t)it * kThreads;
            lv[it] = j < (uint32_t)kGroups4 ? g4[j] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int it = 0; it < kLoadIt; it++) {
            const uint32_t j = tid + (uint32_t)it * kThreads;
            uint32_t raw[4] = {lv[it].x, lv[it].y, lv[it].z, lv[it].w};
            if (flags) {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const uint32_t f = raw[b] & (raw[b] >> 1) & (raw[b] >> 2) & (raw[b] >> 3) & 0x11111111u;   // nibble == 0xF
                    raw[b] &= ~((f << 4) - f);
                }
            }
            if (4 * j + 3 < (uint32_t)kTileWords) {   // a whole 16-byte group: one LDS store
                *reinterpret_cast<uint4*>(&gw[4 * j]) = make_uint4(raw[0], raw[1], raw[2], raw[3]);
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++)
                    if (4 * j + b < (uint32_t)kTileWords) gw[4 * j + b] = raw[b];
            }
        }
    }
    if (flags == 0) {
#pragma unroll
        for (int it = 0; it < kPlaneIt; it++) {
            const uint32_t j = tid + (uint32_t)it * kThreads;
            if (j < (uint32_t)kPlaneWords) {
                gp[j] = make_uint2(ph[it], pl[it]);
            }
        }
    }
    tile_sync<kThreads>();
    const int iupac = flags & 1;

    // Bit4 PAM test: thread tid owns tile words w * kThreads + tid (w < 8), 8
    // starts each; plane PAM test: plane words w * kThreads + tid (w < 2), 32
    // starts each.  Consecutive lanes read consecutive LDS words (no bank
    // conflicts).

    // The group loop, compiled once for tiles of plain bases (plane PAM test,
    // matrix-core or plane compare) and once for flagged tiles (bit4 words),
    // so neither copy carries the other's set-up.
    if (flags == 0) {
        for (uint32_t gi = 0; gi < n_groups; gi++)
            scan_group<W, THREADS, true>(groups[gi], pat_chunk, blockIdx.y, gp, gw, cand, wave_sum, tid, lane, wave,
                                         s0, scan_end, tail, iupac, pats, prec, full_words, mfrag, out, out_count,
                                         out_cap);
    } else {
        for (uint32_t gi = 0; gi < n_groups; gi++)
            scan_group<W, THREADS, false>(groups[gi], pat_chunk, blockIdx.y, gp, gw, cand, wave_sum, tid, lane, wave,
                                          s0, scan_end, tail, iupac, pats, prec, full_words, mfrag, out, out_count,
                                          out_cap);
    }
}

// Persistent variant of the scan: workgroups of two waves stay resident and
// claim runs of 4,096-start tiles from 8 queues, one per XCD, each owning a
// contiguous eighth of the tiles (a workgroup drains its own XCD's range
// first, so consecutive tiles and their halos share that XCD's L2, then
// the others').  The two waves load a tile's planes (and bit4 words when
// needed) together and then scan the groups alternately, each with its own
// window list.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcc_id() {
    // HW_REG_XCC_ID (hardware register 20), bits 3:0: the XCD this wave runs on
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kXcds - 1);
}

// Tiles a workgroup claims at once (one queue atomic per claim).
constexpr uint32_t kClaimTiles = 16;

// The workgroup's next run of up to kClaimTiles tiles [first, end), or
// first == n when every queue is drained.
__device__ __forceinline__ uint2 claim_tiles(uint32_t* heads, uint32_t n) {
    const uint32_t x0 = xcc_id();
    for (uint32_t k = 0; k < kXcds; k++) {
        const uint32_t x = (x0 + k) & (kXcds - 1);
        const uint32_t lo = (uint32_t)((uint64_t)n * x / kXcds), hi = (uint32_t)((uint64_t)n * (x + 1) / kXcds);
        uint32_t* const h = heads + x * kSegStride;
        if (lo >= hi || __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi - lo) continue;
        const uint32_t i = atomicAdd(h, kClaimTiles);
        if (i < hi - lo) return make_uint2(lo + i, lo + min(i + kClaimTiles, hi - lo));
    }
    return make_uint2(n, n);
}

template <int W>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(5, 8))) void plain_queue_kernel(
    const uint32_t* __restrict__ genome, uint64_t scan_end, uint32_t n_tiles, const GroupDev* __restrict__ groups,
    uint32_t n_groups, const PatRec* __restrict__ pats, const PlaneRec* __restrict__ prec,
    const uint32_t* __restrict__ full_words, const uint4* __restrict__ mfrag, const uint32_t* __restrict__ img_hi,
    const uint32_t* __restrict__ img_lo, const uint8_t* __restrict__ img_flags, PlainRaw* __restrict__ out_all,
    uint32_t* __restrict__ counts, uint32_t seg_cap, uint32_t bit4_all, uint32_t* __restrict__ heads) {
    constexpr int kThreads = 128;
    constexpr int kTile = 64 * 64;
    constexpr int kTileWords = kTile / 8 + kHaloWords;
    constexpr int kPlaneWords = (kTileWords + 3) / 4 + 1;
    struct alignas(16) QueueLds {
        uint2 gp[kPlaneWords];          // (hi, lo) plane words at offset 0
        uint32_t gw[kTileWords];        // bit4 words (flagged tiles, bit4 groups)
        uint16_t cand[2][kTile / 2];    // each wave's window list
        uint32_t tile, tile_end;        // the claimed run of tiles
    };
    __shared__ QueueLds lds;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint32_t* const out_count = counts + (blockIdx.x & (kHitSegs - 1)) * kSegStride;
    PlainRaw* const out = out_all + (size_t)(blockIdx.x & (kHitSegs - 1)) * seg_cap;
    uint16_t* const cand = lds.cand[wave];
    uint32_t t = 0, t_end = 0;   // uniform: the tiles claimed and not yet scanned
    for (;;) {
        if (t == t_end) {
            if (tid == 0) {
                const uint2 c = claim_tiles(heads, n_tiles);
                lds.tile = c.x;
                lds.tile_end = c.y;
            }
            __syncthreads();
            t = __builtin_amdgcn_readfirstlane(lds.tile);
            t_end = __builtin_amdgcn_readfirstlane(lds.tile_end);
            __syncthreads();   // read before the next claim writes
        }
        if (t >= n_tiles) break;
        const uint64_t s0 = (uint64_t)t * kTile;
        const uint64_t w0 = s0 >> 3;
        // the tile's planes and flags (and bit4 words where needed), as in
        // plain_scan_kernel
        constexpr int kPlaneIt = (kPlaneWords + kThreads - 1) / kThreads;
        uint32_t ph[kPlaneIt], pl[kPlaneIt];
#pragma unroll
        for (int it = 0; it < kPlaneIt; it++) {
            const uint32_t j = tid + (uint32_t)it * kThreads;
            const bool in = j < (uint32_t)kPlaneWords;
            ph[it] = in ? img_hi[(w0 >> 2) + j] : 0u;
            pl[it] = in ? img_lo[(w0 >> 2) + j] : 0u;
        }
        // the tile holding scan_end runs as a flagged tile, as in plain_scan_kernel
        const bool tail = s0 + (uint64_t)kTile > scan_end;   // uniform
        const int flags = img_flags[s0 >> 12] | (tail ? 2 : 0);   // uniform
        if (bit4_all || flags) {
            constexpr int kGroups4 = (kTileWords + 3) / 4;
            constexpr int kLoadIt = (kGroups4 + kThreads - 1) / kThreads;
            const uint4* const g4 = reinterpret_cast<const uint4*>(genome + w0);
            uint4 lv[kLoadIt];
#pragma unroll
            for (int it = 0; it < kLoadIt; it++) {
                const uint32_t j = tid + (uint32_t)it * kThreads;
                lv[it] = j < (uint32_t)kGroups4 ? g4[j] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int it = 0; it < kLoadIt; it++) {
                const uint32_t j = tid + (uint32_t)it * kThreads;
                uint32_t raw[4] = {lv[it].x, lv[it].y, lv[it].z, lv[it].w};
                if (flags) {
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const uint32_t f = raw[b] & (raw[b] >> 1) & (raw[b] >> 2) & (raw[b] >> 3) & 0x11111111u;
                        raw[b] &= ~((f << 4) - f);
                    }
                }
                if (4 * j + 3 < (uint32_t)kTileWords) {
                    *reinterpret_cast<uint4*>(&lds.gw[4 * j]) = make_uint4(raw[0], raw[1], raw[2], raw[3]);
                } else {
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        if (4 * j + b < (uint32_t)kTileWords) lds.gw[4 * j + b] = raw[b];
                }
            }
        }
        if (flags == 0) {
#pragma unroll
            for (int it = 0; it < kPlaneIt; it++) {
                const uint32_t j = tid + (uint32_t)it * kThreads;
                if (j < (uint32_t)kPlaneWords) lds.gp[j] = make_uint2(ph[it], pl[it]);
            }
        }
        __syncthreads();
        const int iupac = flags & 1;
        if (flags == 0) {
            for (uint32_t gi = wave; gi < n_groups; gi += 2)
                scan_group<W, 64, true>(groups[gi], 0, 0, lds.gp, lds.gw, cand, nullptr, lane, lane, 0, s0, scan_end,
                                        tail, iupac, pats, prec, full_words, mfrag, out, out_count, seg_cap);
        } else {
            for (uint32_t gi = wave; gi < n_groups; gi += 2)
                scan_group<W, 64, false>(groups[gi], 0, 0, lds.gp, lds.gw, cand, nullptr, lane, lane, 0, s0, scan_end,
                                         tail, iupac, pats, prec, full_words, mfrag, out, out_count, seg_cap);
        }
        __syncthreads();   // both waves are done with the tile before the next one loads
        t++;
    }
}

// Scan image of the genome, one wave per 4096-start unit u: the flags of
// words [512 u, 512 u + 512 + halo) after the N-mask (bit 0: an IUPAC letter,
// i.e. a nibble with two or more bits; bit 1: a zero nibble, i.e. masked N or
// padding) and the two bit planes of plane words [128 u, 128 u + 128)
// (positions 32 j .. 32 j + 31; base bit 1 = A/G, base bit 0 = C/G).  A scan
// tile of k units takes the OR of their flags, which covers its words and
// its halo.
__global__ __launch_bounds__(64) void scan_image_kernel(const uint32_t* __restrict__ genome, uint32_t* __restrict__ hi,
                                                         uint32_t* __restrict__ lo, uint8_t* __restrict__ flags_out) {
    constexpr uint32_t kUnitWords = 512;
    constexpr uint32_t kFlagWords = kUnitWords + kHaloWords;
    const uint64_t u = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint4* const g4 = reinterpret_cast<const uint4*>(genome + u * kUnitWords);
    int fl = 0;
    for (uint32_t j = tid; j < (kFlagWords + 3) / 4; j += 64) {
        const uint4 v = g4[j];
        const uint32_t raw[4] = {v.x, v.y, v.z, v.w};
        uint32_t h = 0, l = 0;
#pragma unroll
        for (uint32_t b = 0; b < 4; b++) {
            uint32_t w = raw[b];
            const uint32_t f = w & (w >> 1) & (w >> 2) & (w >> 3) & 0x11111111u;   // nibble == 0xF
            w &= ~(f * 0xFu);
            if (4 * j + b < kFlagWords) {
                // a nibble with two or more bits set makes the word hold more
                // bits than nonzero nibbles
                const uint32_t nz = nz_nibbles(w);
                fl |= (__popc(w) != __popc(nz) ? 1 : 0) | (nz != 0x88888888u ? 2 : 0);
            }
            h |= gather_nibble_bit0(((w >> 2) | (w >> 3)) & 0x11111111u) << (8 * b);
            l |= gather_nibble_bit0(((w >> 1) | (w >> 3)) & 0x11111111u) << (8 * b);
        }
        if (j < kUnitWords / 4) {
            hi[u * (kUnitWords / 4) + j] = h;
            lo[u * (kUnitWords / 4) + j] = l;
        }
    }
    fl = (__ballot(fl & 1) != 0 ? 1 : 0) | (__ballot(fl & 2) != 0 ? 2 : 0);
    if (tid == 0) flags_out[u] = (uint8_t)fl;
}

// Packs the hit segments into one array: segment s goes after segments
// n words of p set to v (the static search clears its counters with this
// kernel rather than a memset)
__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// 0..s-1 (counts already checked against the segment capacity).
__global__ __launch_bounds__(256) void plain_compact_kernel(const PlainRaw* __restrict__ in,
                                                            const uint32_t* __restrict__ counts, uint32_t seg_cap,
                                                            PlainRaw* __restrict__ out) {
    __shared__ uint32_t off;
    const uint32_t seg = blockIdx.x;
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        for (uint32_t t = 0; t < seg; t++) o += counts[t * kSegStride];
        off = o;
    }
    __syncthreads();
    const uint32_t n = min(counts[seg * kSegStride], seg_cap);
    const PlainRaw* src = in + (size_t)seg * seg_cap;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out[off + i] = src[i];
}

}  // namespace

struct PlainPlan {
    struct Launch {
        int W;
        uint32_t group_begin, n_groups;
        uint32_t max_pats;         // largest group of the launch
        uint32_t bit4_all;         // a group compares bit4 words on plain-base tiles
    };
    std::vector<GroupDev> groups;
    std::vector<Launch> launches;
    GroupDev* d_groups = nullptr;
    PatRec* d_pats = nullptr;
    PlaneRec* d_prec = nullptr;
    uint32_t* d_full = nullptr;
    uint4* d_frag = nullptr;
    uint32_t max_L = 0;
};

namespace {

// Matrix-core operand fragments of a plane group (layout and weights as in
// mfma_chunk): for each block of 16 plane records, 64 lanes x 16 bytes of fp4
// nibbles.  Every plane group with a window position is scored exactly.
bool build_mfma_frags(GroupDev& G, const std::vector<const Pattern*>& gp, std::vector<uint4>& frags) {
    if (!G.planes || G.wmask == 0) return false;
    // fp4 (e2m1) codes of +1 / -1 and of the bias weights by value in units of
    // 0.25 (bias feature 0.5 x weight): 12 = 6.0, 8 = 4.0, 6 = 3.0, 4 = 2.0,
    // 3 = 1.5, 2 = 1.0, 1 = 0.5
    constexpr uint8_t kPlus = 0x2, kMinus = 0xA;
    static const int kUnits[7] = {12, 8, 6, 4, 3, 2, 1};
    static const uint8_t kCode[7] = {0x7, 0x6, 0x5, 0x4, 0x3, 0x2, 0x1};
    // weights on the (h, l, hl) features: [genome == b] - (b == T)
    auto weights = [](uint32_t b, uint8_t w[3]) {
        w[0] = w[1] = w[2] = 0;
        if (b == 1) { w[0] = kMinus; w[1] = kMinus; w[2] = kPlus; }        // T: -h - l + hl (+1 in the bias)
        else if (b == 2) { w[1] = kPlus; w[2] = kMinus; }                  // C: l - hl
        else if (b == 4) { w[0] = kPlus; w[2] = kMinus; }                  // A: h - hl
        else { w[2] = kPlus; }                                             // G: hl
    };
    const int Lw = __builtin_popcount(G.wmask);
    G.mfma = 1;
    G.frag_begin = (uint32_t)frags.size();
    const uint32_t nblk = G.prec_padded / 16;
    for (uint32_t blk = 0; blk < nblk; blk++) {
        for (uint32_t lane = 0; lane < 64; lane++) {
            const uint32_t c = lane >> 4, slot = blk * 16 + (lane & 15u);
            uint32_t dw[4] = {0, 0, 0, 0};
            // bias slots: 32 per pattern, 8 in dword 3 of each lane group
            int units_left;
            if (slot < gp.size()) {
                const Pattern& p = *gp[slot];
                int nT = 0;
                for (uint32_t j = 0; j < 8; j++) {
                    const uint32_t pos = c + 4 * j;
                    if (!((G.wmask >> pos) & 1u)) continue;
                    uint8_t w[3];
                    weights(p.b4[G.pk_lo + pos], w);
                    for (int f = 0; f < 3; f++) dw[f] |= (uint32_t)w[f] << (4 * j);
                }
                for (uint32_t pos = 0; pos < 32; pos++)
                    if (((G.wmask >> pos) & 1u) && p.b4[G.pk_lo + pos] == 1) nT++;
                // 0.5 (nT - Lw + M) + 0.25 in units of 0.25
                units_left = 2 * (nT - Lw + (int)std::min<uint32_t>(p.M, 64u)) + 1;
            } else {
                units_left = -12 * 32;   // filler: -96, never passes
            }
            // the pattern's bias split greedily over its 32 slots (lane group
            // 0 first); every lane of the column computes the same split
            const int sign = units_left < 0 ? -1 : 1;
            int mag = sign * units_left;
            for (uint32_t g = 0; g < 4; g++)
                for (uint32_t j = 0; j < 8; j++) {
                    uint8_t code = 0;
                    for (int k = 0; k < 7; k++)
                        if (mag >= kUnits[k]) {
                            code = kCode[k];
                            mag -= kUnits[k];
                            break;
                        }
                    if (code && sign < 0) code |= 0x8;
                    if (g == c) dw[3] |= (uint32_t)code << (4 * j);
                }
            if (mag != 0) throw Error(COF_E_UNSUPPORTED, "matrix-core bias does not fit its slots");
            frags.push_back(make_uint4(dw[0], dw[1], dw[2], dw[3]));
        }
    }
    return true;
}

}  // namespace

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
    std::vector<PatRec> recs;
    std::vector<PlaneRec> precs;
    std::vector<uint32_t> full;
    std::vector<uint4> frags;
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
            // filter code replicated over the 8 nibbles of a word
            uint32_t fv = kCodec.enc_mixed[(uint8_t)f[k]];
            G.pam_pos[G.n_pam] = k;
            G.pam_mask[G.n_pam] = fv * 0x11111111u;
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
        // 2-bit planes: positions that are N in every pattern of the group
        // leave the window; the rest must be plain bases in every pattern and
        // fit one 32-bit word
        std::vector<bool> all_n(L, true);
        for (uint32_t pi : kv.second)
            for (uint32_t k = 0; k < L; k++) all_n[k] = all_n[k] && q.pats[pi].b4[k] == 0xF;
        uint32_t pk_lo = L, pk_hi = 0;
        for (uint32_t k = 0; k < L; k++)
            if (!all_n[k]) { pk_lo = std::min(pk_lo, k); pk_hi = std::max(pk_hi, k + 1); }
        if (pk_lo >= pk_hi) pk_lo = pk_hi = 0;
        bool planes = pk_hi - pk_lo <= 32;
        for (uint32_t pi : kv.second)
            for (uint32_t k = pk_lo; planes && k < pk_hi; k++)
                if (!all_n[k] && !single_base(q.pats[pi].b4[k])) planes = false;
        G.planes = planes ? 1u : 0u;
        G.pk_lo = pk_lo;
        G.wmask = 0;
        for (uint32_t k = pk_lo; k < pk_hi; k++)
            if (!all_n[k]) G.wmask |= 1u << (k - pk_lo);
        G.prec_begin = (uint32_t)precs.size();
        for (uint32_t pi : kv.second) {
            const Pattern& p = q.pats[pi];
            uint32_t cw[8] = {0}, fw[8] = {0};
            for (uint32_t k = k_lo; k < k_hi; k++) {
                if (excl[k]) continue;
                uint32_t r = k - k_lo;
                cw[r / 8] |= (uint32_t)p.b4[k] << (4 * (r % 8));
            }
            for (uint32_t k = 0; k < L; k++) fw[k / 8] |= (uint32_t)p.b4[k] << (4 * (k % 8));
            PatRec r{};
            for (int w = 0; w < 8; w++) r.w[w] = cw[w];
            r.thr = (int32_t)L - (int32_t)n_excl - (int32_t)p.M;
            r.M = p.M;
            r.pid = pi;
            recs.push_back(r);
            PlaneRec pr{0, 0, 64u - std::min<uint32_t>(p.M, 64u), pi};
            if (planes)
                for (uint32_t k = pk_lo; k < pk_hi; k++) {
                    if (all_n[k]) continue;
                    const uint32_t v = p.b4[k];   // T C A G = 1 2 4 8 -> (0,0) (0,1) (1,0) (1,1)
                    pr.ph |= (uint32_t)((v & 0xC) != 0) << (k - pk_lo);
                    pr.pl |= (uint32_t)((v & 0xA) != 0) << (k - pk_lo);
                }
            precs.push_back(pr);
            full.insert(full.end(), fw, fw + 8);
            pat_index++;
        }
        while ((precs.size() - G.prec_begin) % 16) precs.push_back(PlaneRec{0, 0, 200u, 0u});
        G.prec_padded = (uint32_t)(precs.size() - G.prec_begin);
        std::vector<const Pattern*> gpats;
        for (uint32_t pi : kv.second) gpats.push_back(&q.pats[pi]);
        build_mfma_frags(G, gpats, frags);
        by_w[(int)G.W].push_back(G);
    }
    for (auto& kv : by_w) {
        uint32_t mp = 0, bit4 = 0;
        for (const GroupDev& G : kv.second) {
            mp = std::max(mp, G.pat_count);
            bit4 |= G.planes ? 0u : 1u;
        }
        PlainPlan::Launch l{kv.first, (uint32_t)pp->groups.size(), (uint32_t)kv.second.size(), mp, bit4};
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
        up(pp->d_pats, recs);
        up(pp->d_prec, precs);
        up(pp->d_full, full);
        if (!frags.empty()) up(pp->d_frag, frags);
    }
    return pp;
}

void plain_plan_destroy(PlainPlan* pp) {
    if (!pp) return;
    (void)hipFree(pp->d_groups);
    (void)hipFree(pp->d_pats);
    (void)hipFree(pp->d_prec);
    (void)hipFree(pp->d_full);
    (void)hipFree(pp->d_frag);
    delete pp;
}

bool plain_plan_empty(const PlainPlan* pp) { return !pp || pp->launches.empty(); }

uint64_t plain_tile_nucl() { return (uint64_t)kMaxThreads * 64; }
uint64_t plain_pad_words() { return kMaxTileWords + 16; }

namespace {

template <int W>
void launch_plain(const PlainTune& t, dim3 grid, hipStream_t st, const uint32_t* d_genome, uint64_t scan_end,
                  uint64_t first, const GroupDev* g, uint32_t n_groups, uint32_t bit4_all, const PatRec* pats,
                  const PlaneRec* prec, const uint32_t* full, const uint4* frag, const Workspace& ws, PlainRaw* out,
                  uint32_t* count, uint32_t cap) {
#define COF_PLAIN_SHAPE(T)                                                                                  \
    case T:                                                                                                 \
        hipLaunchKernelGGL((plain_scan_kernel<W, T>), grid, dim3(T), 0, st, d_genome, scan_end, first, g, n_groups, \
                           pats, prec, full, frag, ws.img_hi, ws.img_lo, ws.img_flags, out, count, cap,      \
                           t.pat_chunk, bit4_all);                                                          \
        break;
    switch (t.threads) {
        COF_PLAIN_SHAPE(64) COF_PLAIN_SHAPE(128) COF_PLAIN_SHAPE(256) COF_PLAIN_SHAPE(512)
        default: throw Error(COF_E_INVALID, "plain workgroup size must be 64, 128, 256 or 512");
    }
#undef COF_PLAIN_SHAPE
}

// The persistent scan: as many two-wave workgroups as stay resident on the
// device, draining the per-XCD tile queues (heads cleared by the caller).
template <int W>
void launch_queue(hipStream_t st, const uint32_t* d_genome, uint64_t scan_end, const GroupDev* g, uint32_t n_groups,
                  uint32_t bit4_all, const PatRec* pats, const PlaneRec* prec, const uint32_t* full, const uint4* frag,
                  const Workspace& ws, PlainRaw* out, uint32_t* count, uint32_t cap) {
    static int per_cu = 0;
    if (!per_cu) {
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, plain_queue_kernel<W>, 128, 0));
        per_cu = std::max(per_cu, 1);
    }
    const uint64_t n_tiles = (scan_end + 4095) / 4096;
    if (n_tiles > 0xFFFFFFFFull) throw Error(COF_E_UNSUPPORTED, "genome too large for the tile queues");
    const uint32_t grid = (uint32_t)std::min<uint64_t>(n_tiles, (uint64_t)per_cu * (uint64_t)std::max(ws.cus, 1));
    hipLaunchKernelGGL((plain_queue_kernel<W>), dim3(grid), dim3(128), 0, st, d_genome, scan_end, (uint32_t)n_tiles,
                       g, n_groups, pats, prec, full, frag, ws.img_hi, ws.img_lo, ws.img_flags, out, count, cap,
                       bit4_all, ws.plain_heads);
}

}  // namespace

// Scans start positions [0, scan_end) of the resident genome.  Device words
// past the genome must be zero up to plain_pad_words() beyond the last tile.
Workspace::~Workspace() {
    for (void* p : {plain_out, plain_out2, (void*)plain_count, (void*)plain_heads, queue, bulge_out, tiles, chunks, (void*)counters, (void*)sums,
                    qs[0], qs[1], (void*)qsteps[0], (void*)qsteps[1], q_first, (void*)q_first_steps, (void*)chunk_lut,
                    (void*)tile_order, (void*)tile_cand, (void*)tile_cand_prev, (void*)tile_steps, (void*)rounds,
                    (void*)round_ctr, sq[0], sq[1], (void*)screen_ctr, (void*)img_hi, (void*)img_lo,
                    (void*)img_flags})
        if (p) (void)hipFree(p);
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : rev) if (e) (void)hipEventDestroy(e);
    for (auto& x : bulge_exec) if (x) (void)hipGraphExecDestroy(x);
    if (st2) (void)hipStreamDestroy(st2);
    if (st3) (void)hipStreamDestroy(st3);
    if (h_plain) (void)hipHostFree(h_plain);
    if (h_counts) (void)hipHostFree(h_counts);
    if (h_bulge) (void)hipHostFree(h_bulge);
}

void Workspace::ensure_host(void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap && *p) return;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    size_t want = std::max<size_t>(bytes + bytes / 4, 1u << 20);
    HIPCHECK(hipHostMalloc(p, want, hipHostMallocDefault));
    *cap = want;
}

void Workspace::ensure_round_events(uint32_t r) {
    while (rev.size() < 2 * (size_t)r) {
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        rev.push_back(e);
    }
}

void Workspace::ensure_events() {
    if (ev[0]) return;
    for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
    HIPCHECK(hipMalloc(&counters, 4 * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&sums, 8 * sizeof(unsigned long long)));
    HIPCHECK(hipMalloc(&plain_count, kHitSegs * kSegStride * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&plain_heads, kXcds * kSegStride * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&round_ctr, 2 * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&screen_ctr, 2 * sizeof(uint32_t)));
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
}

void ws_grow(void** ptr, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap && *ptr) return;
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    HIPCHECK(hipMalloc(ptr, need * elem));
    *cap = need;
}

namespace {

// Scans of all contexts on one device run one after another: a scan fills
// the GPU, so overlapping two gains nothing, and back-to-back scans keep each
// context's event timing its own kernels.  A context's stream waits for the
// event recorded after the previous scan on the device (any context), while
// the host stays free to convert the previous context's hits.
struct ScanOrder {
    std::mutex m;
    hipEvent_t last[64] = {};
};
ScanOrder& scan_order() {
    static ScanOrder o;
    return o;
}

// Builds the scan image for starts [0, n_nib) unless the current one covers
// them: units up to the last one of the widest tile shape, plus one for the
// planes of the last tile's halo.
void ensure_scan_image(Workspace& ws, const uint32_t* d_genome, uint64_t n_nib, hipStream_t st) {
    const uint64_t widest = plain_tile_nucl();
    const uint64_t units = (n_nib + widest - 1) / widest * (widest / 4096) + 1;
    if (ws.img_units >= units) return;
    if (ws.img_cap_units < units) {
        for (void* p : {(void*)ws.img_hi, (void*)ws.img_lo, (void*)ws.img_flags})
            if (p) (void)hipFree(p);
        ws.img_hi = ws.img_lo = nullptr;
        ws.img_flags = nullptr;
        ws.img_cap_units = 0;
        HIPCHECK(hipMalloc(&ws.img_hi, units * 128 * sizeof(uint32_t)));
        HIPCHECK(hipMalloc(&ws.img_lo, units * 128 * sizeof(uint32_t)));
        HIPCHECK(hipMalloc(&ws.img_flags, units));
        ws.img_cap_units = units;
    }
    hipLaunchKernelGGL(scan_image_kernel, dim3((uint32_t)units), dim3(64), 0, st, d_genome, ws.img_hi, ws.img_lo,
                       ws.img_flags);
    HIPCHECK(hipGetLastError());
    ws.img_units = units;
}

}  // namespace

// Queues the scan of start positions [0, scan_end) and the copy of the
// per-segment hit counts to the host; plain_finish completes it.
void plain_launch(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
                  uint64_t* n_launches, const PlainTune& tune) {
    ws.plain_pending = false;
    if (plain_plan_empty(pp) || scan_end == 0) return;
    ws.ensure_events();
    if (!ws.plain_out || (tune.hit_cap && tune.hit_cap != ws.plain_cap_tuned)) {
        // first scan, or a tuned starting size: (re)size the hit buffer
        if (ws.plain_out) (void)hipFree(ws.plain_out);
        ws.plain_out = nullptr;
        size_t cap = 0;
        ws_grow(&ws.plain_out, &cap, tune.hit_cap ? tune.hit_cap : (1u << 20), sizeof(PlainRaw));
        ws.plain_cap = (uint32_t)cap;
        ws.plain_cap_tuned = tune.hit_cap;
    }
    size_t hc = ws.h_counts_cap;
    ws.ensure_host((void**)&ws.h_counts, &hc, kHitSegs * kSegStride * sizeof(uint32_t));
    ws.h_counts_cap = hc;
    const uint64_t kTile = (uint64_t)tune.threads * 64;
    const uint64_t n_tiles = (scan_end + kTile - 1) / kTile;
    PlainRaw* d_out = (PlainRaw*)ws.plain_out;
    const uint32_t seg_cap = ws.plain_cap / kHitSegs;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    ScanOrder& order = scan_order();
    std::lock_guard<std::mutex> order_lock(order.m);
    hipEvent_t& last = order.last[dev & 63];
    if (last) HIPCHECK(hipStreamWaitEvent(st, last, 0));
    else HIPCHECK(hipEventCreateWithFlags(&last, hipEventDisableTiming));
    ensure_scan_image(ws, d_genome, std::max(scan_end, ws.genome_nib), st);
    HIPCHECK(hipMemsetAsync(ws.plain_count, 0, kHitSegs * kSegStride * sizeof(uint32_t), st));
    HIPCHECK(hipEventRecord(ws.ev[0], st));
    const bool use_queue = tune.queue && tune.threads == 64 && tune.pat_chunk == 0;
    for (const auto& l : pp->launches) {
        if (use_queue) {
            const GroupDev* g = pp->d_groups + l.group_begin;
            HIPCHECK(hipMemsetAsync(ws.plain_heads, 0, kXcds * kSegStride * sizeof(uint32_t), st));
#define COF_QUEUE_LAUNCH(WW)                                                                                    \
    case WW:                                                                                                    \
        launch_queue<WW>(st, d_genome, scan_end, g, l.n_groups, l.bit4_all, pp->d_pats, pp->d_prec, pp->d_full, \
                         pp->d_frag, ws, d_out, ws.plain_count, seg_cap);                                       \
        break;
            switch (l.W) {
                COF_QUEUE_LAUNCH(1) COF_QUEUE_LAUNCH(2) COF_QUEUE_LAUNCH(3) COF_QUEUE_LAUNCH(4)
                COF_QUEUE_LAUNCH(5) COF_QUEUE_LAUNCH(6) COF_QUEUE_LAUNCH(7) COF_QUEUE_LAUNCH(8)
                default: throw Error(COF_E_UNSUPPORTED, "compare window too wide");
            }
#undef COF_QUEUE_LAUNCH
            HIPCHECK(hipGetLastError());
            if (n_launches) (*n_launches)++;
            continue;
        }
        // grids of at most 2^20 tiles keep the tile index in 32 bits
        for (uint64_t t0 = 0; t0 < n_tiles; t0 += (1u << 20)) {
            uint32_t nb = (uint32_t)std::min<uint64_t>(n_tiles - t0, 1u << 20);
            uint64_t first = t0 * kTile;
            const GroupDev* g = pp->d_groups + l.group_begin;
            uint32_t ny = tune.pat_chunk ? (l.max_pats + tune.pat_chunk - 1) / tune.pat_chunk : 1u;
            dim3 grid(nb, std::max(1u, ny));
#define COF_PLAIN_LAUNCH(WW)                                                                                   \
    case WW:                                                                                                   \
        launch_plain<WW>(tune, grid, st, d_genome, scan_end, first, g, l.n_groups, l.bit4_all, pp->d_pats,     \
                         pp->d_prec, pp->d_full, pp->d_frag, ws, d_out, ws.plain_count, seg_cap);          \
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
    HIPCHECK(hipEventRecord(ws.ev[1], st));
    HIPCHECK(hipEventRecord(last, st));
    HIPCHECK(hipMemcpyAsync(ws.h_counts, ws.plain_count, kHitSegs * kSegStride * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, st));
    ws.plain_pending = true;
    ws.plain_scan_end = scan_end;
}

// Waits for a launched scan and brings its hits to the host; a segment that
// overflowed grows every segment and the scan runs again.
void plain_finish(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, hipStream_t st, RawView<PlainRaw>& out,
                  double* kernel_ms, uint64_t* n_launches, const PlainTune& tune) {
    out = RawView<PlainRaw>{};
    if (!ws.plain_pending) return;
    for (;;) {
        HIPCHECK(hipStreamSynchronize(st));
        ws.plain_pending = false;
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]));
        if (kernel_ms) *kernel_ms += ms;
        const uint32_t seg_cap = ws.plain_cap / kHitSegs;
        uint64_t total = 0;
        uint32_t most = 0;
        for (uint32_t k = 0; k < kHitSegs; k++) {
            total += ws.h_counts[k * kSegStride];
            most = std::max(most, ws.h_counts[k * kSegStride]);
        }
        if (most > seg_cap) {
            // a segment overflowed: grow every segment past the fullest one and scan again
            const uint64_t need = ((uint64_t)most + most / 4 + 1024) * kHitSegs;
            if (need > 0xFFFFFFF0ull) throw Error(COF_E_NOMEM, "plain search: too many hits");
            size_t cap = ws.plain_cap;
            ws_grow(&ws.plain_out, &cap, need, sizeof(PlainRaw));
            ws.plain_cap = (uint32_t)cap;
            plain_launch(pp, ws, d_genome, ws.plain_scan_end, st, n_launches, tune);
            continue;
        }
        const uint32_t cnt = (uint32_t)total;
        ws.ensure_host(&ws.h_plain, &ws.h_plain_cap, (size_t)cnt * sizeof(PlainRaw));
        if (cnt) {
            size_t cap2 = ws.plain_cap2;
            ws_grow(&ws.plain_out2, &cap2, cnt, sizeof(PlainRaw));
            ws.plain_cap2 = (uint32_t)cap2;
            hipLaunchKernelGGL(plain_compact_kernel, dim3(kHitSegs), dim3(256), 0, st, (const PlainRaw*)ws.plain_out,
                               ws.plain_count, seg_cap, (PlainRaw*)ws.plain_out2);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(ws.h_plain, ws.plain_out2, (size_t)cnt * sizeof(PlainRaw), hipMemcpyDeviceToHost,
                                    st));
        }
        HIPCHECK(hipStreamSynchronize(st));
        out = RawView<PlainRaw>{(const PlainRaw*)ws.h_plain, cnt};
        return;
    }
}

void plain_run(PlainPlan* pp, Workspace& ws, const uint32_t* d_genome, uint64_t scan_end, hipStream_t st,
               RawView<PlainRaw>& out, double* kernel_ms, uint64_t* n_launches, const PlainTune& tune) {
    plain_launch(pp, ws, d_genome, scan_end, st, n_launches, tune);
    plain_finish(pp, ws, d_genome, st, out, kernel_ms, n_launches, tune);
}

// ---------------------------------------------------------------------------
// Allocation-free plain search: the same scan, its hit segments, counters and
// packed copy in the static `scratch` array instead of the workspace.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_static_m;
void* g_static_ptr[kMaxStaticDevices] = {};
}  // namespace

static_assert(kStaticSegCap * 64ull * sizeof(PlainRaw) == kStaticSegBytes, "static scratch layout");
static_assert(kHitSegs == 64 && kSegStride == 32, "static scratch counter layout");

void plain_static_init(int device) {
    if (device < 0 || device >= (int)kMaxStaticDevices)
        throw Error(COF_E_INVALID, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_static_m);
    if (g_static_ptr[device]) return;
    int cur = -1;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != device) throw Error(COF_E_INVALID, "plain_static_init: device is not current");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(COF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_static_ptr[device] = p;
}

uint64_t plain_static_max_hits() { return (uint64_t)kHitSegs * kStaticSegCap; }

void plain_search_static(PlainPlan* pp, const Workspace& ws, const uint32_t* d_genome, uint64_t scan_end,
                         hipStream_t st, const PlainTune& tune, uint32_t* h_counts, PlainRaw* h_stage,
                         RawView<PlainRaw>& out, uint64_t* n_launches) {
    out = RawView<PlainRaw>{};
    if (plain_plan_empty(pp) || scan_end == 0) return;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_static_m);
        if (dev >= 0 && dev < (int)kMaxStaticDevices) base = g_static_ptr[dev];
    }
    if (!base) throw Error(COF_E_INVALID, "static scratch not set up on this device (cof_open does it)");
    const uint64_t widest = plain_tile_nucl();
    if (ws.img_units < (scan_end + widest - 1) / widest * (widest / 4096) + 1)
        throw Error(COF_E_INVALID, "the genome's scan image is missing (it is built when a genome is loaded)");
    uint32_t* const counts = (uint32_t*)base;
    PlainRaw* const segs = (PlainRaw*)((uint8_t*)base + kStaticCountBytes);
    PlainRaw* const packed = (PlainRaw*)((uint8_t*)base + kStaticCountBytes + kStaticSegBytes);
    const uint32_t n_counts = kHitSegs * kSegStride;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(8), dim3(256), 0, st, counts, n_counts, 0u);
    HIPCHECK(hipGetLastError());
    const uint64_t kTile = (uint64_t)tune.threads * 64;
    const uint64_t n_tiles = (scan_end + kTile - 1) / kTile;
    for (const auto& l : pp->launches) {
        for (uint64_t t0 = 0; t0 < n_tiles; t0 += (1u << 20)) {
            uint32_t nb = (uint32_t)std::min<uint64_t>(n_tiles - t0, 1u << 20);
            uint64_t first = t0 * kTile;
            const GroupDev* g = pp->d_groups + l.group_begin;
            uint32_t ny = tune.pat_chunk ? (l.max_pats + tune.pat_chunk - 1) / tune.pat_chunk : 1u;
            dim3 grid(nb, std::max(1u, ny));
#define COF_PLAIN_LAUNCH(WW)                                                                                   \
    case WW:                                                                                                   \
        launch_plain<WW>(tune, grid, st, d_genome, scan_end, first, g, l.n_groups, l.bit4_all, pp->d_pats,     \
                         pp->d_prec, pp->d_full, pp->d_frag, ws, segs, counts, kStaticSegCap);             \
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
    HIPCHECK(hipMemcpyAsync(h_counts, counts, kStaticCountBytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    uint64_t total = 0;
    uint32_t most = 0;
    for (uint32_t k = 0; k < kHitSegs; k++) {
        total += h_counts[k * kSegStride];
        most = std::max(most, h_counts[k * kSegStride]);
    }
    if (most > kStaticSegCap)
        throw Error(COF_E_NOMEM, "plain hits exceed the static scratch: a hit segment holds at most " +
                                     std::to_string(kStaticSegCap) + " records and one received " +
                                     std::to_string(most) + " (cof_search has no such limit)");
    if (total) {
        hipLaunchKernelGGL(plain_compact_kernel, dim3(kHitSegs), dim3(256), 0, st, (const PlainRaw*)segs, counts,
                           kStaticSegCap, packed);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(h_stage, packed, total * sizeof(PlainRaw), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    out = RawView<PlainRaw>{h_stage, (size_t)total};
}

void plain_prepare_image(Workspace& ws, const uint32_t* d_genome, hipStream_t st) {
    if (ws.genome_nib) ensure_scan_image(ws, d_genome, ws.genome_nib, st);
}

}  // namespace cof
