// Brute-force bulge alignment enumerator: a test checker for the CPU oracle
// (oracle/cof_ref.cpp).  Test infrastructure only; built by tests/bulge_brute.py.
//
// It shares no code and no algorithm with the oracle's bulge path.  The
// oracle (like the reference, cas-offinder-lib/src/search.rs:779-897) runs a
// PAM quick check, a Myers filter, a DP table and a backward DFS pruned by
// that table, then checks the PAM along each alignment.  Here, for every end
// position j and every start s of its text window, a forward recursion
// extends the alignment op by op (diagonal, RNA bulge, DNA bulge) and keeps
// only what the per-op rules and the per-type caps allow.  There is no DP
// table, no Myers score and no quick check (unless asked for, see below): an
// alignment is emitted iff it satisfies the rules themselves.
//
// Rules, each from the reference text:
//   * match test cmp_loose_bit4 (traceback.rs:41-46): text 0 never matches,
//     pattern N (0xF) matches any other text, text N mismatches a specific
//     pattern base, otherwise bit AND;
//   * no diagonal mismatch at a pattern N (traceback.rs:330-344);
//   * RNA bulge (pattern char consumed alone) forbidden at an N position and
//     next to an N position on either side (traceback.rs:346-366);
//   * DNA bulge (text char consumed alone) only between two pattern
//     positions, not with N on both sides (traceback.rs:368-381);
//   * DP cells on text 0 (padding) are infinite (traceback.rs:203-210), so
//     no state of the path may sit right after a padding nibble inside the
//     window: consumed chars are never 0, and a state that consumed no text
//     yet (leading RNA bulges) must not follow a 0 inside the window;
//   * caps: mismatches <= M, DNA bulges <= D, RNA bulges <= R (their sum is
//     then <= M + D + R, the DP bound of traceback.rs:269-327, so the DP
//     prune removes nothing else);
//   * the full PAM check of classify_myers_candidate (search.rs:634-671):
//     the effective filter cell of every diagonal op accepts the genome base;
//   * window: ends j with j + 1 >= L, text window [max(0, j + 1 - W), j]
//     with W = L + D (search.rs:815-817, 612-616).
// The Myers filter (search.rs:841-877) is a lower bound of every alignment
// cost, so it removes nothing either.  The quick PAM check
// (search.rs:703-755) is implied by the full check when a pattern's N cells
// are one run at one end of the pattern; for other patterns it is a separate
// gate, applied here only when quick_gate is set.
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

struct Enum {
    const uint8_t* text;
    uint64_t n;
    const uint8_t* pat;
    const uint8_t* isn;
    const uint8_t* filt;
    uint32_t L, M, D, R;
    uint64_t j, s, ts;
    char ops[256];
    FILE* out;
    uint64_t emitted;
};

bool loose(uint8_t p, uint8_t t) {
    if (t == 0) return false;
    if (p == 0xF) return true;
    if (t == 0xF) return false;
    return (p & t) != 0;
}

bool filter_accepts(uint8_t f, uint8_t g) {
    if (f == 0xF) return g != 0;
    if (g == 0 || g == 0xF) return false;
    return (g & f) != 0;
}

// i pattern chars and text [s, t) consumed in k ops
void extend(Enum& e, uint32_t i, uint64_t t, uint32_t k, uint32_t mm, uint32_t db, uint32_t rb) {
    const int64_t rest_p = (int64_t)e.L - i, rest_t = (int64_t)(e.j + 1) - (int64_t)t;
    // what is left must be coverable with the bulges left
    if (rest_t - rest_p > (int64_t)(e.D - db) || rest_p - rest_t > (int64_t)(e.R - rb)) return;
    if (i == e.L) {
        if (rest_t != 0) return;
        fprintf(e.out, "%llu %llu %u %u %u %.*s\n", (unsigned long long)e.j, (unsigned long long)e.s, mm, db, rb,
                (int)k, e.ops);
        e.emitted++;
        return;
    }
    if (t <= e.j) {                                   // diagonal: pattern i against text t
        const uint8_t tv = e.text[t];
        if (tv != 0 && filter_accepts(e.filt[i], tv)) {
            const uint32_t mis = loose(e.pat[i], tv) ? 0 : 1;
            if (!(mis && e.isn[i]) && mm + mis <= e.M) {
                e.ops[k] = mis ? 'S' : 'M';
                extend(e, i + 1, t + 1, k + 1, mm + mis, db, rb);
            }
        }
    }
    if (rb < e.R && !e.isn[i] && !(i + 1 < e.L && e.isn[i + 1]) && !(i >= 1 && e.isn[i - 1])) {
        if (!(t > e.ts && e.text[t - 1] == 0)) {      // RNA bulge: pattern i alone
            e.ops[k] = 'R';
            extend(e, i + 1, t, k + 1, mm, db, rb + 1);
        }
    }
    if (db < e.D && i >= 1 && t <= e.j && !(e.isn[i - 1] && e.isn[i]) && e.text[t] != 0) {
        e.ops[k] = 'D';                               // DNA bulge: text t alone
        extend(e, i, t + 1, k + 1, mm, db + 1, rb);
    }
}

// check_pam_quick / precompute_pam_precheck (search.rs:703-777), only used
// when quick_gate is set
bool quick_gate(const Enum& e, uint64_t j, uint64_t total) {
    uint32_t off = 0;
    while (off < e.L && !e.isn[off]) off++;
    std::vector<uint8_t> pf;
    for (uint32_t p = off; p < e.L; p++)
        if (e.isn[p]) pf.push_back(e.filt[p]);
    const int64_t nd = off == 0 ? e.D + 1 : 1, nr = off == 0 ? e.R + 1 : 1;
    for (int64_t bd = 0; bd < nd; bd++)
        for (int64_t br = 0; br < nr; br++) {
            const int64_t span = (int64_t)e.L + bd - br;
            if (span <= 0 || span > (int64_t)j + 1) continue;
            const uint64_t a = j + 1 - (uint64_t)span;
            bool ok = true;
            for (size_t k = 0; k < pf.size() && ok; k++) {
                const uint64_t g = a + off + k;
                ok = g < total && filter_accepts(pf[k], e.text[g]);
            }
            if (ok) return true;
        }
    return false;
}

}  // namespace

extern "C" {

// Every alignment of one pattern (bit4 codes `pat`, N flags `isn`, effective
// filter codes `filt`, length L) ending at any j in [0, n) of `text` (one
// nibble per byte), written to out_path as "j s mm db rb OPS" lines (ops in
// alignment order).  Returns the number of lines, -1 if the file cannot be
// written.
int64_t brute_bulge_enum(const uint8_t* text, uint64_t n, const uint8_t* pat, const uint8_t* isn,
                         const uint8_t* filt, uint32_t L, uint32_t M, uint32_t D, uint32_t R, int use_quick_gate,
                         const char* out_path) {
    FILE* f = fopen(out_path, "w");
    if (!f || L == 0 || L + D > 200) {
        if (f) fclose(f);
        return -1;
    }
    Enum e{text, n, pat, isn, filt, L, M, D, R, 0, 0, 0, {}, f, 0};
    const uint64_t W = (uint64_t)L + D;
    for (uint64_t j = L - 1; j < n; j++) {
        if (e.text[j] == 0) continue;                 // the end state reads text j
        if (use_quick_gate && !quick_gate(e, j, n)) continue;
        e.j = j;
        e.ts = j + 1 >= W ? j + 1 - W : 0;
        for (uint64_t s = e.ts; s <= j; s++) {
            e.s = s;
            extend(e, 0, s, 0, 0, 0, 0);
        }
    }
    fclose(f);
    return (int64_t)e.emitted;
}

}  // extern "C"
