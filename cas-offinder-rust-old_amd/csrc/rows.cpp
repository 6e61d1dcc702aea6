// Host-side conversion of raw device hits into reference output rows.
//
// Plain path: the device emits every (start, pattern) whose window passes the
// writer predicate on the resident genome.  Here each hit is placed in the
// virtual batch the reference would have reported it from (search.rs:1525-1572
// batcher, :540-546 head-overlap start limit, :1449-1454 drop rule), the
// predicate is re-evaluated on that batch's view (bytes past the batch end
// read as 0, as the reference's zero-filled batch buffer does), and the row
// is formatted as main.rs:158-242 does.
//
// Bulge path: the device already scanned per virtual batch (window floor and
// PAM limit); here the chromosome-end drop rule of convert_matches_myers
// (search.rs:1316-1374) is applied and rows are formatted (main.rs:145-157).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <functional>
#include <mutex>
#include <thread>

#include "cof_internal.hpp"

namespace cof {

BatchIndex::BatchIndex(const Genome& g, const std::vector<Batch>& bs) {
    nonhead_batch.assign(g.chunks.size(), -1);
    head_batch.assign(g.chunks.size(), -1);
    for (size_t b = 0; b < bs.size(); b++) {
        if (bs[b].has_head) head_batch[bs[b].first] = (int64_t)b;
        for (size_t c = bs[b].first + (bs[b].has_head ? 1 : 0); c < bs[b].last; c++) nonhead_batch[c] = (int64_t)b;
        uint64_t e = bs[b].last < g.chunks.size() ? g.chunks[bs[b].last].byte_off : g.n_bytes;
        end_nib.push_back(e * 2);
        active_begin.push_back(g.chunks[bs[b].first + (bs[b].has_head ? 1 : 0)].byte_off * 2);
    }
}

namespace {

// Writer predicate of main.rs:158-194 evaluated in pattern orientation on the
// N-masked window (search.rs:512-531), with nibbles at or past view_end = 0.
bool plain_predicate(const Genome& g, const Pattern& p, uint64_t pos, uint64_t view_end, uint32_t* mm_out) {
    int32_t pop = 0;
    uint32_t amm = 0;
    for (uint32_t k = 0; k < p.L; k++) {
        uint64_t i = pos + k;
        uint8_t v = i < view_end ? nib(g.packed.data(), i) : 0;
        if (v == 0xF) v = 0;
        uint8_t pb = p.b4[k];
        pop += __builtin_popcount(pb & v);
        uint8_t fc = (uint8_t)p.eff_filter[k];
        bool sb = single_base(v);
        if (!is_n_char((char)fc) && !(sb && (v & kCodec.enc_mixed[fc]))) return false;
        if (pb != 0xF && !(sb && (v & pb))) amm++;
    }
    int32_t kmm = (int32_t)p.L - pop;
    if (kmm > (int32_t)p.M || amm > p.M) return false;
    *mm_out = amm;
    return true;
}

}  // namespace

namespace {

// Persistent worker pool for the converters (creating 16 threads per search
// costs about a millisecond).  After a job the workers spin for a short while
// before sleeping, so a conversion that follows soon after (the plain then the
// bulge hits of one search, or the next search's) starts without a
// condition-variable wake-up.
class Pool {
public:
    static Pool& get() {
        static Pool pool;
        return pool;
    }
    unsigned size() const { return (unsigned)workers_.size() + 1; }
    // runs f(t) for t in [0, n) on the workers and the calling thread
    void run(unsigned n, const std::function<void(unsigned)>& f) {
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            n_tasks_ = n;
            next_ = 0;
            done_ = 0;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [&] { return done_ == n_tasks_; });
        job_ = nullptr;
    }

private:
    Pool() {
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        unsigned nw = std::min(16u, hw) - 1;
        for (unsigned w = 0; w < nw; w++)
            workers_.emplace_back([this] {
                uint64_t seen = 0;
                for (;;) {
                    const auto t0 = std::chrono::steady_clock::now();
                    while (gen_.load(std::memory_order_acquire) == seen && !stop_.load(std::memory_order_relaxed) &&
                           std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(300))
                        std::this_thread::yield();
                    std::unique_lock<std::mutex> lk(m_);
                    cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
                    if (stop_) return;
                    seen = gen_.load();
                    lk.unlock();
                    work();
                }
            });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    void work() {
        for (;;) {
            unsigned t;
            const std::function<void(unsigned)>* f;
            {
                std::lock_guard<std::mutex> lk(m_);
                if (!job_ || next_ >= n_tasks_) return;
                t = next_++;
                f = job_;
            }
            (*f)(t);
            std::lock_guard<std::mutex> lk(m_);
            if (++done_ == n_tasks_) done_cv_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(unsigned)>* job_ = nullptr;
    unsigned n_tasks_ = 0, next_ = 0, done_ = 0;
    std::atomic<uint64_t> gen_{0};
    std::atomic<bool> stop_{false};
};

// Runs fn(begin, end, out, aux) -> count over slices of [0, n) in parallel,
// once: slice [b0, b1) writes the hits it keeps to out/aux positions b0,
// b0 + 1, ... of the final arrays (sized for every raw hit), leaving holes
// after them.  The converters keep almost every raw hit, so the few holes
// below the kept total are then filled with kept hits from above it (hit
// order is unspecified by the interface; out[i] and aux[i] move together).
template <typename Fn>
void parallel_convert(size_t n, HitVec& out, AuxVec& aux, Fn fn) {
    if (!n) return;
    const size_t base = out.size();
    Pool& pool = Pool::get();
    const unsigned nt = n < 4096 ? 1u : pool.size() * 4;
    std::vector<size_t> cnt(nt);
    auto lo_of = [&](unsigned t) { return n * t / nt; };
    out.resize(base + n);
    aux.resize(base + n);
    cof_hit* const o = out.data() + base;
    HitAux* const a = aux.data() + base;
    auto body = [&](unsigned t) {
        const size_t b0 = lo_of(t), b1 = lo_of(t + 1);
        cnt[t] = fn(b0, b1, o + b0, a + b0);
    };
    if (nt == 1) body(0);
    else pool.run(nt, body);
    size_t kept = 0;
    for (unsigned t = 0; t < nt; t++) kept += cnt[t];
    if (kept < n) {
        // the holes in [0, kept) are as many as the kept hits in [kept, n):
        // pair them, holes upwards from slice 0, kept hits downwards from the
        // last slice
        unsigned st = nt - 1;
        size_t s_end = lo_of(st) + cnt[st];   // one past the next kept hit to move
        for (unsigned t = 0; t < nt && lo_of(t) < kept; t++) {
            const size_t h_end = std::min(lo_of(t + 1), kept);
            for (size_t h = lo_of(t) + cnt[t]; h < h_end; h++) {
                while (s_end <= std::max(lo_of(st), kept)) {
                    if (st == 0) throw Error(COF_E_INVALID, "internal: hit conversion lost count of its holes");
                    st--;
                    s_end = lo_of(st) + cnt[st];
                }
                s_end--;
                o[h] = o[s_end];
                a[h] = a[s_end];
            }
        }
    }
    out.resize(base + kept);
    aux.resize(base + kept);
}
}  // namespace

bool place_plain_hit(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                     const PlainRaw& r, cof_hit& h, HitAux& a) {
    const Pattern& p = q.pats[r.pattern];
    size_t c = g.chunk_of(r.pos);
    const Chunk& ch = g.chunks[c];
    uint64_t chrom_pos = ch.start + (r.pos - ch.nib_begin());
    bool past_end = chrom_pos + p.L > ch.end;
    const int64_t b = bi.nonhead_batch[c], hb = bi.head_batch[c];
    if (b < 0 && hb < 0) return false;
    bool next_new_chrom = c + 1 < g.chunks.size() && g.chunks[c + 1].start == 0;
    uint32_t mm = 0;
    uint64_t view_end = 0;
    bool emit = false;
    // The device evaluated the predicate on the resident genome, which
    // equals the batch view whenever the window ends inside it; only
    // windows running past the view end are evaluated again here.
    auto check = [&](uint64_t ve) {
        view_end = ve;
        if (r.pos + p.L <= ve) { mm = r.flags; return true; }
        return plain_predicate(g, p, r.pos, ve, &mm);
    };
    // reported by batch b, where c is not the head
    if (b >= 0) {
        bool last_in_b = c + 1 == bs[(size_t)b].last;
        if (!((last_in_b || next_new_chrom) && past_end)) emit = check(bi.end_nib[(size_t)b]);
    }
    // reported by batch hb, where c is the head (search.rs:540-546); the
    // two cases exclude each other (c is the last chunk of batch b, so
    // the first case needs the window inside c, the second a start within
    // L - 1 of c's end)
    if (!emit && hb >= 0) {
        uint64_t c1 = g.chunks[c + 1].byte_off * 2 - ch.nib_begin();
        uint64_t min_start = c1 >= p.L - 1 ? c1 - (p.L - 1) : 0;
        uint64_t local = r.pos - ch.nib_begin();
        if (local >= min_start && !(next_new_chrom && past_end)) emit = check(bi.end_nib[(size_t)hb]);
    }
    if (!emit) return false;
    h = cof_hit{};
    h.chrom_pos = chrom_pos;
    h.chrom = ch.chrom;
    h.pattern = r.pattern;
    h.mismatches = (uint16_t)mm;
    a = HitAux{r.pos, view_end};
    return true;
}

void convert_plain(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                   RawView<PlainRaw> raw, HitVec& out_all, AuxVec& aux_all) {
    parallel_convert(raw.size(), out_all, aux_all, [&](size_t b0, size_t b1, cof_hit* out, HitAux* aux) -> size_t {
        size_t k = 0;
        cof_hit h;
        HitAux a;
        for (size_t ri = b0; ri < b1; ri++) {
            if (!place_plain_hit(g, q, bs, bi, raw[ri], h, a)) continue;
            if (out) {
                out[k] = h;
                aux[k] = a;
            }
            k++;
        }
        return k;
    });
}

uint64_t convert_plain_into(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                            RawView<PlainRaw> raw, cof_hit* out) {
    uint64_t k = 0;
    cof_hit h;
    HitAux a;
    for (size_t ri = 0; ri < raw.size(); ri++) {
        if (!place_plain_hit(g, q, bs, bi, raw[ri], h, a)) continue;
        if (out) out[k] = h;
        k++;
    }
    return k;
}

void convert_bulge(const Genome& g, const Query& q, const std::vector<Batch>& bs, const BatchIndex& bi,
                   RawView<BulgeRaw> raw, HitVec& out_all, AuxVec& aux_all) {
    parallel_convert(raw.size(), out_all, aux_all, [&](size_t b0, size_t b1, cof_hit* out, HitAux* aux) -> size_t {
    size_t k = 0;
    for (size_t ri = b0; ri < b1; ri++) {
        const BulgeRaw& r = raw[ri];
        const Pattern& p = q.pats[r.pattern];
        uint64_t span = (uint64_t)p.L + r.db - r.rb;
        uint64_t j = r.start + span - 1;
        // batch whose scanned end range holds j
        size_t b = (size_t)(std::upper_bound(bi.active_begin.begin(), bi.active_begin.end(), j) - bi.active_begin.begin());
        if (b == 0) continue;
        b -= 1;
        size_t c = g.chunk_of(r.start);
        if (c < bs[b].first) c = bs[b].first;
        const Chunk& ch = g.chunks[c];
        uint64_t chrom_pos = ch.start + (r.start - ch.nib_begin());
        bool is_end_chrom = c + 1 == bs[b].last || g.chunks[c + 1].start == 0;
        bool past_end = chrom_pos + span > ch.end;
        if (is_end_chrom && past_end) continue;
        if (out) {
            cof_hit h{};
            h.chrom_pos = chrom_pos;
            h.chrom = ch.chrom;
            h.pattern = r.pattern;
            h.mismatches = r.mm;
            h.dna_bulge = r.db;
            h.rna_bulge = r.rb;
            h.n_ops = r.n_ops;
            ops_reverse(r.ops, r.n_ops, h.ops);      // device stack order -> alignment order
            out[k] = h;
            aux[k] = HitAux{r.start, bi.end_nib[b]};
        }
        k++;
    }
    return k;
    });
}

// Match.dna_seq / Match.rna_seq of a final hit (chrom_chunk.rs:19-29).
// Plain path (convert_matches, search.rs:1456-1475): the window of the
// N-masked batch view decoded (genome N and bytes past the view as NUL) and
// the pattern decoded; bulge path (convert_matches_myers, search.rs:1352-1372
// with traceback.rs:250-290): the aligned strings with '-' at bulges.
// Reverse-strand hits reverse-complement both (bit4ops.rs:207-232).
void hit_seqs(const Genome& g, const Query& q, const cof_hit& h, const HitAux& a, std::string& dna, std::string& rna) {
    const Pattern& p = q.pats[h.pattern];
    rna.clear();
    dna.clear();
    if (!p.bulge) {
        for (uint32_t k = 0; k < p.L; k++) {
            uint64_t i = a.nib + k;
            uint8_t v = i < a.view_end ? nib(g.packed.data(), i) : 0;
            if (v == 0xF) v = 0;
            dna.push_back(kCodec.dec[v]);
            rna.push_back(kCodec.dec[p.b4[k]]);
        }
    } else {
        uint64_t t = a.nib;
        uint32_t pi = 0;
        for (uint32_t k = 0; k < h.n_ops; k++) {
            uint32_t op = (uint32_t)(h.ops[k / 32] >> (2 * (k % 32))) & 3;
            if (op <= 1) { rna.push_back(kCodec.dec[p.b4[pi++]]); dna.push_back(kCodec.dec[nib(g.packed.data(), t++)]); }
            else if (op == 3) { rna.push_back(kCodec.dec[p.b4[pi++]]); dna.push_back('-'); }
            else { rna.push_back('-'); dna.push_back(kCodec.dec[nib(g.packed.data(), t++)]); }
        }
    }
    if (!p.forward) {
        std::reverse(dna.begin(), dna.end());
        for (auto& c : dna) c = comp_char(c);
        std::reverse(rna.begin(), rna.end());
        for (auto& c : rna) c = comp_char(c);
    }
}

// Rows of main.rs:130-247.
void write_tsv(const Genome& g, const Query& q, const HitVec& hits, const AuxVec& aux, FILE* f) {
    fputs("#Bulge type\tcrRNA\tDNA\tChromosome\tPosition\tDirection\tMismatches\tBulge Size\n", f);
    std::string line, rna, dna;
    for (size_t hi = 0; hi < hits.size(); hi++) {
        const cof_hit& h = hits[hi];
        const Pattern& p = q.pats[h.pattern];
        hit_seqs(g, q, h, aux[hi], dna, rna);
        const char* type = "X";
        uint32_t bsize = 0;
        if (!p.bulge) {
            for (size_t k = 0; k < dna.size(); k++) {         // main.rs:198-217
                char& d = dna[k];
                bool rn = is_n_char(rna[k]);
                if (d == 0) d = rn ? 'N' : 'n';
                else if (rn) {}
                else if (is_n_char(d)) d = 'n';
                else if (!cmp_chars((uint8_t)d, (uint8_t)rna[k])) d = (char)(d | 0x20);
            }
        } else {
            uint32_t tot = h.dna_bulge + h.rna_bulge;
            type = tot == 0 ? "X" : h.dna_bulge > 0 ? "DNA" : "RNA";
            bsize = tot;
        }
        uint64_t shift = 0;
        if (!p.forward) {
            uint64_t a = p.pam_len_out > 0 ? p.pam_len_out - 1 : 0;
            shift = a > h.dna_bulge ? a - h.dna_bulge : 0;
        }
        line.clear();
        line += type; line += '\t';
        line += rna; line += '\t';
        line.append(dna.data(), dna.size()); line += '\t';
        line += g.chrom_names[h.chrom]; line += '\t';
        line += std::to_string(h.chrom_pos + shift); line += '\t';
        line += p.forward ? '+' : '-'; line += '\t';
        line += std::to_string(h.mismatches); line += '\t';
        line += std::to_string(bsize); line += '\n';
        fwrite(line.data(), 1, line.size(), f);
    }
}

}  // namespace cof
