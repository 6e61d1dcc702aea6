// Query preparation: validation (cas-offinder-cli/src/cli_utils.rs:37-125 and
// the asserts of search.rs:1502-1507), forward + reverse-complement patterns
// (main.rs:253-259), effective filters (search.rs:1213-1225) and the
// output position shift length (main.rs:80-91).
#include <map>
#include <tuple>

#include "cof_internal.hpp"

namespace cof {

static bool mixed_base(const std::string& s) {
    for (unsigned char c : s) if (kCodec.enc_mixed[c] == 0) return false;
    return !s.empty();
}

Query prepare_query(const cof_query* q) {
    if (!q || q->n_guides == 0 || !q->guides || !q->filters || !q->max_mismatches)
        throw Error(COF_E_INVALID, "query needs at least one guide with filter and mismatch limit");
    Query out;
    out.n_guides = q->n_guides;
    out.batch_nucl = q->batch_nucl ? q->batch_nucl : (1ull << 30);
    out.flags = q->flags;
    if (out.batch_nucl % kChunkNucl != 0 || out.batch_nucl < 2 * kChunkNucl || out.batch_nucl > (1ull << 31))
        throw Error(COF_E_INVALID, "batch_nucl must be a multiple of 65536 in [131072, 2^31]");
    // The run a guide belongs to: guides of one reference run share filter
    // and limits; the position shift uses that run's first guide.
    std::map<std::tuple<std::string, uint32_t, uint32_t, uint32_t>, uint32_t> run_pam;
    std::vector<Pattern> fwd, rev;
    for (uint32_t i = 0; i < q->n_guides; i++) {
        if (!q->guides[i] || !q->filters[i]) throw Error(COF_E_INVALID, "null guide or filter");
        std::string gde = q->guides[i], flt = q->filters[i];
        if (!mixed_base(flt)) throw Error(COF_E_INVALID, "Pattern in input file needs to be a mixed base string");
        if (!mixed_base(gde)) throw Error(COF_E_INVALID, "Pattern in input file needs to be a mixed base string");
        if (gde.size() != flt.size()) throw Error(COF_E_INVALID, "All patters in input file must be same length");
        uint32_t M = q->max_mismatches[i];
        uint32_t D = q->max_dna_bulges ? q->max_dna_bulges[i] : 0;
        uint32_t R = q->max_rna_bulges ? q->max_rna_bulges[i] : 0;
        uint32_t L = (uint32_t)gde.size();
        bool bulge = D > 0 || R > 0;
        if (!bulge && M >= L)
            throw Error(COF_E_INVALID, "max mismatches must be below the pattern length (the reference panics on the zero tail)");
        if (bulge) {
            if (L > 64) throw Error(COF_E_UNSUPPORTED, "bulge search supports patterns up to 64 nt (myers.rs:33)");
            if (D > 16 || R > 16 || M > 64) throw Error(COF_E_UNSUPPORTED, "bulge limits above 16 are not supported");
            if (L + D > 96) throw Error(COF_E_UNSUPPORTED, "alignment longer than 96 ops");
        } else if (L > 256) {
            throw Error(COF_E_UNSUPPORTED, "plain search supports patterns up to 256 nt");
        }
        auto key = std::make_tuple(flt, M, D, R);
        if (!run_pam.count(key)) {
            uint32_t n = 0;
            for (auto it = gde.rbegin(); it != gde.rend() && is_n_char(*it); ++it) n++;
            run_pam[key] = n;
        }
        Pattern p;
        p.L = L; p.M = M; p.D = D; p.R = R;
        p.guide = i;
        p.bulge = bulge;
        p.pam_len_out = run_pam[key];
        p.ascii = gde;
        p.eff_filter = flt;
        p.forward = true;
        for (char c : gde) p.b4.push_back(kCodec.enc_mixed[(uint8_t)c]);
        Pattern r = p;
        r.ascii = revcomp(gde);
        r.eff_filter = revcomp(flt);
        r.forward = false;
        r.b4.clear();
        for (char c : r.ascii) r.b4.push_back(kCodec.enc_mixed[(uint8_t)c]);
        fwd.push_back(p);
        rev.push_back(r);
        if (bulge) out.any_bulge = true; else out.any_plain = true;
    }
    out.pats = fwd;
    out.pats.insert(out.pats.end(), rev.begin(), rev.end());
    return out;
}

}  // namespace cof
