// cof-cli: drop-in for cas-offinder-cli (cas-offinder-cli/src/main.rs:55-305)
// on one MI355X.  Usage: cof-cli {input|-} {C|G|A} {output|-}
// The device letter is one character (cli_utils.rs:4-16) and picks the
// reference path whose output the run reproduces: C the CPU path's 4 Mnt
// batches, G and A the GPU path's 1 Gnt batches (search.rs:21-50); the search
// itself always runs on HIP device 0 (COF_DEVICE overrides, COF_BATCH_NUCL
// overrides the batch size).  Input parsing follows
// cas-offinder-cli/src/cli_utils.rs:37-148.  A run that writes to a file also
// writes the run summary <output>.log (log_writer.rs:27-53, main.rs:277-304).
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "cof.h"

namespace {

std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

bool parse_u32(const std::string& s, uint32_t& v) {
    if (s.empty() || s.size() > 10) return false;
    uint64_t x = 0;
    for (char c : s) {
        if (c < '0' || c > '9') return false;
        x = x * 10 + (uint64_t)(c - '0');
    }
    if (x > 0xFFFFFFFFull) return false;
    v = (uint32_t)x;
    return true;
}

int die(const char* msg) {
    fprintf(stderr, "error: %s\n", msg);
    return 1;
}

// Usage text of main.rs:17-46; the device list names the HIP devices the
// library sees.
void usage() {
    fprintf(stderr,
            "\nCas-OFFinder 2 - cof-cli (MI355X)\n\n"
            "Usage: cof-cli [options] {input_filename|-} {C|G|A}[device_id(s)] {output_filename|-}\n"
            "(C: using CPUs, G: using GPUs, A: using accelerators)\n\n"
            "Example input file:\n/var/chromosomes/human_hg19\nNNNNNNNNNNNNNNNNNNNNNRG\n"
            "GGCCGACCTGTCGCTGACGCNNN 5\nCGCCAGCGTCAGCGACAGGTNNN 5\n"
            "ACGGCGCCAGCGTCAGCGACNNN 5\nGTCGCTGACGCTGGCGCCGTNNN 5\n\n"
            "Available device list:\n");
    int n = 0;
    if (cof_device_count(&n) == COF_OK)
        for (int d = 0; d < n; d++) {
            char name[256];
            if (cof_device_name(d, name, sizeof name) == COF_OK) fprintf(stderr, "Type: GPU, ID: %d, <%s>\n", d, name);
        }
}

// Run summary with the fields and layout of log_writer.rs:27-53.
struct RunLog {
    std::string genome_path, search_filter, device_label, algorithm;
    uint64_t genome_size = 0, n_matches = 0;
    size_t n_patterns = 0;
    uint32_t max_mismatches = 0, max_dna_bulges = 0, max_rna_bulges = 0;
    double total_elapsed_secs = 0;
};

bool write_log(const std::string& path, const RunLog& l) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "=== Cas-OFFinder Rust ===\nRun date (unix timestamp): %llu\n\n", (unsigned long long)time(nullptr));
    fprintf(f, "Algorithm: %s\nDevice: %s\n\nInput:\n", l.algorithm.c_str(), l.device_label.c_str());
    fprintf(f, "  Genome: %s\n  Genome size: %llu bytes\n  Patterns: %zu\n  Search filter: %s\n",
            l.genome_path.c_str(), (unsigned long long)l.genome_size, l.n_patterns, l.search_filter.c_str());
    fprintf(f, "  Max mismatches: %u\n  Max DNA bulges: %u\n  Max RNA bulges: %u\n\n", l.max_mismatches,
            l.max_dna_bulges, l.max_rna_bulges);
    fprintf(f, "Results:\n  Matches found: %llu\n  Total elapsed: %.3fs\n", (unsigned long long)l.n_matches,
            l.total_elapsed_secs);
    return fclose(f) == 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        usage();
        return argc < 2 ? 0 : 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::string dev = argv[2];
    if (dev.size() != 1 || (dev[0] != 'C' && dev[0] != 'G' && dev[0] != 'A'))
        return die("2nd argument must be one of {C|G|A}");
    std::string text;
    {
        FILE* f = strcmp(argv[1], "-") ? fopen(argv[1], "rb") : stdin;
        if (!f) return die("cannot open input file");
        char b[65536];
        size_t n;
        while ((n = fread(b, 1, sizeof b, f)) > 0) text.append(b, n);
        if (f != stdin) fclose(f);
    }
    std::vector<std::string> lines;
    size_t s = 0;
    for (size_t i = 0; i <= text.size(); i++) {
        if (i == text.size() || text[i] == '\n') {
            if (i == text.size() && s == i) break;
            std::string ln = text.substr(s, i - s);
            if (!ln.empty() && ln.back() == '\r') ln.pop_back();
            lines.push_back(ln);
            s = i + 1;
        }
    }
    if (lines.size() < 2) return die("Input file must contain at least 3 lines");
    auto parts = split_ws(lines[1]);
    std::string filter;
    uint32_t D = 0, R = 0;
    if (parts.size() == 1) filter = parts[0];
    else if (parts.size() == 3) {
        filter = parts[0];
        if (!parse_u32(parts[1], D)) return die("bulge_dna must be a non-negative integer");
        if (!parse_u32(parts[2], R)) return die("bulge_rna must be a non-negative integer");
    } else return die("2nd line must be: <search_filter> [<bulge_dna> <bulge_rna>]");
    std::vector<std::string> guides;
    uint32_t M = 0;
    bool have_mm = false, have_info = false, using_info = false;
    for (size_t li = 2; li < lines.size(); li++) {
        auto lp = split_ws(lines[li]);
        if (lp.size() != 2 && lp.size() != 3)
            return die("Pattern line must have following elements: <patterns> <mismatches> [<label>]");
        uint32_t cm;
        if (!parse_u32(lp[1], cm)) return die("2nd element of each pattern in mismatches must be an unsigned integer");
        if (!have_mm) { M = cm; have_mm = true; }
        if (!have_info) { using_info = lp.size() == 3; have_info = true; }
        if (cm != M) return die("In this version of cas-offinder, all mismatches on every line of input file must be the same");
        if (using_info && lp.size() < 3)
            return die("Pattern lines in input file must be consistently have either 2 or 3 elements, no mixing and matching allowed");
        if (lp[0].size() != filter.size()) return die("All patters in input file must be same length");
        guides.push_back(lp[0]);
    }
    if (!have_mm) return die("Input file must contain at least 1 pattern line");
    std::vector<const char*> gp, fp;
    std::vector<uint32_t> mm(guides.size(), M), dd(guides.size(), D), rr(guides.size(), R);
    for (auto& g : guides) { gp.push_back(g.c_str()); fp.push_back(filter.c_str()); }
    cof_query q{};
    q.n_guides = (uint32_t)guides.size();
    q.guides = gp.data();
    q.filters = fp.data();
    q.max_mismatches = mm.data();
    q.max_dna_bulges = dd.data();
    q.max_rna_bulges = rr.data();
    q.batch_nucl = dev[0] == 'C' ? (1ull << 22) : (1ull << 30);   // search.rs:21-22, 44-50
    if (const char* bn = getenv("COF_BATCH_NUCL")) q.batch_nucl = strtoull(bn, nullptr, 10);
    int device = 0;
    if (const char* d = getenv("COF_DEVICE")) device = atoi(d);
    cof_ctx* ctx = nullptr;
    cof_status st = cof_open(device, &q, &ctx);
    if (st != COF_OK) return die(cof_strerror(st));
    st = cof_load_genome_file(ctx, lines[0].c_str());
    if (st != COF_OK) {
        fprintf(stderr, "error: %s: %s\n", cof_strerror(st), cof_last_error(ctx));
        cof_close(ctx);
        return 1;
    }
    uint64_t n = 0;
    st = cof_search(ctx, nullptr, 0, &n);
    if (st != COF_OK && st != COF_E_OVERFLOW) {
        fprintf(stderr, "error: %s: %s\n", cof_strerror(st), cof_last_error(ctx));
        cof_close(ctx);
        return 1;
    }
    st = cof_write_tsv(ctx, argv[3]);
    cof_stats stats{};
    cof_get_stats(ctx, &stats);
    cof_close(ctx);
    if (st != COF_OK) return die(cof_strerror(st));
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "Completed in %gs (search %g ms, %llu rows)\n", secs, stats.search_ms, (unsigned long long)n);
    if (strcmp(argv[3], "-") != 0) {
        RunLog l;
        l.genome_path = lines[0];
        struct stat sb;
        l.genome_size = stat(lines[0].c_str(), &sb) == 0 ? (uint64_t)sb.st_size : 0;
        l.n_patterns = guides.size();
        l.search_filter = filter;
        l.max_mismatches = M;
        l.max_dna_bulges = D;
        l.max_rna_bulges = R;
        l.device_label = dev[0] == 'C' ? "CPU" : dev[0] == 'G' ? "GPU" : "Accelerator";
        l.algorithm = (D > 0 || R > 0) ? "Myers bit-parallel" : "popcount (legacy)";
        l.n_matches = n;
        l.total_elapsed_secs = secs;
        std::string lp = std::string(argv[3]) + ".log";
        if (!write_log(lp, l)) fprintf(stderr, "warning: failed to write log file '%s'\n", lp.c_str());
    }
    return 0;
}
