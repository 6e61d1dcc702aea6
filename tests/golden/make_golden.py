"""Writes the committed golden vectors: oracle rows for small seeded synthetic genomes.

Each case is a genome spec (seed, chromosome lengths, line width, N gaps,
repeats, IUPAC letters, planted near-hits; generated by libcofsynth, the same
seeded generator the tests and the benchmark use) plus a reference-format
query.  The oracle (oracle/cof_ref.cpp, a restatement of the reference's CPU
path) is run on the FASTA and its sorted rows are stored gzip-compressed as
tests/golden/<case>.tsv.gz.  tests/test_golden.py re-runs this script into a
temporary folder and requires byte-identical output, and the GPU tests compare
libcof's rows with these files.

    python tests/golden/make_golden.py [out_dir]
"""
import gzip
import os
import random
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cas-offinder-rust-old_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import cofpy  # noqa: E402


def _guides(seed, n, length):
    rng = random.Random(seed)
    return ["".join(rng.choice("ACGT") for _ in range(length)) for _ in range(n)]


def _case(name, seed, lens, pam, three, glen, n_guides, M, D=0, R=0, line_width=60, n_frac=0.02, gap=(50, 3000),
          repeats=None, iupac=0.0, plants_per_guide=12, batch_nucl=1 << 22):
    guides = _guides(seed * 7 + 1, n_guides, glen)
    if three:
        filt = "N" * glen + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, glen, max(M, 1), plants_per_guide) for g in guides]
    else:
        filt = pam + "N" * glen
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), glen, max(M, 1), plants_per_guide) for g in guides]
    return {"name": name, "genome": {"seed": seed, "lens": lens, "line_width": line_width, "n_frac": n_frac,
                                     "gap": gap, "repeats": repeats, "iupac": iupac, "plants": plants},
            "filter": filt, "patterns": pats, "M": M, "D": D, "R": R, "batch_nucl": batch_nucl}


def _mixed_case(name, seed, lens, n_guides, max_bulge_edits=6, repeats=(0.10, 20, 300, 0.10), batch_nucl=1 << 22):
    """One query mixing what one reference run cannot (cli_utils.rs:97-108: one filter, length and M per run):
    the batched workload's guide mix (length U{17..30}, PAM in {NRG, NGG, NNGRRT, TTTN}, M U{0..6}, 20% with
    D, R in U{1,2}), except that a bulge guide draws M from 0..max_bulge_edits - D - R: a 17-nt guide at
    M6 D2 R2 aligns ~3e7 ways per Mbp of this genome, too many rows for a fixture.  The expected rows are the
    union of one oracle run per (filter, M, D, R) set."""
    rng = random.Random(seed)
    guides = []
    while len(guides) < n_guides:
        glen = rng.randint(17, 30)
        pam = rng.choice(["NRG", "NGG", "NNGRRT", "TTTN"])
        g = "".join(rng.choice("ACGT") for _ in range(glen))
        mm = rng.randint(0, 6)
        d, r = (rng.randint(1, 2), rng.randint(1, 2)) if rng.random() < 0.2 else (0, 0)
        if d + r:
            mm = min(mm, max_bulge_edits - d - r)
        if pam == "TTTN":
            guides.append(("NNNN" + g, pam + "N" * glen, mm, d, r, (pam + g, 4, glen, min(mm, 3) or 1, 4)))
        else:
            guides.append((g + "N" * len(pam), "N" * glen + pam, mm, d, r, (g + pam, 0, glen, min(mm, 3) or 1, 4)))
    plants = [x[5] for x in guides]
    return {"name": name, "genome": {"seed": seed, "lens": lens, "line_width": 60, "n_frac": 0.02, "gap": (50, 3000),
                                     "repeats": repeats, "iupac": 0.0, "plants": plants},
            "mixed": [x[:5] for x in guides], "batch_nucl": batch_nucl}


CASES = [
    _case("plain_nrg20_m4", 501, [300_000, 170_001], "NRG", True, 20, 4, 4, iupac=0.0005),
    _case("plain_tttn20_m3_w61", 502, [250_000, 90_000], "TTTN", False, 20, 4, 3, line_width=61),
    _case("plain_nngrrt21_m5_seams", 503, [65520 * 3 + 5, 131_041], "NNGRRT", True, 21, 3, 5,
          batch_nucl=1 << 17),
    _case("plain_alln24_m9_iupac", 508, [200_000, 100_003], "", True, 24, 2, 9, line_width=61, iupac=0.001),
    _case("bulge_ngg20_m3_d1r1", 504, [150_000, 60_001], "NGG", True, 20, 2, 3, 1, 1, repeats=(0.1, 4, 300, 0.1)),
    _case("bulge_tttn20_m2_d1r1", 505, [120_000, 50_000], "TTTN", False, 20, 2, 2, 1, 1, iupac=0.0005),
    _case("bulge_nrg20_m4_d2r2", 506, [40_000, 20_001], "NRG", True, 20, 2, 4, 2, 2, plants_per_guide=2),
    _case("bulge_nrg28_m5_d2r1_w33", 507, [90_000, 30_007], "NRG", True, 28, 2, 5, 2, 1),
    _mixed_case("mixed_len17_30_4pams_m0_6_bulge20pct", 509, [200_000, 150_001, 65520 + 4000], 64, max_bulge_edits=7),
]


def run_sets(case):
    """(filter, M, D, R, patterns) per reference run of a case: one run, or one per set of a mixed case."""
    if "mixed" not in case:
        return [(case["filter"], case["M"], case["D"], case["R"], case["patterns"])]
    sets = {}
    for pat, filt, mm, d, r in case["mixed"]:
        sets.setdefault((filt, mm, d, r), []).append(pat)
    return [k + (v,) for k, v in sets.items()]


def make_genome(spec):
    g = spec
    return cofpy.SynthGenome(g["seed"], g["lens"], line_width=g["line_width"], n_frac=g["n_frac"],
                             gap=tuple(g["gap"]), repeats=tuple(g["repeats"]) if g["repeats"] else None,
                             iupac_frac=g["iupac"], plants=[tuple(p) for p in g["plants"]])


def input_text(run, fasta):
    filt, M, D, R, pats = run
    line2 = filt + (f" {D} {R}" if D or R else "")
    return f"{fasta}\n{line2}\n" + "".join(f"{p} {M}\n" for p in pats)


def main(out_dir=HERE):
    import oracle
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as d:
        for case in CASES:
            fa = os.path.join(d, case["name"] + ".fa")
            make_genome(case["genome"]).write_fasta(fa)
            tsv = os.path.join(d, case["name"] + ".tsv")
            rows = []
            for run in run_sets(case):
                oracle.run(input_text(run, fa), tsv, batch_nucl=case["batch_nucl"])
                rows += cofpy.read_rows(tsv)
            rows.sort()
            data = ("\n".join(rows) + "\n").encode() if rows else b""
            with open(os.path.join(out_dir, case["name"] + ".tsv.gz"), "wb") as f:
                f.write(gzip.compress(data, compresslevel=9, mtime=0))
            print(f"{case['name']}: {len(rows)} rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
