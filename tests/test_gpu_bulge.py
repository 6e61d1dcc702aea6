"""Bulge search (DNA/RNA bulges) on the GPU vs the CPU oracle: sorted rows must be identical.

Starts from the reference's own bulge test (integration_bulge.rs:208-259:
N x25 + GG with 1 DNA / 1 RNA bulge, CPU rows == GPU rows, non-empty, 8
columns, type in {X, DNA, RNA}) and widens it with seeded synthetic genomes:
3' and 5' PAMs (PAM-last and PAM-first quick checks), bulge limits 0-2 per
type, N gaps and IUPAC letters, odd chunks and small virtual batches (window
floors and PAM limits at batch seams), and a mixed plain + bulge query.
"""
import os
import random

import pytest

import cofpy
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def oracle_rows(tmp_path, genome, filt_line, guides, mm, batch_nucl):
    in_text = f"{genome}\n{filt_line}\n" + "".join(f"{g} {mm}\n" for g in guides)
    oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=batch_nucl)
    return cofpy.read_rows(tmp_path / "ref.tsv")


def gpu_rows(tmp_path, genome, filt, guides, mms, ds, rs, batch_nucl, filters=None):
    s = cofpy.Search(guides, filters or [filt] * len(guides), mms, ds, rs, batch_nucl=batch_nucl)
    s.load_file(genome)
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    st = s.stats()
    s.close()
    return cofpy.read_rows(tmp_path / "gpu.tsv"), st


def test_reference_bulge_query(tmp_path):
    genome = os.path.join(GOLDEN, "upstream1000.fa")
    filt = "NNNNNNNNNNNNNNNNNNNNNNNNNGG"
    guide = "CCGTGGTTCAACATTTGCTTAGCANNN"
    want = oracle_rows(tmp_path, genome, filt + " 1 1", [guide], 5, 1 << 30)
    got, st = gpu_rows(tmp_path, genome, filt, [guide], [5], [1], [1], 1 << 30)
    assert got == want
    assert len(got) > 0
    for row in got:
        cols = row.split("\t")
        assert len(cols) == 8 and cols[0] in ("X", "DNA", "RNA")


CASES = [
    # pam, 3', glen, M, D, R
    ("NRG", True, 20, 4, 2, 2),
    ("NGG", True, 20, 3, 1, 0),
    ("NGG", True, 20, 3, 0, 1),
    ("TTTN", False, 20, 3, 1, 1),
    ("NNGRRT", True, 21, 3, 1, 2),
    ("NRG", True, 28, 5, 2, 1),     # W = 33: 64-bit window kernels
    ("NGG", True, 20, 3, 3, 2),     # D + R > 4: traceback diagonal masks in LDS, screen skipped (D > 2)
    ("TTTN", False, 20, 2, 1, 4),   # 5' PAM, R > 2
]


@pytest.mark.parametrize("pam,three,glen,M,D,R", CASES)
def test_synthetic_bulge(tmp_path, pam, three, glen, M, D, R):
    rng = random.Random(glen * 31 + M + D * 7 + R * 13)
    guides = ["".join(rng.choice("ACGT") for _ in range(glen)) for _ in range(2)]
    if three:
        filt = "N" * glen + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, glen, M, 15) for g in guides]
    else:
        filt = pam + "N" * glen
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), glen, M, 15) for g in guides]
    sg = cofpy.SynthGenome(77 + glen + M, [60_000, 41_000], n_frac=0.02, gap=(20, 400), iupac_frac=0.001,
                           repeats=(0.1, 4, 300, 0.1), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    want = oracle_rows(tmp_path, fa, f"{filt} {D} {R}", pats, M, 1 << 30)
    got, st = gpu_rows(tmp_path, fa, filt, pats, [M] * 2, [D] * 2, [R] * 2, 1 << 30)
    assert len(want) > 0
    assert got == want


@pytest.mark.parametrize("batch_nucl", [1 << 17, 3 << 16])
def test_bulge_batch_seams_odd_chunks(tmp_path, batch_nucl):
    guides = ["GGCCGACCTGTCGCTGACGC"]
    sg = cofpy.SynthGenome(5, [65520 * 2 + 40, 65521, 90_000], line_width=61, n_frac=0.01, gap=(10, 200),
                           plants=[(guides[0] + "TGG", 0, 20, 3, 40)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    want = oracle_rows(tmp_path, fa, "NNNNNNNNNNNNNNNNNNNNNRG 1 1", [g + "NNN" for g in guides], 3, batch_nucl)
    got, _ = gpu_rows(tmp_path, fa, "NNNNNNNNNNNNNNNNNNNNNRG", [g + "NNN" for g in guides], [3], [1], [1], batch_nucl)
    assert got == want and len(want) > 0


def test_mixed_plain_and_bulge_query(tmp_path):
    # one context holding a plain run and a bulge run == union of two reference runs
    g1, g2 = "GGCCGACCTGTCGCTGACGC", "ACCTGTCGCTGACGCGGCCG"
    sg = cofpy.SynthGenome(9, [120_000], plants=[(g1 + "AGG", 0, 20, 4, 20), (g2 + "CGG", 0, 20, 2, 20)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    f = "NNNNNNNNNNNNNNNNNNNNNRG"
    want = sorted(oracle_rows(tmp_path, fa, f, [g1 + "NNN"], 4, 1 << 30) +
                  oracle_rows(tmp_path, fa, f + " 1 1", [g2 + "NNN"], 2, 1 << 30))
    got, _ = gpu_rows(tmp_path, fa, f, [g1 + "NNN", g2 + "NNN"], [4, 2], [0, 1], [0, 1], 1 << 30)
    assert got == want


def test_count_only_equals_record_mode(tmp_path):
    # the device-side drop rule + checksum equals conversion on the host;
    # small chromosomes and 2-chunk batches put alignments across chromosome
    # ends and batch seams
    guides = ["GGCCGACCTGTCGCTGACGC", "TTGACCGATCGGATCCAGTT"]
    sg = cofpy.SynthGenome(21, [65520 + 30, 300, 65520 * 2, 45_000, 23], n_frac=0.01, gap=(5, 100),
                           repeats=(0.1, 3, 300, 0.1), plants=[(g + "GGG", 0, 20, 3, 25) for g in guides])
    pats = [g + "NNN" for g in guides]
    res = []
    for count_only in (False, True):
        s = cofpy.Search(pats, ["NNNNNNNNNNNNNNNNNNNNNRG"] * 2, [4, 4], [2, 2], [2, 2], batch_nucl=1 << 17,
                         count_only=count_only)
        s.load_synth(sg)
        n = s.search()
        st = s.stats()
        res.append((st.bulge_hits, st.bulge_checksum, n))
        s.close()
    assert res[0][0] > 1000
    assert res[0][:2] == res[1][:2]
    assert res[0][2] == res[0][0] and res[1][2] == 0


def test_set_query_reuses_resident_genome(tmp_path):
    # a re-queried context gives the same rows as a fresh one
    g1, g2 = "GGCCGACCTGTCGCTGACGC", "ACCTGTCGCTGACGCGGCCG"
    sg = cofpy.SynthGenome(13, [90_000, 30_000], plants=[(g1 + "AGG", 0, 20, 3, 20), (g2 + "TTTA", 0, 20, 2, 20)])
    f1, f2 = "NNNNNNNNNNNNNNNNNNNNNRG", "NNNNNNNNNNNNNNNNNNNNNNGRRT"
    a = cofpy.Search([g1 + "NNN"], [f1], [4], [1], [1])
    a.load_synth(sg)
    a.search()
    a.set_query([g2 + "NNNNNN"], [f2], [3])
    a.search()
    a.write_tsv(tmp_path / "a.tsv")
    b = cofpy.Search([g2 + "NNNNNN"], [f2], [3])
    b.load_synth(sg)
    b.search()
    b.write_tsv(tmp_path / "b.tsv")
    assert cofpy.read_rows(tmp_path / "a.tsv") == cofpy.read_rows(tmp_path / "b.tsv")
    a.close()
    b.close()


@pytest.mark.parametrize("glen,pam,D,R", [(25, "NNGRRT", 1, 2), (26, "NNGRRT", 0, 1), (27, "NGG", 2, 1),
                                          (58, "NGG", 3, 1), (59, "NRG", 2, 2)])
def test_window_width_edges(tmp_path, glen, pam, D, R):
    # L + D = 32 and 64 are the last windows of the 32- and 64-bit kernels
    rng = random.Random(glen * 3 + D)
    g = "".join(rng.choice("ACGT") for _ in range(glen))
    filt = "N" * glen + pam
    sg = cofpy.SynthGenome(31 + glen, [30_000, 12_000], n_frac=0.01, gap=(5, 50),
                           plants=[(g + pam, 0, glen, 3, 25)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    M = 3 if glen < 40 else 1
    want = oracle_rows(tmp_path, fa, f"{filt} {D} {R}", [g + "N" * len(pam)], M, 1 << 22)
    got, _ = gpu_rows(tmp_path, fa, filt, [g + "N" * len(pam)], [M], [D], [R], 1 << 22)
    assert len(want) > 0
    assert got == want


def test_budget_screen_keeps_every_alignment(tmp_path):
    # The budget screen between the Myers filter and the traceback drops only
    # candidates with no alignment inside the M / D / R caps.  A mixed query
    # covers every screen shape: M 0-8 (M above 6 passes unchecked), D and R
    # 0-2, a 5' PAM, and 28-nt guides with D = 2 (64-bit windows); rows,
    # counts and checksums must not change with the screen on or off.  (The
    # M = 8 guide, passed unchecked, supplies most of this query's candidates.)
    rng = random.Random(2024)
    pats, filts, ms, ds, rs, plants = [], [], [], [], [], []
    for k in range(12):
        glen = 28 if k % 6 == 5 else 20
        g = "".join(rng.choice("ACGT") for _ in range(glen))
        m, d, r = [0, 2, 3, 4, 5, 6, 8, 4, 3, 4, 2, 4][k], k % 3, (k // 3) % 3
        if k == 7:
            pats.append("TTTN" + g)
            filts.append("TTTN" + "N" * glen)
            plants.append(("TTTA" + g, 4, glen, m, 12))
        else:
            pats.append(g + "NNN")
            filts.append("N" * glen + "NRG")
            plants.append((g + "AGG", 0, glen, m, 12))
        ms.append(m)
        ds.append(d)
        rs.append(r)
    sg = cofpy.SynthGenome(4242, [150_000, 70_001], n_frac=0.01, gap=(20, 300), iupac_frac=0.001,
                           repeats=(0.15, 6, 300, 0.1), plants=plants)
    res = {}
    for screen in (1, 2):
        for count_only in (False, True):
            s = cofpy.Search(pats, filts, ms, ds, rs, batch_nucl=1 << 17, count_only=count_only)
            s.set_tuning(bulge_screen=screen)
            assert s.tuning()["bulge_screen"] == screen
            s.load_synth(sg)
            s.search()
            st = s.stats()
            rows = None
            if not count_only:
                s.write_tsv(tmp_path / "gpu.tsv")
                rows = cofpy.read_rows(tmp_path / "gpu.tsv")
            res[(screen, count_only)] = (rows, st.bulge_hits, st.bulge_checksum, st.n_candidates, st.n_traced)
            s.close()
    on, off = res[(1, False)], res[(2, False)]
    assert len(on[0]) > 100
    assert on[0] == off[0]
    assert on[1:4] == off[1:4]
    assert res[(1, True)][1:4] == res[(2, True)][1:4] == on[1:4]
    assert off[4] == off[3]
    assert 0 < on[4] < off[3]


def test_budget_screen_single_group_vs_oracle(tmp_path):
    # one reference-format group (the oracle's scope) with the screen on and
    # off; with M4 D2 R2 on a repeat-rich genome most Myers survivors have no
    # alignment within the caps and are screened out
    rng = random.Random(77)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(3)]
    sg = cofpy.SynthGenome(99, [90_000, 30_001], n_frac=0.02, gap=(20, 300), repeats=(0.2, 4, 300, 0.1),
                           plants=[(g + "TGG", 0, 20, 4, 20) for g in guides])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    pats = [g + "NNN" for g in guides]
    filt = "NNNNNNNNNNNNNNNNNNNNNRG"
    want = oracle_rows(tmp_path, fa, filt + " 2 2", pats, 4, 1 << 30)
    for screen in (1, 2):
        s = cofpy.Search(pats, [filt] * 3, [4] * 3, [2] * 3, [2] * 3, batch_nucl=1 << 30)
        s.set_tuning(bulge_screen=screen)
        s.load_file(fa)
        s.search()
        s.write_tsv(tmp_path / "gpu.tsv")
        st = s.stats()
        s.close()
        assert cofpy.read_rows(tmp_path / "gpu.tsv") == want
        assert (st.n_traced < st.n_candidates // 2) if screen == 1 else (st.n_traced == st.n_candidates)
    assert len(want) > 0
