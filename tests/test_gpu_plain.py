"""Plain mismatch scan on the GPU vs the CPU oracle: sorted rows must be identical.

Cases follow the reference's own tests (integration_bulge.rs:156-202: the
117-row query, CPU == GPU on sorted rows) and widen them with seeded synthetic
genomes covering the PAM shapes of the benchmark grid, N gaps, IUPAC genome
letters, odd FASTA line widths (odd chunks with padding nibbles) and small
virtual batches (many batch seams).
"""
import os

import pytest

import cofpy
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def run_both(tmp_path, genome, filt, guides, mm, batch_nucl=1 << 22):
    in_text = f"{genome}\n{filt}\n" + "".join(f"{g} {mm}\n" for g in guides)
    ref_out = tmp_path / "ref.tsv"
    oracle.run(in_text, ref_out, batch_nucl=batch_nucl)
    want = cofpy.read_rows(ref_out)
    f = filt.split()[0]
    s = cofpy.Search(guides, [f] * len(guides), [mm] * len(guides), batch_nucl=batch_nucl)
    s.load_file(genome)
    s.search()
    out = tmp_path / "gpu.tsv"
    s.write_tsv(out)
    got = cofpy.read_rows(out)
    st = s.stats()
    s.close()
    return got, want, st


def test_reference_117_rows(tmp_path):
    # integration_bulge.rs:159-175: filter N x24, two guides at 11 mismatches -> 117 rows
    for genome in ("upstream1000.fa", "upstream1000.2bit"):
        got, want, _ = run_both(tmp_path, os.path.join(GOLDEN, genome), "N" * 24,
                                ["CCGTGGTTCAACATTTGCTTAGCA", "GATGTTGGTAAGTGGGATATGGCA"], 11)
        assert len(want) == 117
        assert got == want


def _guides(rng, n, length):
    return ["".join(rng.choice("ACGT") for _ in range(length)) for _ in range(n)]


PAMS = [("NRG", True), ("NNGRRT", True), ("TTTN", False)]


@pytest.mark.parametrize("glen", [20, 25])
@pytest.mark.parametrize("pam,three_prime", PAMS)
@pytest.mark.parametrize("mm", [3, 6])
def test_synthetic_pam_grid(tmp_path, glen, pam, three_prime, mm):
    import random
    rng = random.Random(glen * 100 + mm * 7 + len(pam))
    guides = _guides(rng, 6, glen)
    if three_prime:
        filt = "N" * glen + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, glen, mm, 25) for g in guides]
    else:
        filt = pam + "N" * glen
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), glen, mm, 25) for g in guides]
    sg = cofpy.SynthGenome(1000 + glen + mm, [300_000, 170_000, 61_000], n_frac=0.02, gap=(50, 2000),
                           plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    got, want, st = run_both(tmp_path, fa, filt, pats, mm)
    assert len(want) > 0
    assert got == want


def test_conversion_drops_across_slices(tmp_path):
    # 300 chromosomes of 3 kbp tight-packed: the device also scans windows
    # that run from one chromosome into the next, and the conversion drops
    # them (search.rs:1456-1475).  With ~25 k raw hits the converter works in
    # many slices, so the hits dropped near chromosome ends leave holes in
    # many of them, which are filled from the end of the hit array.
    import random
    rng = random.Random(31)
    guides = [g + "NNN" for g in _guides(rng, 4, 20)]
    sg = cofpy.SynthGenome(77, [3000] * 300)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    got, want, st = run_both(tmp_path, fa, "N" * 20 + "NGG", guides, 11)
    assert st.n_raw_hits >= 4096 and st.n_raw_hits > len(want)
    assert got == want


@pytest.mark.parametrize("batch_nucl", [1 << 17, 3 << 16])
def test_batch_seams(tmp_path, batch_nucl):
    # chunks of 65520 nt, chromosomes with tiny tails, batches of 2-3 chunks
    guides = ["GGCCGACCTGTCGCTGACGC", "ACCTGTCGCTGACGCGGCCG"]
    plants = [(g + "AGG", 0, 20, 5, 60) for g in guides]
    sg = cofpy.SynthGenome(42, [65520 * 3 + 5, 65520 + 1, 131_041, 65520 * 2], n_frac=0.01, gap=(20, 500),
                           plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    got, want, _ = run_both(tmp_path, fa, "NNNNNNNNNNNNNNNNNNNNNRG", [g + "NNN" for g in guides], 6,
                            batch_nucl=batch_nucl)
    assert got == want


def test_odd_line_width_and_iupac(tmp_path):
    # odd line width -> odd-sized chunks whose padding nibble sits inside the
    # tight-packed layout; IUPAC genome letters exercise the exact path
    guides = ["GATGTTGGTAAGTGGGATATGGCA"]
    sg = cofpy.SynthGenome(9, [200_001, 77_777], line_width=61, n_frac=0.02, gap=(10, 300), iupac_frac=0.002,
                           plants=[(guides[0], 0, 24, 8, 80)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    for batch in (1 << 17, 1 << 22):
        got, want, _ = run_both(tmp_path, fa, "N" * 24, guides, 9, batch_nucl=batch)
        assert len(want) > 0
        assert got == want
    # IUPAC in guides and filter
    got, want, _ = run_both(tmp_path, fa, "NNNNNNNNNNNNNNNNNNNNNNRS", ["GATGTTGGTRAGTGGGATATGGNN"], 5)
    assert got == want


def test_overflow_retry_and_api(tmp_path):
    guides = ["GGCCGACCTGTCGCTGACGC"]
    sg = cofpy.SynthGenome(5, [250_000], plants=[(guides[0] + "TGG", 0, 20, 2, 30)])
    s = cofpy.Search([g + "NNN" for g in guides], ["NNNNNNNNNNNNNNNNNNNNNRG"], [4])
    s.load_synth(sg)
    n = s.search()
    assert n >= 30
    hits = s.hits()
    assert len(hits) == n
    # a too-small buffer reports the needed count and COF_E_OVERFLOW
    import ctypes as C
    buf = (cofpy.CofHit * 2)()
    need = C.c_uint64()
    st = cofpy.lib().cof_search(s.h, buf, 2, C.byref(need))
    assert st == cofpy.COF_E_OVERFLOW and need.value == n
    s.close()


@pytest.mark.parametrize("case", ["shared_n", "iupac_guide", "window32", "window33", "n_padded"])
def test_plane_window_shapes(tmp_path, case):
    # The scan compares 2-bit planes when a group's patterns are plain bases on
    # a window of at most 32 positions (positions that are N in every pattern
    # left out) and the tile holds only plain bases; other groups and tiles use
    # the bit4 compare.  Each case puts one query on either side of that line;
    # the genome mixes clean tiles with N gaps and odd chunks (zero padding).
    import random
    rng = random.Random(len(case) * 31)

    def g(n):
        return "".join(rng.choice("ACGT") for _ in range(n))

    if case == "shared_n":          # an N at guide position 5 in every pattern: left out of the window
        guides = [g(5) + "N" + g(14) for _ in range(9)]
        filt, pats, mm = "N" * 21 + "RG", [x + "NNN" for x in guides], 3
    elif case == "iupac_guide":     # one pattern with R: the whole group compares bit4
        guides = [g(20) for _ in range(8)] + [g(7) + "R" + g(12)]
        filt, pats, mm = "N" * 21 + "RG", [x + "NNN" for x in guides], 4
    elif case == "window32":        # 32-position window, L = 35
        guides = [g(32) for _ in range(5)]
        filt, pats, mm = "N" * 33 + "RG", [x + "NNN" for x in guides], 12
    elif case == "window33":        # 33 positions: bit4
        guides = [g(33) for _ in range(5)]
        filt, pats, mm = "N" * 34 + "RG", [x + "NNN" for x in guides], 12
    else:                           # 5' PAM, guides padded with N on both sides
        guides = [g(18) for _ in range(11)]
        filt, pats, mm = "TTTN" + "N" * 20, ["NNNN" + x + "NN" for x in guides], 4
    if case == "n_padded":
        plants = [("TTTA" + x + "AC", 4, 18, 2, 20) for x in guides]
    else:                           # guide letters sampled from their IUPAC sets, PAM AGG
        plants = [(x + "AGG", 0, len(x), 2, 20) for x in guides]
    sg = cofpy.SynthGenome(77 + len(case), [260_001, 131_071, 9_999], line_width=61, n_frac=0.02, gap=(30, 3000),
                           plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    got, want, st = run_both(tmp_path, fa, filt, pats, mm)
    assert len(want) > 0
    assert got == want


def test_async_search_matches_sync(tmp_path):
    # cof_search_async / cof_search_wait on two interleaved contexts give the
    # hits of cof_search; a context with a search in flight refuses other calls
    guides = ["GGCCGACCTGTCGCTGACGC", "ACCTGTCGCTGACGCGGCCG"]
    sg = cofpy.SynthGenome(7, [300_000, 90_001], n_frac=0.01, gap=(20, 500),
                           plants=[(g + "AGG", 0, 20, 4, 40) for g in guides])
    pats = [g + "NNN" for g in guides]
    a = cofpy.Search(pats, ["NNNNNNNNNNNNNNNNNNNNNRG"] * 2, [5, 5])
    b = cofpy.Search(pats, ["NNNNNNNNNNNNNNNNNNNNNRG"] * 2, [3, 3])
    a.load_synth(sg)
    b.load_synth(sg)
    want_a, want_b = a.search(), b.search()
    def key(s):     # hit order depends on the order workgroups append; the set does not
        return sorted((h.chrom, h.chrom_pos, h.pattern, h.mismatches) for h in s.hits())

    ha, hb = key(a), key(b)
    for _ in range(2):
        a.search_async()
        b.search_async()
        with pytest.raises(cofpy.CofError) as e:
            a.write_tsv(tmp_path / "x.tsv")
        assert e.value.status == cofpy.COF_E_INVALID
        assert a.wait() == want_a
        assert b.wait() == want_b
        assert key(a) == ha
        assert key(b) == hb
    with pytest.raises(cofpy.CofError):
        a.wait()                       # nothing in flight
    a.search_async()                   # left in flight: close waits for it
    a.close()
    b.close()
    assert want_a > want_b > 0


@pytest.mark.parametrize("glen,n,mms,threads,chunk,pam", [
    (20, 150, (0, 2, 4, 6, 7), 64, 0, "RG"),   # 10 blocks of 16 patterns per strand: passes of 8 + 2
    (20, 150, (1, 5), 256, 48, "RG"),          # slices of 3 blocks, four waves per tile
    (25, 37, (3, 5, 6), 64, 0, "RG"),          # 25-position window: positions 0-19 filter, planes decide
    (17, 20, (5,), 128, 16, "RG"),             # one block per slice
    (20, 40, (1, 3), 256, 0, "NN"),            # no PAM: every start is a window, lists of two half tiles
])
def test_matrix_core_pattern_blocks(tmp_path, glen, n, mms, threads, chunk, pam):
    # Plane groups on plain-base tiles are scored 16 windows x 16 patterns per
    # matrix-core product over window positions 0-19 (exact up to 20
    # positions, a filter checked on the planes beyond), with mixed mismatch
    # limits per pattern and filler patterns past the group's last one; rows
    # must equal the oracle's.
    import random
    rng = random.Random(glen * 1000 + n)
    guides = ["".join(rng.choice("ACGT") for _ in range(glen)) for _ in range(n)]
    mm = [mms[i % len(mms)] for i in range(n)]
    filt = "N" * (glen + 1) + pam
    pats = [x + "NNN" for x in guides]
    plants = [(x + "AGG", 0, glen, min(m + 1, 8), 2) for x, m in zip(guides, mm)]
    sg = cofpy.SynthGenome(glen + n, [200_001, 70_007], line_width=61, n_frac=0.01, gap=(30, 900), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    # the reference takes one mismatch limit per input file: one oracle run
    # per limit, rows merged
    want = []
    for m in sorted(set(mm)):
        in_text = f"{fa}\n{filt}\n" + "".join(f"{p} {m}\n" for p, pm in zip(pats, mm) if pm == m)
        oracle.run(in_text, tmp_path / f"ref{m}.tsv", batch_nucl=1 << 22)
        want += cofpy.read_rows(tmp_path / f"ref{m}.tsv")
    want.sort()
    s = cofpy.Search(pats, [filt] * n, mm, batch_nucl=1 << 22)
    s.set_tuning(plain_threads=threads, plain_pat_chunk=chunk)
    s.load_file(fa)
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    s.close()
    assert len(want) > 0
    assert cofpy.read_rows(tmp_path / "gpu.tsv") == want


def test_scan_image_follows_genome_reloads(tmp_path):
    # The scan image (bit planes and unit flags of the resident genome) is
    # built by the first plain search after a genome load: one context loads
    # three genomes in turn (larger, smaller, larger again), and each search
    # must equal a fresh context's and the oracle's rows.
    import random
    rng = random.Random(5)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(24)]
    pats = [g + "NNN" for g in guides]
    filt = "N" * 21 + "RG"
    n = len(pats)
    specs = [(31, [300_001, 120_000]), (32, [70_001]), (33, [410_000, 9_999, 150_001])]
    s = cofpy.Search(pats, [filt] * n, [4] * n, batch_nucl=1 << 22)
    for seed, lens in specs:
        sg = cofpy.SynthGenome(seed, lens, line_width=61, n_frac=0.02, gap=(30, 900),
                               plants=[(g + "AGG", 0, 20, 3, 3) for g in guides])
        fa = tmp_path / f"g{seed}.fa"
        sg.write_fasta(fa)
        s.load_file(fa)
        s.search()
        s.write_tsv(tmp_path / f"reuse{seed}.tsv")
        fresh = cofpy.Search(pats, [filt] * n, [4] * n, batch_nucl=1 << 22)
        fresh.load_file(fa)
        fresh.search()
        fresh.write_tsv(tmp_path / f"fresh{seed}.tsv")
        fresh.close()
        oracle.run(f"{fa}\n{filt}\n" + "".join(f"{p} 4\n" for p in pats), tmp_path / f"ref{seed}.tsv",
                   batch_nucl=1 << 22)
        want = cofpy.read_rows(tmp_path / f"ref{seed}.tsv")
        assert len(want) > 0
        assert cofpy.read_rows(tmp_path / f"reuse{seed}.tsv") == want
        assert cofpy.read_rows(tmp_path / f"fresh{seed}.tsv") == want
    s.close()


@pytest.mark.parametrize("pam,three_prime", [("TTTV", False), ("NGCG", True), ("NBHY", True), ("SMN", True)])
def test_pam_sets_with_c_on_plain_tiles(tmp_path, pam, three_prime):
    # filter codes holding C (C, Y, S, M, B, H, V) on a genome with no IUPAC
    # letters, so nearly every tile takes the bit-plane PAM test and each
    # of its base selects (T, C, A, G) is exercised
    import random
    rng = random.Random(len(pam) * 13 + three_prime)
    guides = _guides(rng, 4, 20)
    if three_prime:
        filt = "N" * 20 + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, 20, 3, 30) for g in guides]
    else:
        filt = pam + "N" * 20
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), 20, 3, 30) for g in guides]
    sg = cofpy.SynthGenome(2000 + len(pam), [300_000, 170_001], n_frac=0.005, gap=(50, 2000), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    got, want, _ = run_both(tmp_path, fa, filt, pats, 4)
    assert len(want) >= 30
    assert got == want
