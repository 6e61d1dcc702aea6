"""Pins the CPU oracle's bulge path against an independent brute force.

The oracle (oracle/cof_ref.cpp) restates the reference's bulge search: quick
PAM check (search.rs:703-755), Myers filter (:841-877), traceback_all
(traceback.rs:181-381), the full PAM check of classify_myers_candidate
(search.rs:597-688), convert_matches_myers' drop rule (:1316-1374), the
per-batch scan with head chunks (:1525-1572, :921-950) and the writer's bulge
branch (main.rs:145-157, 224-242).  tests/bulge_brute.py derives the same rows
with none of that machinery: every op string forward from every start, caps
and per-op rules only, one batch-free text, and a hit kept iff its alignment
lies in one chromosome.  Agreement on every row, at several batch sizes, pins
the quick check (it may drop nothing), the full PAM check, the drop rule and
the seams.  A guide with an N inside it (where the quick check is a gate of
its own) is checked with the gate restated, and the case shows that the gate
matters there.
"""
import random

import pytest

import bulge_brute
import cofpy
import oracle


def _genome(tmp_path, seed, glen, pam, three, M, line_width=60, n_guides=2, lens=None):
    rng = random.Random(seed)
    guides = ["".join(rng.choice("ACGT") for _ in range(glen)) for _ in range(n_guides)]
    if three:
        filt = "N" * glen + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, glen, M, 6) for g in guides]
    else:
        filt = pam + "N" * glen
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), glen, M, 6) for g in guides]
    # two chromosomes with several chunks (seams inside a chromosome) and an
    # odd-length one; chunk tails are longer than a text window
    sg = cofpy.SynthGenome(seed, lens or [65520 * 2 + 701, 3001, 70_001], line_width=line_width, n_frac=0.02,
                           gap=(5, 200), iupac_frac=0.002, repeats=(0.05, 3, 300, 0.15), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    return fa, filt, pats


def _oracle(tmp_path, fa, filt, pats, M, D, R, batch_nucl):
    in_text = f"{fa}\n{filt} {D} {R}\n" + "".join(f"{p} {M}\n" for p in pats)
    oracle.run(in_text, tmp_path / "o.tsv", batch_nucl=batch_nucl)
    return cofpy.read_rows(tmp_path / "o.tsv")


CASES = [
    # pam, 3', guide length, M, D, R, FASTA line width
    ("NRG", True, 20, 3, 2, 2, 60),
    ("NGG", True, 20, 2, 1, 0, 60),
    ("NGG", True, 20, 2, 0, 1, 60),
    ("NNGRRT", True, 21, 3, 1, 2, 60),
    ("TTTN", False, 20, 3, 2, 1, 60),
    ("TTTN", False, 18, 2, 0, 2, 60),
    ("NRG", True, 17, 4, 1, 1, 61),      # odd chunks: a 0 nibble after each
    ("NGG", True, 28, 3, 2, 2, 60),      # W = 33
]


@pytest.mark.parametrize("pam,three,glen,M,D,R,lw", CASES)
def test_bulge_oracle_vs_brute_force(tmp_path, pam, three, glen, M, D, R, lw):
    fa, filt, pats = _genome(tmp_path, glen * 7 + M * 3 + D * 5 + R + lw, glen, pam, three, M, lw)
    want = bulge_brute.rows(fa, filt, pats, M, D, R)
    assert len(want) > 0
    assert any(r.startswith(("DNA", "RNA")) for r in want)
    for batch in (1 << 17, 1 << 30):       # 2-chunk batches (a seam at every chunk) and one batch
        assert _oracle(tmp_path, fa, filt, pats, M, D, R, batch) == want, batch


def test_bulge_batch_size_invariance(tmp_path):
    fa, filt, pats = _genome(tmp_path, 5, 20, "NRG", True, 3)
    res = [_oracle(tmp_path, fa, filt, pats, 3, 1, 2, b) for b in (1 << 17, 3 << 16, 1 << 22, 1 << 30)]
    assert res[0] == res[1] == res[2] == res[3] and len(res[0]) > 0


def test_quick_check_gate_internal_n(tmp_path):
    # N inside the guide: the quick check's PAM offset is that N, and its PAM
    # cells are the filter at the pattern's N cells from there on, read at
    # consecutive positions (search.rs:761-777, 703-755), which the full PAM
    # check does not imply
    rng = random.Random(8)
    g = "".join(rng.choice("ACGT") for _ in range(20))
    gn = g[:6] + "N" + g[7:]
    sg = cofpy.SynthGenome(8, [65520 + 901, 40_001], n_frac=0.01, gap=(5, 100),
                           plants=[(g + "AGG", 0, 20, 2, 40), (g[:8] + "AG" + g[10:] + "CGG", 0, 20, 1, 40)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    filt = "N" * 20 + "NRG"
    pats = [gn + "NNN"]
    got = _oracle(tmp_path, fa, filt, pats, 2, 1, 1, 1 << 17)
    assert got == bulge_brute.rows(fa, filt, pats, 2, 1, 1, quick_gate=True)
    assert len(got) > 0
    assert got != bulge_brute.rows(fa, filt, pats, 2, 1, 1, quick_gate=False)


def test_brute_force_refuses_short_chunks(tmp_path):
    sg = cofpy.SynthGenome(1, [65520 + 20, 5000])   # a 20-nt last line starts a chunk of its own
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    with pytest.raises(ValueError):
        bulge_brute.rows(fa, "N" * 20 + "NGG", ["A" * 20 + "NNN"], 1, 1, 1)
