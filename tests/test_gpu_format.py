"""Device-side TSV formatting (cof_write_tsv) against the host writer and the oracle.

The device formatter reproduces the reference writer's rows (cas-offinder-cli
main.rs:130-247): plain rows with the N/lowercase rules of :198-217 and the
reverse-strand position shift of :224-242, bulge rows with the aligned strings
and type/size columns of :145-157.  Every case checks that the device bytes
equal the host writer's bytes exactly and that the sorted rows equal the
oracle's.  Cases: a plain query whose rows span several device batches, long
chromosome names with spaces, a bulge query with both bulge types, and a
search with no hits (header only).
"""
import os
import random

import pytest

import cofpy
import oracle

pytestmark = pytest.mark.gpu


def _check(tmp_path, s, in_text, min_rows=1):
    s.write_tsv(tmp_path / "dev.tsv")
    st = s.stats()
    assert st.write_device == 1
    s.write_tsv(tmp_path / "host.tsv", host=True)
    assert s.stats().write_device == 0
    dev = open(tmp_path / "dev.tsv", "rb").read()
    host = open(tmp_path / "host.tsv", "rb").read()
    assert dev == host
    oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=1 << 22)
    want = cofpy.read_rows(tmp_path / "ref.tsv")
    assert cofpy.read_rows(tmp_path / "dev.tsv") == want
    assert len(want) >= min_rows
    return len(want)


def test_plain_rows_over_several_batches(tmp_path):
    # 12 guides of 8 nt + NRG at M4 on 2 Mbp: ~5e5 rows, more than one 2^18-row device batch
    rng = random.Random(11)
    guides = ["".join(rng.choice("ACGT") for _ in range(8)) for _ in range(12)]
    sg = cofpy.SynthGenome(77, [1_200_000, 800_001], line_width=61, n_frac=0.01, gap=(20, 500),
                           iupac_frac=0.001)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    filt = "NNNNNNNNNRG"
    pats = [g + "NNN" for g in guides]
    n = len(pats)
    s = cofpy.Search(pats, [filt] * n, [4] * n, [0] * n, [0] * n, batch_nucl=1 << 22)
    s.load_file(fa)
    s.search()
    rows = _check(tmp_path, s, f"{fa}\n{filt}\n" + "".join(f"{p} 4\n" for p in pats), min_rows=(1 << 18) + 1)
    assert s.stats().n_hits == rows
    s.close()


def _named_fasta(path, seed, names, lens):
    rng = random.Random(seed)
    with open(path, "w") as f:
        for nm, L in zip(names, lens):
            f.write(f">{nm}\n")
            seq = "".join(rng.choice("ACGTACGTACGTNacgtRY") for _ in range(L))
            for k in range(0, L, 70):
                f.write(seq[k:k + 70] + "\n")


def test_long_names_and_bulges(tmp_path):
    names = ["chr1 " + "x" * 300 + " assembled molecule", "scaffold_2 | unplaced | v1.1", "c"]
    fa = tmp_path / "named.fa"
    _named_fasta(fa, 5, names, [150_000, 90_001, 40_000])
    rng = random.Random(3)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(4)]
    filt = "NNNNNNNNNNNNNNNNNNNNNRG"
    pats = [g + "NNN" for g in guides]
    n = len(pats)
    for D, R, M in ((0, 0, 9), (1, 1, 5), (2, 0, 5)):
        s = cofpy.Search(pats, [filt] * n, [M] * n, [D] * n, [R] * n, batch_nucl=1 << 22)
        s.load_file(fa)
        s.search()
        sub = tmp_path / f"d{D}r{R}"
        sub.mkdir()
        line2 = filt if D + R == 0 else f"{filt} {D} {R}"
        _check(sub, s, f"{fa}\n{line2}\n" + "".join(f"{p} {M}\n" for p in pats))
        rows = open(sub / "dev.tsv").read().splitlines()[1:]
        assert {r.split("\t")[3] for r in rows} <= set(names)
        if D + R:
            assert any(r.startswith("DNA\t") for r in rows)
        if R:
            assert any(r.startswith("RNA\t") for r in rows)
        s.close()


def test_header_only_without_hits(tmp_path):
    sg = cofpy.SynthGenome(9, [50_000])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    s = cofpy.Search(["ACGTACGTACGTACGTACGTNNN"], ["NNNNNNNNNNNNNNNNNNNNNRG"], [0], [0], [0])
    s.load_file(fa)
    s.search()
    s.write_tsv(tmp_path / "dev.tsv")
    s.write_tsv(tmp_path / "host.tsv", host=True)
    s.close()
    assert open(tmp_path / "dev.tsv", "rb").read() == open(tmp_path / "host.tsv", "rb").read()
    assert open(tmp_path / "dev.tsv").read().count("\n") == 1
