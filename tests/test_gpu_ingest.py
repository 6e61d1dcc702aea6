"""Device FASTA and 2bit ingest (cof_load_genome_file_ex) give the oracle's genome.

FASTA files with several line shapes are loaded both ways (device encoder and
host reader) and searched with a dense all-N query whose rows print the
genome text around every hit; both must give the oracle's rows for the same
file (the oracle's reader restates read_fasta.rs:9-71 and is pinned by the
reference's upstream1000 fixtures in test_oracle.py).  Also: the reference
query and fixtures, irregular files (host fallback, or refusal when the
device path is required), a folder of *.fa files (read_fasta_folder.rs:8-23),
and 2bit files (the reference fixture and synthetic files with N-block edge
cases, read_2bit.rs:33-118) decoded on the device.
"""
import os
import random

import pytest

import cofpy
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def _rows(path, flags, guides, filt, mm, tmp_path, tag):
    s = cofpy.Search(guides, [filt] * len(guides), [mm] * len(guides), batch_nucl=1 << 22)
    s.load_file(path, flags)
    s.search()
    st = s.stats()
    out = tmp_path / f"{tag}.tsv"
    s.write_tsv(out)
    s.close()
    return cofpy.read_rows(out), st


def _dense_query(seed):
    rng = random.Random(seed)
    return ["".join(rng.choice("ACGT") for _ in range(12)) for _ in range(3)], "N" * 12, 5


def _oracle_rows(path, guides, filt, mm, tmp_path):
    oracle.run(f"{path}\n{filt}\n" + "".join(f"{g} {mm}\n" for g in guides), tmp_path / "ref.tsv",
               batch_nucl=1 << 22)
    return cofpy.read_rows(tmp_path / "ref.tsv")


def _compare(path, tmp_path, seed=1, expect_device=True):
    guides, filt, mm = _dense_query(seed)
    want = _oracle_rows(path, guides, filt, mm, tmp_path)
    host, st_h = _rows(path, cofpy.COF_LOAD_HOST, guides, filt, mm, tmp_path, "h")
    dev, st_d = _rows(path, 0, guides, filt, mm, tmp_path, "d")
    assert st_h.load_device == 0
    assert st_d.load_device == (1 if expect_device else 0)
    assert len(want) > 1000
    assert dev == want
    assert host == want


@pytest.mark.parametrize("width", [60, 61, 80, 1000])
def test_device_ingest_equals_host(tmp_path, width):
    sg = cofpy.SynthGenome(300 + width, [400_000, 65_536 * 2 + 7, 1_003], line_width=width, n_frac=0.03,
                           gap=(20, 900), iupac_frac=0.001)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    _compare(fa, tmp_path, width)


def test_crlf_lowercase_and_empty_chromosome(tmp_path):
    sg = cofpy.SynthGenome(77, [200_000, 50_001], line_width=70, n_frac=0.02, gap=(10, 500))
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    text = open(fa).read().splitlines()
    rng = random.Random(3)
    out = []
    for ln in text:
        if not ln.startswith(">") and rng.random() < 0.3:
            ln = ln.lower()
        out.append(ln)
    out.insert(0, ">empty_chrom")
    with open(tmp_path / "crlf.fa", "w", newline="") as f:
        f.write("\r\n".join(out) + "\r\n")
    _compare(tmp_path / "crlf.fa", tmp_path, 5)


def test_irregular_lines_fall_back(tmp_path):
    rng = random.Random(9)
    lines = [">chrA"]
    for _ in range(3000):
        lines.append("".join(rng.choice("ACGTN") for _ in range(rng.randint(40, 90))))
    fa = tmp_path / "irr.fa"
    with open(fa, "w") as f:
        f.write("\n".join(lines) + "\n")
    _compare(fa, tmp_path, 7, expect_device=False)
    guides, filt, mm = _dense_query(7)
    s = cofpy.Search(guides, [filt] * 3, [mm] * 3)
    with pytest.raises(cofpy.CofError) as e:
        s.load_file(fa, cofpy.COF_LOAD_DEVICE)
    assert e.value.status == cofpy.COF_E_UNSUPPORTED
    s.close()


def test_folder_of_fasta_files(tmp_path):
    # read_fasta_folder.rs:8-23: every regular *.fa file of the folder, other
    # files and sub-folders skipped; the folder goes through the host reader
    d = tmp_path / "genome"
    d.mkdir()
    for k, (lens, width) in enumerate([([90_001, 3_000], 60), ([70_000], 61), ([65_536 * 2 + 11], 80)]):
        sg = cofpy.SynthGenome(900 + k, lens, line_width=width, n_frac=0.02, gap=(20, 400), iupac_frac=0.001)
        sg.write_fasta(d / f"part{k}.fa")
    cofpy.SynthGenome(5, [50_000]).write_fasta(d / "skipped.fasta")
    (d / "notes.txt").write_text("not a genome\n")
    (d / "sub.fa").mkdir()
    guides, filt, mm = _dense_query(13)
    want = _oracle_rows(d, guides, filt, mm, tmp_path)
    got, st = _rows(d, 0, guides, filt, mm, tmp_path, "f")
    assert st.load_device == 0
    assert len(want) > 1000 and got == want
    with pytest.raises(cofpy.CofError) as e:
        _rows(d, cofpy.COF_LOAD_DEVICE, guides, filt, mm, tmp_path, "x")
    assert e.value.status == cofpy.COF_E_UNSUPPORTED


def test_reference_fixture_both_ways(tmp_path):
    # integration_bulge.rs:159-175 query on upstream1000.fa: 117 rows either way
    fa = os.path.join(GOLDEN, "upstream1000.fa")
    guides = ["CCGTGGTTCAACATTTGCTTAGCA", "GATGTTGGTAAGTGGGATATGGCA"]
    h, st_h = _rows(fa, cofpy.COF_LOAD_HOST, guides, "N" * 24, 11, tmp_path, "h")
    d, st_d = _rows(fa, cofpy.COF_LOAD_DEVICE, guides, "N" * 24, 11, tmp_path, "d")
    assert len(h) == 117 and d == h and st_d.load_device == 1


def test_reference_2bit_fixture_both_ways(tmp_path):
    # the same query on upstream1000.2bit: device 2bit decode == host reader, 117 rows
    tb = os.path.join(GOLDEN, "upstream1000.2bit")
    guides = ["CCGTGGTTCAACATTTGCTTAGCA", "GATGTTGGTAAGTGGGATATGGCA"]
    h, st_h = _rows(tb, cofpy.COF_LOAD_HOST, guides, "N" * 24, 11, tmp_path, "h")
    d, st_d = _rows(tb, cofpy.COF_LOAD_DEVICE, guides, "N" * 24, 11, tmp_path, "d")
    assert len(h) == 117 and d == h and st_d.load_device == 1 and st_h.load_device == 0


def test_2bit_device_decode_edge_cases(tmp_path):
    # pieces of 64 Knt with N blocks across, on and at the ends of piece
    # boundaries, unsorted and overlapping blocks (the reader's one-block-back
    # walk), mask blocks, an empty and a 3-nt chromosome
    import twobit
    from test_twobit_cpu import P, _records
    rng = random.Random(17)
    recs = _records(21) + [("chrF", "".join(rng.choice("ACGT") for _ in range(2 * P + 77)),
                            [(P + 100, 50), (P - 10, 200), (5, 10), (2 * P + 70, 7)], [(3, 4)])]
    path = tmp_path / "e.2bit"
    twobit.write_2bit(path, recs)
    _compare(path, tmp_path, 11)
    guides, filt, mm = _dense_query(11)
    got, _ = _rows(path, 0, guides, filt, mm, tmp_path, "o")
    oracle.run(f"{path}\n{filt}\n" + "".join(f"{g} {mm}\n" for g in guides), tmp_path / "ref.tsv", batch_nucl=1 << 22)
    assert got == cofpy.read_rows(tmp_path / "ref.tsv")


def test_2bit_truncated_file_is_an_error(tmp_path):
    import twobit
    path = tmp_path / "t.2bit"
    twobit.write_2bit(path, [("c1", "ACGT" * 1000, [], [])])
    data = open(path, "rb").read()
    open(path, "wb").write(data[:-10])
    guides, filt, mm = _dense_query(3)
    for flags in (0, cofpy.COF_LOAD_HOST):
        s = cofpy.Search(guides, [filt] * 3, [mm] * 3)
        with pytest.raises(cofpy.CofError) as e:
            s.load_file(path, flags)
        assert e.value.status == cofpy.COF_E_IO
        s.close()
