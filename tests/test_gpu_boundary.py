"""The C ABI as a drop-in for the reference's search() contract, on the GPU.

* Match records: cof_search + cof_hit_seqs + cof_chrom_name carry every field
  of Match (chrom_chunk.rs:19-29); rows rebuilt from them by the writer's
  rules (main.rs:130-247) equal the oracle's rows, so a Rust caller can keep
  its writer unchanged.
* Streaming: chunks handed over one by one (cof_stream_*, the channel of
  search(), search.rs:1490-1500) give the rows of the whole genome.
* Batch seam: the reference's batch loop driven from outside, one
  cof_search_batch per SearchChunkInfo (the batcher of search.rs:1525-1572
  restated here), gives the oracle's rows at that batch size.
* COF_E_OVERFLOW writes nothing; a plain scan whose hit segments overflow
  grows them, scans again and still gives the oracle's rows.
"""
import ctypes as C
import random

import pytest

import cofpy
import oracle

pytestmark = pytest.mark.gpu

B4 = {c: v for c, v in zip("TCAGRYSWKMBDHVN", [1, 2, 4, 8, 12, 3, 10, 5, 9, 6, 11, 13, 7, 14, 15])}


def _cmp_chars(dna, rna):
    # bit4ops.rs:233-237: the DNA side maps A/C/G/T only
    d = {"A": 4, "C": 2, "G": 8, "T": 1}.get(chr(dna).upper(), 0) if dna else 0
    return (d & B4.get(chr(rna).upper(), 0)) != 0


def writer_rows(s, guides, bulge):
    """The writer's rows (main.rs:130-247) from the Match fields the ABI returns."""
    hits = s.hits()
    seqs = s.hit_seqs(0, len(hits))
    pam_len = len(guides[0]) - len(guides[0].rstrip("Nn"))
    rows = []
    for h, (dna, rna) in zip(hits, seqs):
        fwd = h.pattern < len(guides)
        if bulge:
            tot = h.dna_bulge + h.rna_bulge
            btype = "X" if tot == 0 else ("DNA" if h.dna_bulge > 0 else "RNA")
            size, dna_out = tot, dna.decode()
        else:
            btype, size = "X", 0
            out, amm = [], 0
            for d, r in zip(dna, rna):
                rn = chr(r) in "Nn"
                if not rn and not _cmp_chars(d, r):
                    amm += 1
                if d == 0:
                    out.append("N" if rn else "n")
                elif rn:
                    out.append(chr(d))
                elif chr(d) in "Nn":
                    out.append("n")
                elif not _cmp_chars(d, r):
                    out.append(chr(d | 0x20))
                else:
                    out.append(chr(d))
            assert amm == h.mismatches
            dna_out = "".join(out)
        pos = h.chrom_pos if fwd else h.chrom_pos + max(max(pam_len - 1, 0) - h.dna_bulge, 0)
        rows.append(f"{btype}\t{rna.decode()}\t{dna_out}\t{s.chrom_name(h.chrom)}\t{pos}\t{'+' if fwd else '-'}\t"
                    f"{h.mismatches}\t{size}")
    return sorted(rows)


def _oracle(tmp_path, genome, filt_line, guides, mm, batch_nucl):
    oracle.run(f"{genome}\n{filt_line}\n" + "".join(f"{g} {mm}\n" for g in guides), tmp_path / "ref.tsv",
               batch_nucl=batch_nucl)
    return cofpy.read_rows(tmp_path / "ref.tsv")


def _setup(tmp_path, pam, three, D, R, seed=3):
    rng = random.Random(seed)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(2)]
    if three:
        filt = "N" * 20 + pam
        pats = [g + "N" * len(pam) for g in guides]
        plants = [(g + pam, 0, 20, 4, 25) for g in guides]
    else:
        filt = pam + "N" * 20
        pats = ["N" * len(pam) + g for g in guides]
        plants = [(pam + g, len(pam), 20, 4, 25) for g in guides]
    sg = cofpy.SynthGenome(seed, [65520 * 2 + 901, 40_001, 3_001], n_frac=0.02, gap=(5, 300), iupac_frac=0.002,
                           repeats=(0.05, 3, 300, 0.1), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    return sg, fa, filt, pats


CASES = [("NRG", True, 0, 0, 5), ("NRG", True, 1, 1, 3), ("TTTN", False, 2, 1, 3), ("NNGRRT", True, 0, 0, 6)]


@pytest.mark.parametrize("pam,three,D,R,M", CASES)
def test_match_records_rebuild_oracle_rows(tmp_path, pam, three, D, R, M):
    sg, fa, filt, pats = _setup(tmp_path, pam, three, D, R)
    bulge = D + R > 0
    want = _oracle(tmp_path, fa, f"{filt} {D} {R}" if bulge else filt, pats, M, 1 << 22)
    s = cofpy.Search(pats, [filt] * len(pats), [M] * len(pats), [D] * len(pats), [R] * len(pats),
                     batch_nucl=1 << 22)
    s.load_file(fa)
    s.search()
    got = writer_rows(s, pats, bulge)
    assert s.chrom_name(999) is None
    s.close()
    assert len(want) > 0 and got == want


def _chunks(sg):
    ptr, nb, chunks, n, names = sg.layout()
    data = C.string_at(ptr, nb)
    out = []
    for k in range(n):
        c = chunks[k]
        out.append((names[c.chrom].decode(), c.start, c.end, data[c.byte_offset:c.byte_offset + (c.end - c.start + 1) // 2]))
    return out


@pytest.mark.parametrize("D,R", [(0, 0), (1, 1)])
def test_stream_chunks(tmp_path, D, R):
    sg, fa, filt, pats = _setup(tmp_path, "NRG", True, D, R, seed=7)
    want = _oracle(tmp_path, fa, f"{filt} {D} {R}" if D + R else filt, pats, 4, 1 << 22)
    s = cofpy.Search(pats, [filt] * 2, [4] * 2, [D] * 2, [R] * 2, batch_nucl=1 << 22)
    s.stream(_chunks(sg))
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    assert cofpy.read_rows(tmp_path / "gpu.tsv") == want and len(want) > 0
    # chunks out of order are refused
    L = cofpy.lib()
    assert L.cof_stream_begin(s.h) == cofpy.COF_OK
    name, st, en, data = _chunks(sg)[1]
    assert L.cof_stream_chunk(s.h, name.encode(), st, en, data) == cofpy.COF_E_INVALID
    s.close()


def reference_batches(n_chunks, batch_nucl):
    """(members, has_head) per batch, as the batcher of search.rs:1525-1572 cuts them."""
    cps = batch_nucl // 65536
    out, buf, head = [], [], False
    for i in range(n_chunks):
        buf.append(i)
        if len(buf) == cps:
            out.append((buf, head))
            buf, head = [buf[-1]], True
    if buf and not (head and len(buf) == 1):
        out.append((buf, head))
    return out


@pytest.mark.parametrize("D,R,M", [(0, 0, 5), (1, 1, 3), (2, 0, 3)])
@pytest.mark.parametrize("batch_nucl", [1 << 17, 3 << 16])
def test_batch_loop(tmp_path, D, R, M, batch_nucl):
    sg, fa, filt, pats = _setup(tmp_path, "NRG", True, D, R, seed=11 + D)
    want = _oracle(tmp_path, fa, f"{filt} {D} {R}" if D + R else filt, pats, M, batch_nucl)
    chunks = _chunks(sg)
    s = cofpy.Search(pats, [filt] * 2, [M] * 2, [D] * 2, [R] * 2)
    got = []
    for members, head in reference_batches(len(chunks), batch_nucl):
        names, table, packed = [], [], b""
        for i in members:
            name, st, en, data = chunks[i]
            if st == 0 or not names:
                names.append(name)
            table.append((len(names) - 1, st, en, len(packed)))
            packed += data
        s.search_batch(packed, table, names, head)
        got += writer_rows(s, pats, D + R > 0)
    s.close()
    assert len(want) > 0 and sorted(got) == want


def test_overflow_writes_nothing(tmp_path):
    sg, fa, filt, pats = _setup(tmp_path, "NRG", True, 0, 0)
    s = cofpy.Search(pats, [filt] * 2, [5] * 2)
    s.load_synth(sg)
    n = s.search()
    assert n > 4
    buf = (cofpy.CofHit * 4)()
    C.memset(buf, 0xAB, C.sizeof(buf))
    need = C.c_uint64()
    assert cofpy.lib().cof_search(s.h, buf, 4, C.byref(need)) == cofpy.COF_E_OVERFLOW
    assert need.value == n
    assert bytes(buf) == b"\xab" * C.sizeof(buf)
    # a too-small string stride is refused before anything is written
    dna = C.create_string_buffer(b"\x11" * 64)
    rna = C.create_string_buffer(b"\x11" * 64)
    lens = (C.c_uint32 * 4)()
    assert cofpy.lib().cof_hit_seqs(s.h, 0, 2, dna, rna, 10, lens) == cofpy.COF_E_OVERFLOW
    assert dna.raw[:64] == b"\x11" * 64
    s.close()


def test_plain_hit_segments_regrow(tmp_path):
    # 16 records per segment: every segment overflows, the scan runs again
    # with grown segments, and the rows are still the oracle's
    rng = random.Random(21)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(2)]
    sg = cofpy.SynthGenome(21, [300_000, 120_001], n_frac=0.01, gap=(10, 200))
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    filt = "N" * 20
    want = _oracle(tmp_path, fa, filt, guides, 10, 1 << 22)
    launches = []
    for cap in (0, 64 * 16):
        s = cofpy.Search(guides, [filt] * 2, [10] * 2, batch_nucl=1 << 22)
        s.set_tuning(plain_hit_cap=cap)
        s.load_file(fa)
        s.search()
        launches.append(s.stats().n_plain_launches)
        if cap == 0:
            assert writer_rows(s, guides, False) == want
            # a plain Match holds NUL where the masked genome is N (search.rs:512-531, 1456-1475)
            assert any(0 in d for d, _ in s.hit_seqs())
        s.write_tsv(tmp_path / "gpu.tsv")
        s.close()
        assert cofpy.read_rows(tmp_path / "gpu.tsv") == want
    assert len(want) > 64 * 16 * 4
    assert launches[1] > launches[0]
