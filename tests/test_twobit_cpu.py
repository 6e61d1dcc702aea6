"""2bit edge cases on the CPU: the oracle's 2bit reader (read_2bit.rs:33-118)
on synthetic files against what the format fixes.

With non-overlapping N blocks the decoded text is the sequence with the block
bases set to N, whatever the piece boundaries (pieces of 64 Knt; blocks that
start inside one piece and end in a later one, start exactly on a piece
boundary, or reach the chromosome end), mask blocks change nothing, and the
chunk table lists one row per 64 Knt piece.  The reader itself is pinned by the
reference's upstream1000.2bit fixture in test_oracle.py; these files are the
inputs of the device 2bit decoder tests.
"""
import random

import oracle
import twobit

P = 1 << 16


def _records(seed):
    rng = random.Random(seed)

    def seq(n):
        return "".join(rng.choice("ACGT") for _ in range(n))

    return [
        ("chrA", seq(2 * P + 5), [(0, 7), (P - 3, 10), (2 * P, 5)], [(100, 50)]),
        ("chrB", seq(3), [], []),
        ("chrEmpty", "", [], []),
        ("chrC", seq(P), [(P - 1, 1)], []),
        ("chrD", seq(200_003), [(P, 3 * P // 2), (199_000, 1_003)], [(0, 200_003)]),
        ("chrE", seq(P + 1), [(10, 5), (20, 5), (P, 1)], []),
    ]


def test_oracle_2bit_edge_cases(tmp_path):
    recs = _records(5)
    path = tmp_path / "e.2bit"
    twobit.write_2bit(path, recs)
    seq, table = oracle.decode_genome(path)
    want = b"".join(twobit.expected_text(r[1], r[2]) for r in recs)
    assert seq == want
    rows = []
    for name, s, _, _ in recs:
        for a in range(0, len(s), P):
            rows.append(f"{name}\t{a}\t{min(a + P, len(s))}")
    assert table.splitlines() == rows


def test_pack_dna_known_bytes():
    # T C A G -> 00 01 10 11, first base in the top bits; a partial byte is zero-padded
    assert twobit.pack_dna("TCAG") == bytes([0b00011011])
    assert twobit.pack_dna("GGGGA") == bytes([0xFF, 0b10000000])
