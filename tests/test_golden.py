"""The committed golden vectors are reproducible: re-running tests/golden/make_golden.py
(seeded synthetic genomes + the oracle) must give byte-identical files."""
import gzip
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402


def test_golden_vectors_reproduce(tmp_path):
    make_golden.main(str(tmp_path))
    for case in make_golden.CASES:
        name = case["name"] + ".tsv.gz"
        want = gzip.decompress(open(os.path.join(GOLDEN, name), "rb").read())
        got = gzip.decompress(open(tmp_path / name, "rb").read())
        assert got == want, case["name"]


@pytest.mark.parametrize("case", make_golden.CASES, ids=[c["name"] for c in make_golden.CASES])
def test_golden_rows_well_formed(case):
    rows = gzip.decompress(open(os.path.join(GOLDEN, case["name"] + ".tsv.gz"), "rb").read()).decode().splitlines()
    assert rows == sorted(rows) and len(rows) > 0
    bulge = case["D"] or case["R"]
    for r in rows[:2000]:
        cols = r.split("\t")
        assert len(cols) == 8
        assert cols[0] in (("X", "DNA", "RNA") if bulge else ("X",))
        assert cols[5] in ("+", "-")
        assert int(cols[6]) <= case["M"]
