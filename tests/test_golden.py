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
    runs = make_golden.run_sets(case)
    bulge = any(d or r for _, _, d, r, _ in runs)
    max_m = max(m for _, m, _, _, _ in runs)
    for r in rows[:2000]:
        cols = r.split("\t")
        assert len(cols) == 8
        assert cols[0] in (("X", "DNA", "RNA") if bulge else ("X",))
        assert cols[5] in ("+", "-")
        assert int(cols[6]) <= max_m


def test_mixed_case_spans_the_batched_mix():
    # the mixed case covers every PAM, lengths across 17-30, M 0-6 and bulge guides
    case = next(c for c in make_golden.CASES if "mixed" in c)
    pams = {f.strip("N") for _, f, _, _, _ in case["mixed"]}
    assert pams == {"RG", "GG", "GRRT", "TTT"}
    lens = [len(f) - len(f.strip("N")) + len(f.strip("N")) for _, f, _, _, _ in case["mixed"]]
    glens = [len(p.strip("N")) for p, _, _, _, _ in case["mixed"]]
    assert min(glens) <= 19 and max(glens) >= 28 and len(lens) == len(glens)
    assert {m for _, _, m, _, _ in case["mixed"]} == set(range(7))
    assert sum(1 for _, _, _, d, r in case["mixed"] if d or r) >= 5
    assert len(make_golden.run_sets(case)) > 10
