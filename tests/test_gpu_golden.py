"""libcof on the GPU reproduces the committed golden vectors (oracle rows for
seeded synthetic genomes: plain and bulge queries, 3'/5' PAMs, IUPAC letters,
odd line widths, small virtual batches, 32- and 64-bit bulge windows, and one
query mixing lengths 17-30, four PAMs, M 0-6 and bulge guides, whose rows are
the union of one oracle run per reference-expressible set)."""
import gzip
import os
import sys

import pytest

import cofpy

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", make_golden.CASES, ids=[c["name"] for c in make_golden.CASES])
def test_gpu_matches_golden(tmp_path, case):
    want = gzip.decompress(open(os.path.join(GOLDEN, case["name"] + ".tsv.gz"), "rb").read()).decode().splitlines()
    if "mixed" in case:
        # one context for the whole mix (per-guide filter, length, M, D, R)
        pats, filts, mms, ds, rs = (list(x) for x in zip(*case["mixed"]))
    else:
        n = len(case["patterns"])
        pats, filts, mms, ds, rs = (case["patterns"], [case["filter"]] * n, [case["M"]] * n, [case["D"]] * n,
                                    [case["R"]] * n)
    s = cofpy.Search(pats, filts, mms, ds, rs, batch_nucl=case["batch_nucl"])
    s.load_synth(make_golden.make_genome(case["genome"]))
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    s.write_tsv(tmp_path / "host.tsv", host=True)
    s.close()
    assert cofpy.read_rows(tmp_path / "gpu.tsv") == want
    assert open(tmp_path / "gpu.tsv", "rb").read() == open(tmp_path / "host.tsv", "rb").read()
