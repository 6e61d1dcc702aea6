"""Batched query (the config-4 mix) on the GPU vs the union of reference runs.

One libcof context holds guides that no single reference run can express
(cli_utils.rs:97-108 allows one filter, one length and one M per run):
lengths 17-30, PAMs NRG/NGG/NNGRRT/TTTN (3' and 5'), M 0-6 and a fifth of the
guides with DNA/RNA bulges.  Its sorted rows must equal the sorted union of
one oracle run per (filter, M, D, R) set, on a repeat-rich genome with
2-chunk virtual batches (a seam at every chunk), in record mode and, for the
bulge part, in count-only mode (count and checksum).
"""
import os
import sys

import pytest

import cofpy
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu


def test_batched_mix_equals_union_of_reference_runs(tmp_path):
    case = make_golden._mixed_case("gpu_mix", 611, [600_000, 400_001, 65520 * 3 + 777, 50_000], 96,
                                   max_bulge_edits=7)
    fa = tmp_path / "g.fa"
    make_golden.make_genome(case["genome"]).write_fasta(fa)
    want, n_bulge_rows = [], 0
    for run in make_golden.run_sets(case):
        oracle.run(make_golden.input_text(run, fa), tmp_path / "ref.tsv", batch_nucl=1 << 17)
        rows = cofpy.read_rows(tmp_path / "ref.tsv")
        want += rows
        if run[2] or run[3]:
            n_bulge_rows += len(rows)
    want.sort()
    pats, filts, mms, ds, rs = (list(x) for x in zip(*case["mixed"]))
    assert sum(1 for d, r in zip(ds, rs) if d or r) >= 8 and n_bulge_rows > 1000
    s = cofpy.Search(pats, filts, mms, ds, rs, batch_nucl=1 << 17)
    s.load_file(fa)
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    rec = s.stats()
    assert cofpy.read_rows(tmp_path / "gpu.tsv") == want
    # count-only: the plain part is returned, the bulge part counted and checksummed
    s.set_flags(cofpy.COF_FLAG_COUNT_ONLY)
    n_plain = s.search()
    cnt = s.stats()
    s.close()
    assert n_plain == len(want) - n_bulge_rows
    assert (cnt.bulge_hits, cnt.bulge_checksum) == (rec.bulge_hits, rec.bulge_checksum) == (n_bulge_rows, rec.bulge_checksum)
