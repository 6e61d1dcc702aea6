"""cof-cli end to end: reference-format input file -> TSV, compared with the oracle's TSV.

The CLI is the drop-in for cas-offinder-cli (main.rs:55-305, input format
cli_utils.rs:37-148): the same input text drives both, and the sorted data
rows and the header must be identical.  Covers a plain query, a bulge query,
stdin input and a malformed input (status and message, no crash).
"""
import os
import subprocess

import pytest

import cofpy
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "cas-offinder-rust-old_amd", "_build", "cof-cli")
pytestmark = pytest.mark.gpu


def _genome(tmp_path):
    guide = "GGCCGACCTGTCGCTGACGC"
    sg = cofpy.SynthGenome(41, [180_000, 70_001], n_frac=0.02, gap=(30, 800), iupac_frac=0.0005,
                           plants=[(guide + "AGG", 0, 20, 4, 30)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    return fa, guide


def _run_cli(args, stdin=None):
    env = dict(os.environ, COF_BATCH_NUCL=str(1 << 22))
    return subprocess.run([CLI] + args, input=stdin, capture_output=True, text=True, env=env, timeout=120)


@pytest.mark.parametrize("filt_line,mm", [("NNNNNNNNNNNNNNNNNNNNNRG", 5), ("NNNNNNNNNNNNNNNNNNNNNRG 1 1", 3)])
def test_cli_matches_oracle(tmp_path, filt_line, mm):
    fa, guide = _genome(tmp_path)
    text = f"{fa}\n{filt_line}\n{guide}NNN {mm}\nCGCCAGCGTCAGCGACAGGTNNN {mm}\n"
    (tmp_path / "in.txt").write_text(text)
    oracle.run(text, tmp_path / "ref.tsv", batch_nucl=1 << 22)
    r = _run_cli([str(tmp_path / "in.txt"), "G", str(tmp_path / "out.tsv")])
    assert r.returncode == 0, r.stderr
    got = open(tmp_path / "out.tsv").read().splitlines()
    want = open(tmp_path / "ref.tsv").read().splitlines()
    assert got[0] == want[0]                      # header
    assert sorted(got[1:]) == sorted(want[1:])
    assert len(want) > 1
    # run summary beside the output (log_writer.rs:27-53)
    log = open(str(tmp_path / "out.tsv") + ".log").read().splitlines()
    bulge = len(filt_line.split()) == 3
    assert log[0] == "=== Cas-OFFinder Rust ==="
    assert log[1].startswith("Run date (unix timestamp): ") and int(log[1].split(": ")[1]) > 1_600_000_000
    assert log[3] == "Algorithm: " + ("Myers bit-parallel" if bulge else "popcount (legacy)")
    assert log[4] == "Device: GPU"
    assert log[6:14] == ["Input:", f"  Genome: {fa}", f"  Genome size: {os.path.getsize(fa)} bytes",
                         "  Patterns: 2", f"  Search filter: {filt_line.split()[0]}", f"  Max mismatches: {mm}",
                         f"  Max DNA bulges: {1 if bulge else 0}", f"  Max RNA bulges: {1 if bulge else 0}"]
    assert log[15:17] == ["Results:", f"  Matches found: {len(want) - 1}"]
    assert log[17].startswith("  Total elapsed: ") and log[17].endswith("s")
    # stdin input, stdout output
    os.remove(str(tmp_path / "out.tsv") + ".log")
    r2 = _run_cli(["-", "G", "-"], stdin=text)
    assert r2.returncode == 0, r2.stderr
    assert sorted(r2.stdout.splitlines()[1:]) == sorted(want[1:])
    assert not any(p.name.endswith(".log") for p in tmp_path.iterdir())     # no summary for stdout output


def test_cli_rejects_bad_input(tmp_path):
    fa, guide = _genome(tmp_path)
    # mismatched pattern length (cli_utils.rs:104-108)
    r = _run_cli(["-", "G", str(tmp_path / "o.tsv")], stdin=f"{fa}\nNNNNNNNNNNNNNNNNNNNNNRG\n{guide}NN 3\n")
    assert r.returncode != 0 and "error" in r.stderr.lower()
    r = _run_cli(["-", "X", str(tmp_path / "o.tsv")], stdin=f"{fa}\nNNNNNNNNNNNNNNNNNNNNNRG\n{guide}NNN 3\n")
    assert r.returncode != 0
