"""cof-cli end to end: reference-format input file -> TSV, compared with the oracle's TSV.

The CLI is the drop-in for cas-offinder-cli (main.rs:55-305, input format
cli_utils.rs:37-148): the same input text drives both, and the sorted data
rows and the header must be identical.  Covers a plain query, a bulge query,
stdin input and a malformed input (status and message, no crash).
"""
import os
import random
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


def test_device_letter_picks_reference_batches(tmp_path):
    # C reproduces the CPU path's 4 Mnt batches, G the GPU path's 1 Gnt ones
    # (search.rs:21-50).  11-nt lines make every full chunk 65527 nt (odd), so
    # a 0 nibble ends each chunk; at a 4 Mnt batch seam the start whose window
    # is the chunk's last 23 nt plus that 0 nibble is reported by neither
    # batch (the first drops it at the batch end, the second starts at
    # chunk 1 - (L - 1), search.rs:540-546, 1449-1454), while one 1 Gnt batch
    # reports it.  A guide copy planted there shows the difference.
    rng = random.Random(5)
    guide = "".join(rng.choice("ACGT") for _ in range(24))
    chunk = 65527
    seq = rng.choices("ACGT", k=64 * chunk + 5000)
    at = 63 * chunk + chunk - 23
    seq[at:at + 23] = list(guide[:23])
    seq = "".join(seq)
    fa = tmp_path / "g.fa"
    with open(fa, "w") as f:
        f.write(">chr1\n")
        for i in range(0, len(seq), 11):
            f.write(seq[i:i + 11] + "\n")
    text = f"{fa}\n{'N' * 24}\n{guide} 1\n"
    (tmp_path / "in.txt").write_text(text)
    rows = {}
    for dev, batch in (("C", 1 << 22), ("G", 1 << 30)):
        oracle.run(text, tmp_path / f"ref{dev}.tsv", batch_nucl=batch)
        want = cofpy.read_rows(tmp_path / f"ref{dev}.tsv")
        env = {k: v for k, v in os.environ.items() if k != "COF_BATCH_NUCL"}
        r = subprocess.run([CLI, str(tmp_path / "in.txt"), dev, str(tmp_path / f"out{dev}.tsv")], capture_output=True,
                           text=True, env=env, timeout=120)
        assert r.returncode == 0, r.stderr
        got = cofpy.read_rows(tmp_path / f"out{dev}.tsv")
        assert got == want
        rows[dev] = set(got)
    assert rows["C"] < rows["G"]
    assert any(r.split("\t")[4] == str(at) for r in rows["G"] - rows["C"])
