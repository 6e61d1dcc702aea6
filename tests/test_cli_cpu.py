"""cof-cli without a GPU: usage text and argument checks that return before any device work.

Usage text follows cas-offinder-cli/src/main.rs:17-46 (printed to stderr, exit 0
when no argument is given); a bad device letter is rejected as in
cli_utils.rs:37-60.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "cas-offinder-rust-old_amd", "_build", "cof-cli")


def _run(args):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)


def test_usage_without_arguments():
    r = _run([])
    assert r.returncode == 0
    assert "Usage: cof-cli [options] {input_filename|-} {C|G|A}[device_id(s)] {output_filename|-}" in r.stderr
    assert "(C: using CPUs, G: using GPUs, A: using accelerators)" in r.stderr
    assert "NNNNNNNNNNNNNNNNNNNNNRG\nGGCCGACCTGTCGCTGACGCNNN 5\n" in r.stderr
    assert "Available device list:" in r.stderr


def test_too_few_arguments_and_bad_device(tmp_path):
    r = _run(["in.txt"])
    assert r.returncode == 1 and "Usage:" in r.stderr
    (tmp_path / "in.txt").write_text("g.fa\nNNNRG\nACNNN 1\n")
    for dev in ("X", "G0", "CG", ""):          # one letter of C, G, A (cli_utils.rs:4-16)
        r = _run([str(tmp_path / "in.txt"), dev, str(tmp_path / "o.tsv")])
        assert r.returncode == 1 and "{C|G|A}" in r.stderr
        assert not (tmp_path / "o.tsv").exists()
