"""Pins the CPU oracle (oracle/cof_ref.cpp) before it is trusted as the parity reference.

* reader known-answer tests of the reference (file_test.rs:34-73): FASTA, 2bit
  and folder reads of upstream1000 decode to expected.txt;
* the reference's only numeric hit fixture: 117 rows for the
  integration_bulge.rs:159-170 query;
* plain path vs an independent brute force written from main.rs's writer
  semantics in the forward orientation (reverse hits via revcomp of the
  window, the way the writer sees them);
* Myers score vs the textbook O(mn) semi-global DP under the peq match rule,
  plus the plan-doc vectors (one of which is stale, see below);
* traceback_all vs a brute-force enumeration of every op string.
"""
import itertools
import os
import random

import pytest

import cofpy
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

B4 = {c: v for c, v in zip("TCAGRYSWKMBDHVN", [1, 2, 4, 8, 12, 3, 10, 5, 9, 6, 11, 13, 7, 14, 15])}
DEC = {v: c for c, v in B4.items()}


def b4(c):
    return B4.get(c.upper(), 0)


def comp(c):
    v = b4(c)
    if v == 0:
        return c
    r = DEC[((v << 2) | (v >> 2)) & 0xF]
    return r.lower() if c.islower() else r


def revcomp(s):
    return "".join(comp(c) for c in reversed(s))


# --------------------------------------------------------------------------- readers
@pytest.mark.parametrize("name", ["upstream1000.fa", "upstream1000.2bit", ""])
def test_reader_known_answer(name):
    expected = open(os.path.join(GOLDEN, "expected.txt"), "rb").read()
    seq, table = oracle.decode_genome(os.path.join(GOLDEN, name) if name else GOLDEN)
    assert seq == expected
    assert table.count("\n") == 13   # 13 records, each < 64 Knt


def test_reference_117_rows(tmp_path):
    for g in ("upstream1000.fa", "upstream1000.2bit"):
        n, _ = oracle.run(f"{os.path.join(GOLDEN, g)}\n{'N' * 24}\nCCGTGGTTCAACATTTGCTTAGCA 11\n"
                          "GATGTTGGTAAGTGGGATATGGCA 11\n", tmp_path / "o.tsv")
        assert n == 117


def test_n_blocks_closed_form(tmp_path):
    # FASTA N (and lowercase) -> 0xF, decoded back as 'N' / uppercase
    fa = tmp_path / "n.fa"
    fa.write_text(">c1 descr\nACGTNNNNacgtnRYk\nTT\n>c2\nN\n")
    seq, table = oracle.decode_genome(fa)
    assert seq == b"ACGTNNNNACGTNRYKTTN"
    assert table.splitlines() == ["c1 descr\t0\t18", "c2\t0\t1"]


# --------------------------------------------------------------------------- plain brute force
def read_fasta_simple(path):
    chroms, name, parts = [], None, []
    for line in open(path):
        line = line.rstrip("\n")
        if line.startswith(">"):
            if name is not None:
                chroms.append((name, "".join(parts)))
            name, parts = line[1:], []
        else:
            parts.append("".join(DEC.get(b4(c), "\0") if b4(c) else "\0" for c in line))
    chroms.append((name, "".join(parts)))
    return chroms


def cmp_chars(dna, rna):
    return dna.upper() in "ACGT" and dna != "" and dna != "\0" and (b4(dna) & b4(rna)) != 0


def brute_plain(fa, filt, guides, M):
    """Rows of the plain path from the writer's forward-orientation view (main.rs:158-242)."""
    L = len(filt)
    pam_len = len(guides[0]) - len(guides[0].rstrip("Nn"))
    rows = []
    for name, seq in read_fasta_simple(fa):
        for s in range(0, len(seq) - L + 1):
            win = seq[s:s + L]
            for fwd in (True, False):
                dna = win if fwd else revcomp(win)
                masked = "".join("\0" if c == "N" else c for c in dna)
                for g in guides:
                    rna = "".join(DEC[b4(c)] for c in g)
                    # kernel predicate on the N-masked genome: L - sum popc
                    pop = sum(bin(b4(r) & b4(d)).count("1") for r, d in zip(rna, masked))
                    if L - pop > M:
                        continue
                    if any(f.upper() != "N" and not cmp_chars(d, f) for f, d in zip(filt, masked)):
                        continue
                    amm = sum(1 for r, d in zip(rna, masked) if r != "N" and not cmp_chars(d, r))
                    if amm > M:
                        continue
                    out = []
                    for d, r in zip(masked, rna):
                        if d == "\0":
                            out.append("N" if r == "N" else "n")
                        elif r == "N":
                            out.append(d)
                        elif not cmp_chars(d, r):
                            out.append(d.lower())
                        else:
                            out.append(d)
                    pos = s if fwd else s + max(pam_len - 1, 0)
                    rows.append(f"X\t{rna}\t{''.join(out)}\t{name}\t{pos}\t{'+' if fwd else '-'}\t{amm}\t0")
    return sorted(rows)


@pytest.mark.parametrize("seed", [1, 2])
def test_plain_vs_brute_force(tmp_path, seed):
    rng = random.Random(seed)
    guide = "".join(rng.choice("ACGT") for _ in range(12))
    sg = cofpy.SynthGenome(seed, [3000, 1801], n_frac=0.03, gap=(3, 40), iupac_frac=0.01,
                           plants=[(guide + "AGG", 0, 12, 3, 12), (guide + "CGG", 0, 12, 2, 6)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    for filt, pats, M in (("NNNNNNNNNNNNNRG", [guide + "NNN"], 4),
                          ("NNNNNNNNNNNNNNN", [guide + "NNN", guide[::-1] + "NNN"], 5),
                          ("TTNNNNNNNNNNNNN", ["NN" + guide + "N"], 4)):
        in_text = f"{fa}\n{filt}\n" + "".join(f"{p} {M}\n" for p in pats)
        oracle.run(in_text, tmp_path / "o.tsv")
        got = cofpy.read_rows(tmp_path / "o.tsv")
        want = brute_plain(fa, filt, pats, M)
        assert got == want
        assert len(got) > 0


def test_plain_batch_size_invariance(tmp_path):
    # even chunks, no tiny chunk at a seam: reference CPU (4 Mnt) and GPU
    # (1 Gnt) batch sizes give the same rows as 2-chunk batches
    g = "GGCCGACCTGTCGCTGACGC"
    sg = cofpy.SynthGenome(3, [65520 * 2 + 600, 90_000, 70_000], n_frac=0.01, gap=(10, 100),
                           plants=[(g + "AGG", 0, 20, 4, 40)])
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    in_text = f"{fa}\nNNNNNNNNNNNNNNNNNNNNNRG\n{g}NNN 5\n"
    res = []
    for b in (1 << 17, 1 << 22, 1 << 30):
        oracle.run(in_text, tmp_path / "o.tsv", batch_nucl=b)
        res.append(cofpy.read_rows(tmp_path / "o.tsv"))
    assert res[0] == res[1] == res[2] and len(res[0]) > 0


# --------------------------------------------------------------------------- Myers
def myers_dp(pattern, text):
    """Textbook semi-global edit distance D[m][n] (free text prefix) under the peq rule."""
    def match(p, t):
        tv = b4(t)
        if tv == 0:
            return False
        if p.upper() not in "ACGT":          # wildcard: matches any non-padding text
            return True
        return tv != 15 and (tv & b4(p)) != 0
    m, n = len(pattern), len(text)
    prev = [0] * (n + 1)
    for i in range(1, m + 1):
        cur = [i] + [0] * n
        for j in range(1, n + 1):
            cur[j] = min(prev[j - 1] + (0 if match(pattern[i - 1], text[j - 1]) else 1), prev[j] + 1, cur[j - 1] + 1)
        prev = cur
    return prev[n]


def test_myers_plan_vectors():
    # docs/superpowers/plans/2026-04-16-myers-bulge-support.md:170-380.  All hold
    # except ACGT/ACGN, which is 1 under the current genome-N rule (N matches
    # only pattern wildcards, myers.rs:7-10); TTTTTTTT is pinned by the DP.
    cases = [("ACGT", "ACGT", 0), ("ACGT", "TTACGT", 0), ("ACGT", "ACGN", 1), ("ACGT", "AAGT", 1),
             ("ACGT", "ACTT", 1), ("ACGT", "AAGA", 2), ("ACGT", "ATTC", 3),
             ("ACGTACGTACGTACGTACGT", "ACGTACGTACGTACGTACGT", 0), ("ACGTACGTACGTACGTACGT", "ACGTAAGTACGTACGTACGT", 1),
             ("ACGTACGTACGTACGTACGT", "ACGTACGTACGTACGTACGA", 1), ("ACGTACGT", "ACGTTACGT", 1),
             ("ACGTACGT", "TACGTACGT", 0), ("ACGTACGT", "ACGNTAACGT", 2), ("ACGTACGT", "ACGACGT", 1),
             ("ACGTACGT", "ACGACG", 2), ("ACGTACGT", "ACGTXAAGT", 2)]
    for p, t, want in cases:
        assert oracle.myers_score(p, t) == want, (p, t)
        assert myers_dp(p, t) == want
    assert oracle.myers_score("ACGTACGT", "TTTTTTTT") == myers_dp("ACGTACGT", "TTTTTTTT") == 6


def test_myers_vs_dp_random():
    rng = random.Random(11)
    alpha = "ACGT" * 6 + "NRYX"
    for _ in range(400):
        m = rng.randint(1, 40)
        p = "".join(rng.choice("ACGTACGTACGTNRS") for _ in range(m))
        t = "".join(rng.choice(alpha) for _ in range(rng.randint(1, 45)))
        assert oracle.myers_score(p, t) == myers_dp(p, t), (p, t)


# --------------------------------------------------------------------------- traceback_all
def cmp_loose(p, t):
    pv, tv = b4(p), b4(t)
    if tv == 0:
        return False
    if pv == 15:
        return True
    if tv == 15:
        return False
    return (pv & tv) != 0


def brute_enum(pattern, text, E, D, R, M):
    """Every op string aligning the whole pattern to a text suffix ending at len(text)."""
    m, n = len(pattern), len(text)
    isn = [c in "Nn" for c in pattern]
    out = set()

    def rec(i, j, ops, mm, db, rb):
        if i >= 1 and j >= 1 and b4(text[j - 1]) == 0:
            return                          # DP cell at padding is infinite (traceback.rs:297-300)
        if mm > M or db > D or rb > R or mm + db + rb > E:
            return
        if i == 0:
            fwd = ops[::-1]
            pa, ta, pi, ti = [], [], 0, j
            for op in fwd:
                if op in "MS":
                    pa.append(DEC[b4(pattern[pi])]); ta.append(DEC[b4(text[ti])]); pi += 1; ti += 1
                elif op == "R":
                    pa.append(DEC[b4(pattern[pi])]); ta.append("-"); pi += 1
                else:
                    pa.append("-"); ta.append(DEC[b4(text[ti])]); ti += 1
            out.add(("".join(pa), "".join(ta), mm, db, rb, j))
            return
        if j > 0:
            mc = 0 if cmp_loose(pattern[i - 1], text[j - 1]) else 1
            if not (mc and isn[i - 1]):
                rec(i - 1, j - 1, ops + ("S" if mc else "M"), mm + mc, db, rb)
        if rb < R:
            if not isn[i - 1] and not (i < m and isn[i]) and not (i >= 2 and isn[i - 2]):
                rec(i - 1, j, ops + "R", mm, db, rb + 1)
        if db < D and j > 0 and i < m:
            if not (isn[i - 1] and isn[i]):
                rec(i, j - 1, ops + "D", mm, db + 1, rb)

    rec(m, n, "", 0, 0, 0)
    return out


def test_traceback_all_vs_brute_force():
    rng = random.Random(5)
    for _ in range(300):
        m = rng.randint(2, 8)
        core = "".join(rng.choice("ACGT") for _ in range(m))
        pat = rng.choice([core, core + "NN", "NN" + core, core[:2] + "R" + core[2:]])
        n = rng.randint(1, len(pat) + 3)
        text = "".join(rng.choice("ACGTACGTNRX") for _ in range(n))
        D, R = rng.randint(0, 2), rng.randint(0, 2)
        M = rng.randint(0, 3)
        E = M + D + R
        got = oracle.traceback_all(pat, text, E, D, R, M)
        assert len(got) == len(set(got))
        assert set(got) == brute_enum(pat, text, E, D, R, M), (pat, text, D, R, M)


def test_traceback_plan_vectors():
    # plan doc :621-668 (single-optimum traceback); each optimum is among the
    # enumerated alignments
    assert ("ACGT", "ACGT", 0, 0, 0, 0) in oracle.traceback_all("ACGT", "ACGT", 2, 1, 1, 2)
    assert ("AC-GT", "ACXGT", 0, 1, 0, 0) not in oracle.traceback_all("ACGT", "ACXGT", 2, 1, 1, 2)  # X is padding
    assert ("AC-GT", "ACAGT", 0, 1, 0, 0) in oracle.traceback_all("ACGT", "ACAGT", 2, 1, 1, 2)
    assert ("ACGT", "AC-T", 0, 0, 1, 0) in oracle.traceback_all("ACGT", "ACT", 2, 1, 1, 2)
    assert oracle.traceback_all("ACGT", "TTTT", 1, 1, 1, 1) == []


# --------------------------------------------------------------------------- boundary
def test_library_exports_every_declared_symbol():
    import re
    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "cof.h")).read()
    declared = set(re.findall(r"\b(cof_[a-z_]+)\s*\(", hdr))
    assert declared == set(cofpy.COF_SYMBOLS)
    L = cofpy.lib()
    for s in declared:
        assert hasattr(L, s), s
    assert L.cof_strerror(cofpy.COF_E_OVERFLOW) == b"output buffer too small"
