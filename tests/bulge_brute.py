"""Brute-force bulge search (test checker): rows of a reference bulge run from first principles.

Used to pin the CPU oracle's bulge path (check_pam_quick, Myers filter,
traceback_all, the full PAM check of classify_myers_candidate, the
convert_matches_myers drop rule and the batch seams) without sharing its
algorithm.  The enumeration core (tests/brute/bulge_enum.cpp) walks every
op string forward from every start; this module supplies the genome and
turns alignments into reference rows:

* the genome is the reader's chunk list (the readers are pinned separately
  by the reference's known-answer fixtures) laid out tight-packed as the
  batcher does (search.rs:1376-1427: ceil(n/2) bytes per chunk, so an odd
  chunk leaves a 0 nibble), as ONE text with no batches at all;
* a hit is kept iff its alignment lies inside one chromosome.  That is what
  the reference's per-batch scan plus its drop rule (search.rs:1338-1350:
  drop a hit whose start chunk ends a chromosome or a batch and which runs
  past that chunk's end) amount to, when every chunk is at least a text
  window long: each end position is scanned in exactly one batch, a batch's
  head chunk gives every window its full width, and an alignment can only
  leave its chromosome from that chromosome's last chunk.  rows() checks the
  chunk-size condition and refuses genomes that break it;
* rows follow the writer's bulge branch (main.rs:145-157, 224-242): strings
  decoded from bit4 with gaps, reverse-strand hits reverse-complemented,
  type X / DNA / RNA, size db + rb, reverse-strand position shifted by
  (trailing Ns of the first guide) - 1 - db, saturating.
"""
import bisect
import ctypes as C
import os
import subprocess
import tempfile

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

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


def lib():
    global _lib
    if _lib is None:
        out = os.path.join(HERE, "brute", "_build")
        os.makedirs(out, exist_ok=True)
        so = os.path.join(out, "libbulgebrute.so")
        src = os.path.join(HERE, "brute", "bulge_enum.cpp")
        if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
            subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so, src], check=True)
        L = C.CDLL(so)
        L.brute_bulge_enum.restype = C.c_int64
        L.brute_bulge_enum.argtypes = [C.c_char_p, C.c_uint64, C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_char_p]
        _lib = L
    return _lib


def genome_layout(path):
    """Reader chunks -> (text nibbles, chunk table [(name, start, end, nibble offset, chrom id)])."""
    seq, table = oracle.decode_genome(path)
    chunks, text, k, chrom = [], bytearray(), 0, -1
    for line in table.splitlines():
        name, start, end = line.split("\t")
        start, end = int(start), int(end)
        if start == 0:
            chrom += 1
        n = end - start
        chunks.append((name, start, end, len(text), chrom))
        text += bytes(b4(chr(c)) for c in seq[k:k + n])
        if n & 1:
            text.append(0)
        k += n
    return bytes(text), chunks


def rows(genome, filt, guides, M, D, R, quick_gate=False):
    """Sorted rows of the reference bulge run `genome / filt D R / guides M`."""
    text, chunks = genome_layout(genome)
    L = len(filt)
    W = L + D
    for c in chunks:
        if c[2] - c[1] < W + 1:
            raise ValueError(f"chunk {c[:3]} is shorter than a text window: the batch-free view would not hold")
    offs = [c[3] for c in chunks]
    pats = list(guides) + [revcomp(g) for g in guides]
    n_fwd = len(guides)
    filt_rc = revcomp(filt)
    pam_len = len(guides[0]) - len(guides[0].rstrip("Nn"))
    out = []
    with tempfile.TemporaryDirectory() as d:
        for pi, p in enumerate(pats):
            assert len(p) == L
            f = filt if pi < n_fwd else filt_rc
            pb = bytes(b4(c) for c in p)
            isn = bytes(1 if c in "Nn" else 0 for c in p)
            fb = bytes(b4(c) for c in f)
            path = os.path.join(d, "al.txt")
            n = lib().brute_bulge_enum(text, len(text), pb, isn, fb, L, M, D, R, 1 if quick_gate else 0,
                                       path.encode())
            if n < 0:
                raise RuntimeError("brute force failed")
            for line in open(path):
                j, s, mm, db, rb, ops = line.split()
                j, s, mm, db, rb = int(j), int(s), int(mm), int(db), int(rb)
                cs = chunks[bisect.bisect_right(offs, s) - 1]
                ce = chunks[bisect.bisect_right(offs, j) - 1]
                if cs[4] != ce[4]:
                    continue                          # leaves its chromosome
                pa, ta, pp, t = [], [], 0, s
                for op in ops:
                    if op in "MS":
                        pa.append(DEC[pb[pp]]); ta.append(DEC[text[t]]); pp += 1; t += 1
                    elif op == "R":
                        pa.append(DEC[pb[pp]]); ta.append("-"); pp += 1
                    else:
                        pa.append("-"); ta.append(DEC[text[t]]); t += 1
                rna, dna = "".join(pa), "".join(ta)
                fwd = pi < n_fwd
                if not fwd:
                    rna, dna = revcomp(rna), revcomp(dna)
                pos = cs[1] + (s - cs[3])
                if not fwd:
                    pos += max(max(pam_len - 1, 0) - db, 0)
                btype = "X" if db + rb == 0 else ("DNA" if db > 0 else "RNA")
                out.append(f"{btype}\t{rna}\t{dna}\t{cs[0]}\t{pos}\t{'+' if fwd else '-'}\t{mm}\t{db + rb}")
    return sorted(out)
