"""Writer of UCSC .2bit files for tests (test infrastructure).

Format (version 0, little-endian): signature 0x1A412743, version, sequence
count, reserved; an index of (name length u8, name, record offset u32); then
per record: DNA size, N-block count, N-block starts, N-block sizes, mask-block
count, mask starts, mask sizes, reserved, and the DNA at 2 bits per base,
4 bases per byte, first base in the most significant bits, T C A G = 0 1 2 3.
Blocks are written in the order given (readers sort N blocks by start).
"""
import struct

import numpy as np

_CODE = np.full(256, 0, dtype=np.uint8)
for _c, _v in zip("TCAGtcag", (0, 1, 2, 3, 0, 1, 2, 3)):
    _CODE[ord(_c)] = _v


def pack_dna(seq):
    """2-bit packing of an A/C/G/T string (other letters pack as T)."""
    v = _CODE[np.frombuffer(seq.encode(), dtype=np.uint8)]
    pad = (-len(v)) % 4
    v = np.concatenate([v, np.zeros(pad, dtype=np.uint8)]).reshape(-1, 4)
    return ((v[:, 0] << 6) | (v[:, 1] << 4) | (v[:, 2] << 2) | v[:, 3]).astype(np.uint8).tobytes()


def record_bytes(seq, n_blocks=(), mask_blocks=()):
    out = [struct.pack("<II", len(seq), len(n_blocks))]
    out += [struct.pack(f"<{len(n_blocks)}I", *[b[0] for b in n_blocks]),
            struct.pack(f"<{len(n_blocks)}I", *[b[1] for b in n_blocks])]
    out += [struct.pack("<I", len(mask_blocks)),
            struct.pack(f"<{len(mask_blocks)}I", *[b[0] for b in mask_blocks]),
            struct.pack(f"<{len(mask_blocks)}I", *[b[1] for b in mask_blocks]),
            struct.pack("<I", 0), pack_dna(seq)]
    return b"".join(out)


def write_2bit(path, records):
    """records: list of (name, seq, n_blocks [(start, size)], mask_blocks [(start, size)])."""
    head = struct.pack("<IIII", 0x1A412743, 0, len(records), 0)
    index_len = sum(1 + len(r[0].encode()) + 4 for r in records)
    bodies = [record_bytes(r[1], r[2], r[3]) for r in records]
    off = len(head) + index_len
    index = []
    for r, b in zip(records, bodies):
        nm = r[0].encode()
        index.append(struct.pack("<B", len(nm)) + nm + struct.pack("<I", off))
        off += len(b)
    with open(path, "wb") as f:
        f.write(head + b"".join(index) + b"".join(bodies))


def expected_text(seq, n_blocks):
    """The sequence a reader yields when the N blocks do not overlap: block bases become N."""
    s = bytearray(seq.upper().encode())
    for a, n in n_blocks:
        s[a:a + n] = b"N" * len(s[a:a + n])
    return bytes(s)
