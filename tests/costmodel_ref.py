"""CPU references for the bulge scheduler's cost model (test infrastructure).

* MLP 12 -> 64 -> 64 -> 1 with ReLU: forward, loss sum((y - t)^2) / n and its
  analytic gradient in numpy (float32 or float64), plus central finite
  differences of the same loss.  Parameter layout: W1[64][12], b1[64],
  W2[64][64], b2[64], w3[64], b3.
* The 12 chunk features, restated position by position from their
  definitions (GC, N, 8-mer entropy, 8-mer repeats, CpG, homopolymers, PAM
  hits, guide-seed hits, IUPAC, chunk fill), and the query-side inputs
  (distinct PAM cell lists, guide 8-mer sets).
"""
import math

import numpy as np

IN, H = 12, 64
W1, B1, W2, B2, W3, B3 = 0, 768, 832, 4928, 4992, 5056
NP = 5057


def unpack(p):
    p = np.asarray(p)
    return (p[W1:B1].reshape(H, IN), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, x, dtype=np.float64):
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, dtype=dtype) for a in unpack(p))
    x = np.asarray(x, dtype=dtype).reshape(-1, IN)
    z1 = x @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3


def loss(p, x, t, dtype=np.float64):
    y = forward(p, x, dtype)
    t = np.asarray(t, dtype=dtype)
    return float(np.sum((y - t) ** 2) / len(t))


def loss_grad(p, x, t, dtype=np.float64):
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, dtype=dtype) for a in unpack(p))
    x = np.asarray(x, dtype=dtype).reshape(-1, IN)
    t = np.asarray(t, dtype=dtype)
    n = len(t)
    z1 = x @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    e = y - t
    dy = 2 * e / n
    g = np.zeros(NP, dtype=dtype)
    g[W3:B3] = dy @ h2
    g[B3] = dy.sum()
    dz2 = np.outer(dy, w3) * (z2 > 0)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (dz2 @ w2) * (z1 > 0)
    g[W1:B1] = (dz1.T @ x).ravel()
    g[B1:W2] = dz1.sum(0)
    return float(np.sum(e ** 2) / n), g


def finite_diff(p, x, t, idx, h=1e-6):
    p = np.asarray(p, dtype=np.float64)
    out = []
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        out.append((loss(a, x, t) - loss(b, x, t)) / (2 * h))
    return np.array(out)


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
_ENC = {c: v for c, v in zip("TCAGRYSWKMBDHVN", [1, 2, 4, 8, 12, 3, 10, 5, 9, 6, 11, 13, 7, 14, 15])}
_COMP = str.maketrans("ACGTRYSWKMBDHVNacgtryswkmbdhvn", "TGCAYRSWMKVHDBNtgcayrswmkvhdbn")


def revcomp(s):
    return s.translate(_COMP)[::-1]


def enc(c):
    return _ENC.get(c.upper(), 0)


CODE2 = {1: 0, 2: 1, 4: 2, 8: 3}


def query_inputs(guides, filters, bulge_flags):
    """Distinct PAM cell lists and guide 8-mer sets of a query (forward then reverse patterns)."""
    pats = [(g, f, b) for g, f, b in zip(guides, filters, bulge_flags)]
    pats += [(revcomp(g), revcomp(f), b) for g, f, b in zip(guides, filters, bulge_flags)]
    any_bulge = any(bulge_flags)
    pams, seen, any_set, pam_set = [], set(), set(), set()
    for g, f, b in pats:
        if any_bulge and not b:
            continue
        cells = [(k, enc(f[k])) for k in range(len(f)) if f[k] not in "Nn"][:8]
        if cells:
            key = tuple((k - cells[0][0], c) for k, c in cells)
            if key not in seen and len(pams) < 16:
                seen.add(key)
                pams.append(key)
        b4 = [enc(c) for c in g]
        L = len(b4)

        def code(s):
            v = 0
            for k in range(s, s + 8):
                if b4[k] not in CODE2:
                    return None
                v = (v << 2) | CODE2[b4[k]]
            return v

        for s in range(L - 7):
            c = code(s)
            if c is not None:
                any_set.add(c)
        e = L
        while e > 0 and b4[e - 1] == 15:
            e -= 1
        bb = 0
        while bb < L and b4[bb] == 15:
            bb += 1
        c = None
        if e < L and e >= 8:
            c = code(e - 8)
        elif bb > 0 and bb + 8 <= L:
            c = code(bb)
        if c is not None:
            pam_set.add(c)
    return pams, any_set, pam_set


def chunk_features(v, pams, any_set, pam_set):
    """The 12 features of one chunk given its nibbles v (np.uint8)."""
    v = np.asarray(v, dtype=np.int64)
    n = len(v)
    single = np.isin(v, [1, 2, 4, 8])
    n_acgt = int(single.sum())
    n_gc = int(np.isin(v, [2, 8]).sum())
    n_n = int((v == 15).sum())
    n_iupac = int((~single & (v != 15) & (v != 0)).sum())
    n_cpg = int(((v[:-1] == 2) & (v[1:] == 8)).sum()) if n > 1 else 0
    n_homo = 0
    for p in range(0, n - 5):
        if single[p] and np.all(v[p:p + 6] == v[p]):
            n_homo += 1
    n_pam = 0
    for key in pams:
        last = key[-1][0]
        for p in range(0, n - last):
            ok = True
            for rel, code in key:
                gv = v[p + rel]
                if not (gv != 0 and gv != 15 and (gv & code)):
                    ok = False
                    break
            n_pam += ok
    kmers = []
    c2 = np.full(n, -1, dtype=np.int64)
    for k, val in CODE2.items():
        c2[v == k] = val
    for p in range(0, n - 7):
        w = c2[p:p + 8]
        if np.all(w >= 0):
            km = 0
            for x in w:
                km = (km << 2) | int(x)
            kmers.append(km)
    kmers = np.array(kmers, dtype=np.int64)
    total = len(kmers)
    counts = np.bincount(kmers, minlength=65536) if total else np.zeros(65536, dtype=np.int64)
    if total:
        pr = counts[counts > 0] / total
        ent = float(-(pr * np.log2(pr)).sum())
        rep = int((counts[kmers] >= 4).sum())
    else:
        ent, rep = 0.0, 0
    n_any = sum(1 for k in kmers if int(k) in any_set)
    n_ps = sum(1 for k in kmers if int(k) in pam_set)
    fn = max(n, 1)
    return np.array([
        n_gc / n_acgt if n_acgt else 0.0,
        n_n / fn,
        ent / 16,
        rep / total if total else 0.0,
        n_cpg / fn,
        n_homo / fn,
        n_pam / (fn * len(pams)) if pams else 0.0,
        n_any / fn,
        n_ps / fn,
        math.log2(1 + n_any) / 16,
        n_iupac / fn,
        n / 65536,
    ])


def genome_nibbles(packed, byte_offset, size):
    b = np.frombuffer(packed, dtype=np.uint8, count=(size + 1) // 2, offset=byte_offset)
    out = np.empty(2 * len(b), dtype=np.uint8)
    out[0::2] = b & 0xF
    out[1::2] = b >> 4
    return out[:size]
