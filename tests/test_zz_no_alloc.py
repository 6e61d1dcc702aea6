"""cof_search_static (the allocation-free plain search) against cof_search, on the GPU.

cof_search_static runs the plain scan with its hit segments, counters and
packed hits in a static device array and converts the hits straight into the
caller's buffer.  Its records must equal cof_search's (bit for bit, as a
multiset) on two calls in a row, the second on a smaller genome; a search
whose hits overflow the static array is refused with COF_E_NOMEM and a
message; a too-small caller buffer gets COF_E_OVERFLOW with nothing written.
"""
import ctypes as C
import random

import pytest

import cofpy

pytestmark = pytest.mark.gpu


def _records(hits):
    return sorted(bytes(h) for h in hits)


def _static(s, cap=None):
    L = cofpy.lib()
    n = C.c_uint64()
    assert L.cof_search_static(s.h, None, 0, C.byref(n)) == cofpy.COF_OK, L.cof_last_error(s.h)
    k = n.value if cap is None else cap
    buf = (cofpy.CofHit * max(1, k))()
    st = L.cof_search_static(s.h, buf, k, C.byref(n))
    return st, n.value, buf


def test_static_search_equals_cof_search():
    rng = random.Random(4)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(6)]
    pats = [g + "NNN" for g in guides]
    big = cofpy.SynthGenome(41, [400_000, 65520 * 2 + 77, 9_001], n_frac=0.02, gap=(20, 500), iupac_frac=0.001,
                            plants=[(g + "AGG", 0, 20, 4, 40) for g in guides])
    small = cofpy.SynthGenome(42, [120_001, 5_000], n_frac=0.01, gap=(20, 300),
                              plants=[(g + "TGG", 0, 20, 3, 10) for g in guides])
    # 2-chunk virtual batches: seams and the head-chunk start limit are in play
    s = cofpy.Search(pats, ["N" * 20 + "NRG"] * 6, [5] * 6, batch_nucl=1 << 17)
    for sg in (big, small):
        s.load_synth(sg)
        want = _records(s.hits())
        st, n, buf = _static(s)
        assert st == cofpy.COF_OK
        assert n == len(want) and n > 0
        assert _records(buf[:n]) == want
        # a buffer one short: COF_E_OVERFLOW, the count, nothing written
        short = (cofpy.CofHit * (n - 1))()
        C.memset(short, 0x5A, C.sizeof(short))
        cnt = C.c_uint64()
        assert cofpy.lib().cof_search_static(s.h, short, n - 1, C.byref(cnt)) == cofpy.COF_E_OVERFLOW
        assert cnt.value == n and bytes(short) == b"\x5a" * C.sizeof(short)
    s.close()


def test_static_search_limits():
    # hits beyond the static array are refused with a message; cof_search
    # still serves the query; bulge queries are not served
    rng = random.Random(9)
    guides = ["".join(rng.choice("ACGT") for _ in range(12)) for _ in range(4)]
    sg = cofpy.SynthGenome(3, [2_000_000])
    s = cofpy.Search(guides, ["N" * 12] * 4, [6] * 4)
    s.load_synth(sg)
    n = C.c_uint64()
    assert cofpy.lib().cof_search_static(s.h, None, 0, C.byref(n)) == cofpy.COF_E_NOMEM
    assert b"static scratch" in cofpy.lib().cof_last_error(s.h)
    assert s.search() > 524_288
    s.close()
    b = cofpy.Search(["A" * 20 + "NNN"], ["N" * 20 + "NGG"], [2], [1], [1])
    b.load_synth(sg)
    assert cofpy.lib().cof_search_static(b.h, None, 0, C.byref(n)) == cofpy.COF_E_UNSUPPORTED
    b.close()
