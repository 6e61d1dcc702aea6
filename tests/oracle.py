"""ctypes access to the CPU oracle (oracle/cof_ref.cpp).  Test infrastructure only."""
import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_ref = None


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(ROOT, "oracle", "_build", "libcofref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built (make -C oracle)")
        R = C.CDLL(path)
        R.cofref_run.restype = C.c_int64
        R.cofref_run.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, C.c_uint, C.POINTER(C.c_double),
                                 C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
        R.cofref_decode_genome.restype = C.c_int64
        R.cofref_decode_genome.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]
        R.cofref_myers_score.restype = C.c_int
        R.cofref_myers_score.argtypes = [C.c_char_p, C.c_char_p]
        R.cofref_traceback_all.restype = C.c_int64
        R.cofref_traceback_all.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_char_p, C.c_size_t]
        _ref = R
    return _ref


def run(in_text, out_path, batch_nucl=0, threads=0):
    """Runs the oracle on a reference-format input text; returns (rows written, stats)."""
    err = C.create_string_buffer(1024)
    secs = C.c_double()
    ncand = C.c_uint64()
    n = ref().cofref_run(in_text.encode(), str(out_path).encode(), batch_nucl, threads, C.byref(secs),
                         C.byref(ncand), err, 1024)
    if n < 0:
        raise RuntimeError(err.value.decode())
    return n, {"search_seconds": secs.value, "n_candidates": ncand.value}


def decode_genome(path, cap=1 << 24):
    seq = C.create_string_buffer(cap)
    tab = C.create_string_buffer(1 << 20)
    n = ref().cofref_decode_genome(str(path).encode(), seq, cap, tab, 1 << 20)
    if n < 0:
        raise RuntimeError(f"decode failed: {n}")
    return seq.raw[:n], tab.value.decode()


def myers_score(pattern, text):
    return ref().cofref_myers_score(pattern.encode(), text.encode())


def traceback_all(pattern, text, max_edits, D, R, M):
    buf = C.create_string_buffer(1 << 22)
    n = ref().cofref_traceback_all(pattern.encode(), text.encode(), max_edits, D, R, M, buf, 1 << 22)
    if n < 0:
        raise RuntimeError("buffer too small")
    out = []
    for line in buf.value.decode().splitlines():
        pa, ta, mm, db, rb, ts = line.split("\t")
        out.append((pa, ta, int(mm), int(db), int(rb), int(ts)))
    return out
