"""ctypes bindings of libcof (include/cof.h) and libcofsynth for tests and benchmarks.

The search itself runs in libcof.so (HIP kernels for gfx950 behind a C ABI);
this module only marshals arguments.  Loading fails loudly when the built
libraries are missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

COF_OK, COF_E_INVALID, COF_E_HIP, COF_E_OVERFLOW = 0, 1, 2, 3
COF_E_IO, COF_E_NOGENOME, COF_E_NOMEM, COF_E_UNSUPPORTED = 4, 5, 6, 7
COF_FLAG_COUNT_ONLY = 1
COF_FLAG_NO_GRAPH = 2
COF_FLAG_SYNC_ROUNDS = 4
COF_FLAG_TILE_STATS = 8

# Every symbol include/cof.h declares (checked by the CPU test suite).
COF_SYMBOLS = ["cof_open", "cof_set_query", "cof_load_genome_file", "cof_load_genome", "cof_search",
               "cof_write_tsv", "cof_get_stats", "cof_close", "cof_strerror", "cof_last_error",
               "cof_set_tuning", "cof_get_tuning", "cof_autotune", "cof_get_tile_stats",
               "cof_mlp_create", "cof_mlp_destroy", "cof_mlp_get_params", "cof_mlp_set_params", "cof_mlp_forward",
               "cof_mlp_loss_grad", "cof_mlp_train", "cof_chunk_features", "cof_set_cost_model",
               "cof_load_genome_file_ex", "cof_device_count", "cof_device_name", "cof_write_tsv_ex",
               "cof_search_async", "cof_search_wait", "cof_hit_seqs", "cof_chrom_name", "cof_stream_begin",
               "cof_stream_chunk", "cof_stream_end", "cof_search_batch", "cof_search_static"]
COF_WRITE_HOST = 1
COF_LOAD_HOST = 1
COF_LOAD_DEVICE = 2
MLP_IN = 12
MLP_PARAMS = 5057
COF_MLP_EAGER = 1


class CofTuning(C.Structure):
    _fields_ = [("plain_threads", C.c_uint32), ("plain_pat_chunk", C.c_uint32), ("bulge_first_tiles", C.c_uint32),
                ("bulge_round_tiles", C.c_uint32), ("verify_blocks_per_cu", C.c_uint32),
                ("bulge_queue_cap", C.c_uint32), ("verify_refill", C.c_uint32), ("plain_hit_cap", C.c_uint32),
                ("plain_queue", C.c_uint32), ("bulge_screen", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CofTileStat(C.Structure):
    _fields_ = [("first_end", C.c_uint64), ("n_ends", C.c_uint32), ("candidates", C.c_uint32),
                ("dfs_steps", C.c_uint64), ("chunk", C.c_uint32), ("_pad", C.c_uint32)]


class CofQuery(C.Structure):
    _fields_ = [("n_guides", C.c_uint32),
                ("guides", C.POINTER(C.c_char_p)),
                ("filters", C.POINTER(C.c_char_p)),
                ("max_mismatches", C.POINTER(C.c_uint32)),
                ("max_dna_bulges", C.POINTER(C.c_uint32)),
                ("max_rna_bulges", C.POINTER(C.c_uint32)),
                ("batch_nucl", C.c_uint64),
                ("flags", C.c_uint32)]


class CofChunk(C.Structure):
    _fields_ = [("chrom", C.c_uint32), ("_pad", C.c_uint32), ("start", C.c_uint64),
                ("end", C.c_uint64), ("byte_offset", C.c_uint64)]


class CofHit(C.Structure):
    _fields_ = [("chrom_pos", C.c_uint64), ("chrom", C.c_uint32), ("pattern", C.c_uint32),
                ("mismatches", C.c_uint16), ("dna_bulge", C.c_uint16), ("rna_bulge", C.c_uint16),
                ("n_ops", C.c_uint16), ("ops", C.c_uint64 * 3)]


class CofStats(C.Structure):
    _fields_ = [("search_ms", C.c_double), ("kernel_ms_plain", C.c_double),
                ("kernel_ms_filter", C.c_double), ("kernel_ms_verify", C.c_double),
                ("n_plain_launches", C.c_uint64), ("n_candidates", C.c_uint64),
                ("n_raw_hits", C.c_uint64), ("n_hits", C.c_uint64), ("genome_nucl", C.c_uint64),
                ("bulge_hits", C.c_uint64), ("bulge_checksum", C.c_uint64),
                ("verify_steps", C.c_uint64), ("verify_slots", C.c_uint64), ("n_pipeline_rounds", C.c_uint64), ("bulge_ms", C.c_double),
                ("n_redo_rounds", C.c_uint64), ("verify_busy", C.c_uint64),
                ("convert_ms", C.c_double), ("load_ms", C.c_double), ("load_device", C.c_uint64),
                ("write_ms", C.c_double), ("write_device", C.c_uint64),
                ("graph_instantiations", C.c_uint64), ("graph_updates", C.c_uint64), ("n_traced", C.c_uint64)]


_lib = None
_synth = None


def lib():
    """libcof.so with argument types set; raises if it was not built."""
    global _lib
    if _lib is None:
        path = os.path.join(BUILD, "libcof.so")
        if not os.path.exists(path):
            raise RuntimeError("libcof.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.cof_open.argtypes = [C.c_int, C.POINTER(CofQuery), C.POINTER(C.c_void_p)]
        L.cof_set_query.argtypes = [C.c_void_p, C.POINTER(CofQuery)]
        L.cof_load_genome_file.argtypes = [C.c_void_p, C.c_char_p]
        L.cof_load_genome_file_ex.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32]
        L.cof_load_genome_file_ex.restype = C.c_int
        L.cof_load_genome.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(CofChunk), C.c_uint64,
                                      C.POINTER(C.c_char_p), C.c_uint32]
        L.cof_search.argtypes = [C.c_void_p, C.POINTER(CofHit), C.c_uint64, C.POINTER(C.c_uint64)]
        L.cof_search_static.argtypes = [C.c_void_p, C.POINTER(CofHit), C.c_uint64, C.POINTER(C.c_uint64)]
        L.cof_write_tsv.argtypes = [C.c_void_p, C.c_char_p]
        L.cof_write_tsv_ex.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32]
        L.cof_search_async.argtypes = [C.c_void_p]
        L.cof_search_wait.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.cof_get_stats.argtypes = [C.c_void_p, C.POINTER(CofStats)]
        L.cof_close.argtypes = [C.c_void_p]
        L.cof_strerror.argtypes = [C.c_int]
        L.cof_strerror.restype = C.c_char_p
        L.cof_last_error.argtypes = [C.c_void_p]
        L.cof_last_error.restype = C.c_char_p
        L.cof_set_tuning.argtypes = [C.c_void_p, C.POINTER(CofTuning)]
        L.cof_get_tuning.argtypes = [C.c_void_p, C.POINTER(CofTuning)]
        L.cof_autotune.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(CofTuning)]
        L.cof_get_tile_stats.argtypes = [C.c_void_p, C.POINTER(CofTileStat), C.c_uint64, C.POINTER(C.c_uint64)]
        fp = C.POINTER(C.c_float)
        L.cof_mlp_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]
        L.cof_mlp_destroy.argtypes = [C.c_void_p]
        L.cof_mlp_destroy.restype = None
        L.cof_mlp_get_params.argtypes = [C.c_void_p, fp]
        L.cof_mlp_set_params.argtypes = [C.c_void_p, fp]
        L.cof_mlp_forward.argtypes = [C.c_void_p, fp, C.c_uint64, fp]
        L.cof_mlp_loss_grad.argtypes = [C.c_void_p, fp, fp, C.c_uint64, fp, fp]
        L.cof_mlp_train.argtypes = [C.c_void_p, fp, fp, C.c_uint64, C.c_uint32, C.c_float, C.c_uint32, fp]
        L.cof_chunk_features.argtypes = [C.c_void_p, fp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.cof_set_cost_model.argtypes = [C.c_void_p, C.c_void_p]
        L.cof_hit_seqs.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_char_p, C.c_char_p, C.c_size_t,
                                   C.POINTER(C.c_uint32)]
        L.cof_chrom_name.argtypes = [C.c_void_p, C.c_uint32]
        L.cof_chrom_name.restype = C.c_char_p
        L.cof_stream_begin.argtypes = [C.c_void_p]
        L.cof_stream_chunk.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.c_uint64, C.c_char_p]
        L.cof_stream_end.argtypes = [C.c_void_p]
        L.cof_search_batch.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.POINTER(CofChunk), C.c_uint64,
                                       C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
        for f in ("cof_open", "cof_set_query", "cof_load_genome_file", "cof_load_genome", "cof_search", "cof_write_tsv",
                  "cof_get_stats", "cof_set_tuning", "cof_get_tuning", "cof_autotune", "cof_get_tile_stats",
                  "cof_mlp_create", "cof_mlp_get_params", "cof_mlp_set_params", "cof_mlp_forward", "cof_mlp_loss_grad",
                  "cof_mlp_train", "cof_chunk_features", "cof_set_cost_model", "cof_hit_seqs", "cof_stream_begin",
                  "cof_stream_chunk", "cof_stream_end", "cof_search_batch", "cof_search_static"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def synth_lib():
    global _synth
    if _synth is None:
        path = os.path.join(BUILD, "libcofsynth.so")
        if not os.path.exists(path):
            raise RuntimeError("libcofsynth.so is not built (run __graft_entry__.build())")
        S = C.CDLL(path)
        S.cofsynth_new.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64), C.c_uint32]
        S.cofsynth_new.restype = C.c_void_p
        S.cofsynth_set_n.argtypes = [C.c_void_p, C.c_double, C.c_uint64, C.c_uint64]
        S.cofsynth_set_repeats.argtypes = [C.c_void_p, C.c_double, C.c_uint32, C.c_uint32, C.c_double]
        S.cofsynth_set_iupac.argtypes = [C.c_void_p, C.c_double]
        S.cofsynth_plant.argtypes = [C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        S.cofsynth_build.argtypes = [C.c_void_p]
        S.cofsynth_packed_bytes.argtypes = [C.c_void_p]
        S.cofsynth_packed_bytes.restype = C.c_uint64
        S.cofsynth_packed.argtypes = [C.c_void_p]
        S.cofsynth_packed.restype = C.c_void_p
        S.cofsynth_n_chunks.argtypes = [C.c_void_p]
        S.cofsynth_n_chunks.restype = C.c_uint64
        S.cofsynth_chunks.argtypes = [C.c_void_p]
        S.cofsynth_chunks.restype = C.POINTER(CofChunk)
        S.cofsynth_n_chroms.argtypes = [C.c_void_p]
        S.cofsynth_n_chroms.restype = C.c_uint32
        S.cofsynth_chrom_name.argtypes = [C.c_void_p, C.c_uint32]
        S.cofsynth_chrom_name.restype = C.c_char_p
        S.cofsynth_write_fasta.argtypes = [C.c_void_p, C.c_char_p]
        S.cofsynth_write_2bit.argtypes = [C.c_void_p, C.c_char_p]
        S.cofsynth_free.argtypes = [C.c_void_p]
        _synth = S
    return _synth


class SynthGenome:
    """Seeded synthetic genome (see csrc/synth.cpp for the model)."""

    def __init__(self, seed, chrom_lens, line_width=60, n_frac=0.0, gap=(100, 100), repeats=None,
                 iupac_frac=0.0, plants=()):
        S = synth_lib()
        arr = (C.c_uint64 * len(chrom_lens))(*chrom_lens)
        self.h = S.cofsynth_new(seed, len(chrom_lens), arr, line_width)
        if not self.h:
            raise ValueError("bad synthetic genome spec")
        if n_frac:
            S.cofsynth_set_n(self.h, n_frac, gap[0], gap[1])
        if repeats:
            S.cofsynth_set_repeats(self.h, *repeats)
        if iupac_frac:
            S.cofsynth_set_iupac(self.h, iupac_frac)
        for site, goff, glen, subs, count in plants:
            if S.cofsynth_plant(self.h, site.encode(), goff, glen, subs, count) != 0:
                raise ValueError("bad plant spec")
        S.cofsynth_build(self.h)

    def write_fasta(self, path):
        if synth_lib().cofsynth_write_fasta(self.h, str(path).encode()) != 0:
            raise IOError(path)

    def write_2bit(self, path):
        """.2bit file of the genome (N runs as N blocks; needs a genome without other IUPAC letters)."""
        if synth_lib().cofsynth_write_2bit(self.h, str(path).encode()) != 0:
            raise IOError(path)

    def layout(self):
        S = synth_lib()
        n = S.cofsynth_n_chunks(self.h)
        names = [S.cofsynth_chrom_name(self.h, i) for i in range(S.cofsynth_n_chroms(self.h))]
        return S.cofsynth_packed(self.h), S.cofsynth_packed_bytes(self.h), S.cofsynth_chunks(self.h), n, names

    def packed_bytes(self):
        ptr, nb, _, _, _ = self.layout()
        return C.string_at(ptr, nb)

    def __del__(self):
        if getattr(self, "h", None) and _synth is not None:
            _synth.cofsynth_free(self.h)
            self.h = None


class CofError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{lib().cof_strerror(status).decode()}: {msg}")
        self.status = status


class Search:
    """One libcof context: query -> resident genome -> hits."""

    def __init__(self, guides, filters, mms, dnas=None, rnas=None, batch_nucl=0, device=0, count_only=False,
                 flags=0):
        L = lib()
        q = self._make_query(guides, filters, mms, dnas, rnas, batch_nucl, count_only, flags)
        h = C.c_void_p()
        st = L.cof_open(device, C.byref(q), C.byref(h))
        if st != COF_OK:
            raise CofError(st, "cof_open")
        self.h = h

    def _make_query(self, guides, filters, mms, dnas, rnas, batch_nucl, count_only, flags=0):
        n = len(guides)
        self._keep = [g.encode() for g in guides], [f.encode() for f in filters]
        q = CofQuery()
        q.n_guides = n
        q.guides = (C.c_char_p * n)(*self._keep[0])
        q.filters = (C.c_char_p * n)(*self._keep[1])
        q.max_mismatches = (C.c_uint32 * n)(*mms)
        q.max_dna_bulges = (C.c_uint32 * n)(*(dnas or [0] * n))
        q.max_rna_bulges = (C.c_uint32 * n)(*(rnas or [0] * n))
        q.batch_nucl = batch_nucl
        q.flags = (COF_FLAG_COUNT_ONLY if count_only else 0) | flags
        self._q = q
        self._qargs = (list(guides), list(filters), list(mms), dnas, rnas, batch_nucl)
        self.flags = q.flags
        return q

    def set_flags(self, flags):
        """Same query with other COF_FLAG_* bits (COF_FLAG_COUNT_ONLY included)."""
        g, f, m, d, r, b = self._qargs
        q = self._make_query(g, f, m, d, r, b, False, flags)
        self._check(lib().cof_set_query(self.h, C.byref(q)), "set_query")

    def set_query(self, guides, filters, mms, dnas=None, rnas=None, batch_nucl=0, count_only=False, flags=0):
        q = self._make_query(guides, filters, mms, dnas, rnas, batch_nucl, count_only, flags)
        self._check(lib().cof_set_query(self.h, C.byref(q)), "set_query")

    def _check(self, st, what):
        if st != COF_OK:
            raise CofError(st, f"{what}: {lib().cof_last_error(self.h).decode()}")

    def load_file(self, path, flags=None):
        if flags is None:
            self._check(lib().cof_load_genome_file(self.h, str(path).encode()), "load_genome_file")
        else:
            self._check(lib().cof_load_genome_file_ex(self.h, str(path).encode(), flags), "load_genome_file_ex")

    def load_synth(self, sg):
        ptr, nb, chunks, n, names = sg.layout()
        arr = (C.c_char_p * len(names))(*names)
        self._check(lib().cof_load_genome(self.h, ptr, nb, chunks, n, arr, len(names)), "load_genome")

    def search(self):
        n = C.c_uint64()
        st = lib().cof_search(self.h, None, 0, C.byref(n))
        if st not in (COF_OK, COF_E_OVERFLOW):
            self._check(st, "search")
        return n.value

    def search_async(self):
        """Queues a search (count query); wait() completes it."""
        self._check(lib().cof_search_async(self.h), "search_async")

    def wait(self):
        n = C.c_uint64()
        self._check(lib().cof_search_wait(self.h, C.byref(n)), "search_wait")
        return n.value

    def hits(self):
        n = self.search()
        buf = (CofHit * max(n, 1))()
        self._check(lib().cof_search(self.h, buf, n, None), "search")
        return buf[:n]

    def hit_seqs(self, first=0, n=None, stride=128):
        """[(dna_seq, rna_seq)] of hits [first, first + n) of the last search (bytes; plain DNA may hold NUL)."""
        if n is None:
            n = self.search_count_cached() - first
        dna = C.create_string_buffer(max(1, n * stride))
        rna = C.create_string_buffer(max(1, n * stride))
        lens = (C.c_uint32 * max(1, n))()
        self._check(lib().cof_hit_seqs(self.h, first, n, dna, rna, stride, lens), "hit_seqs")
        return [(dna.raw[k * stride:k * stride + lens[k]], rna.raw[k * stride:k * stride + lens[k]])
                for k in range(n)]

    def search_count_cached(self):
        n = C.c_uint64()
        st = lib().cof_search(self.h, (CofHit * 1)(), 0, C.byref(n))
        if st not in (COF_OK, COF_E_OVERFLOW):
            self._check(st, "search")
        return n.value

    def chrom_name(self, chrom):
        v = lib().cof_chrom_name(self.h, chrom)
        return None if v is None else v.decode()

    def stream(self, chunks):
        """cof_stream_begin, one cof_stream_chunk per (name, start, end, data), cof_stream_end."""
        L = lib()
        self._check(L.cof_stream_begin(self.h), "stream_begin")
        for name, start, end, data in chunks:
            self._check(L.cof_stream_chunk(self.h, name.encode(), start, end, data), "stream_chunk")
        self._check(L.cof_stream_end(self.h), "stream_end")

    def search_batch(self, packed, chunks, names, has_head):
        """cof_search_batch over one batch: packed bytes, [(chrom, start, end, byte_offset)], names."""
        arr = (CofChunk * max(1, len(chunks)))(*[CofChunk(c, 0, s, e, o) for c, s, e, o in chunks])
        nm = (C.c_char_p * max(1, len(names)))(*[x.encode() for x in names])
        n = C.c_uint64()
        self._check(lib().cof_search_batch(self.h, packed, len(packed), arr, len(chunks), nm, len(names),
                                           1 if has_head else 0, C.byref(n)), "search_batch")
        return n.value

    def write_tsv(self, path, host=False):
        """Reference TSV rows of the last search; formatted on the device unless host=True."""
        self._check(lib().cof_write_tsv_ex(self.h, str(path).encode(), COF_WRITE_HOST if host else 0), "write_tsv")

    def stats(self):
        s = CofStats()
        lib().cof_get_stats(self.h, C.byref(s))
        return s

    def set_tuning(self, **kw):
        t = CofTuning(**kw)
        self._check(lib().cof_set_tuning(self.h, C.byref(t)), "set_tuning")

    def tuning(self):
        t = CofTuning()
        self._check(lib().cof_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def tile_stats(self):
        """numpy structured array (first_end, n_ends, candidates, dfs_steps) of the last bulge search."""
        import numpy as np
        n = C.c_uint64()
        self._check(lib().cof_get_tile_stats(self.h, None, 0, C.byref(n)), "tile_stats")
        buf = (CofTileStat * max(1, n.value))()
        self._check(lib().cof_get_tile_stats(self.h, buf, n.value, C.byref(n)), "tile_stats")
        dt = np.dtype([("first_end", "<u8"), ("n_ends", "<u4"), ("candidates", "<u4"), ("dfs_steps", "<u8"),
                       ("chunk", "<u4"), ("_pad", "<u4")])
        return np.frombuffer(bytes(buf), dtype=dt, count=n.value).copy()

    def chunk_features(self):
        """numpy float32 [n_chunks, 12]: cost-model features of every genome chunk for this query."""
        import numpy as np
        n = C.c_uint64()
        self._check(lib().cof_chunk_features(self.h, None, 0, C.byref(n)), "chunk_features")
        out = np.zeros((n.value, MLP_IN), dtype=np.float32)
        self._check(lib().cof_chunk_features(self.h, _fptr(out), n.value, C.byref(n)), "chunk_features")
        return out

    def set_cost_model(self, mlp):
        """Order bulge work by the model's predictions (None: static genome order)."""
        self._check(lib().cof_set_cost_model(self.h, mlp.h if mlp is not None else None), "set_cost_model")

    def autotune(self, sample_nucl=0):
        t = CofTuning()
        self._check(lib().cof_autotune(self.h, sample_nucl, C.byref(t)), "autotune")
        return t.as_dict()

    def close(self):
        if getattr(self, "h", None):
            lib().cof_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """Device-resident FP32 MLP 12 -> 64 -> 64 -> 1 (the bulge scheduler's cost model)."""

    def __init__(self, seed=1, device=0):
        h = C.c_void_p()
        st = lib().cof_mlp_create(device, seed, C.byref(h))
        if st != COF_OK:
            raise CofError(st, "mlp_create")
        self.h = h

    def _check(self, st, what):
        if st != COF_OK:
            raise CofError(st, what)

    def params(self):
        import numpy as np
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(lib().cof_mlp_get_params(self.h, _fptr(p)), "get_params")
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (MLP_PARAMS,)
        self._check(lib().cof_mlp_set_params(self.h, _fptr(p)), "set_params")

    def forward(self, x):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.zeros(len(x), dtype=np.float32)
        self._check(lib().cof_mlp_forward(self.h, _fptr(x), len(x), _fptr(y)), "forward")
        return y

    def loss_grad(self, x, t):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().cof_mlp_loss_grad(self.h, _fptr(x), _fptr(t), len(x), C.byref(loss), _fptr(g)), "loss_grad")
        return loss.value, g

    def train(self, x, t, steps, lr=1e-3, eager=False):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        self._check(lib().cof_mlp_train(self.h, _fptr(x), _fptr(t), len(x), steps, lr,
                                        COF_MLP_EAGER if eager else 0, C.byref(loss)), "train")
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            lib().cof_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def read_rows(path):
    """Sorted data rows of a reference-format TSV (header dropped)."""
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines and lines[0].startswith("#Bulge type"), "missing header"
    return sorted(lines[1:])
