"""Schedules and launch shapes never change results.

The bulge path's rounds run as a two-stream filter/traceback pipeline captured
into a hipGraph (default), the same pipeline launched eagerly, or synchronous
rounds; the plain scan takes several workgroup sizes and pattern slices; the
autotuner picks among them.  Every variant must give the oracle's rows
(reference semantics: search.rs:1054-1191 for the bulge batch loop,
search.rs:322-366 for the plain launch).  Small rounds and a small queue force
many rounds, device-side overflow detection and the redo-in-halves path.
"""
import random

import numpy as np
import pytest

import cofpy
import oracle

pytestmark = pytest.mark.gpu

FILT = "N" * 20 + "NRG"


def _bulge_genome(tmp_path):
    rng = random.Random(2024)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(3)]
    plants = [(g + "AGG", 0, 20, 3, 20) for g in guides]
    sg = cofpy.SynthGenome(4242, [700_000, 500_000, 130_001], n_frac=0.02, gap=(50, 2000),
                           repeats=(0.1, 6, 300, 0.1), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    return fa, [g + "NNN" for g in guides]


def _oracle(tmp_path, fa, pats, M, D, R):
    in_text = f"{fa}\n{FILT} {D} {R}\n" + "".join(f"{p} {M}\n" for p in pats)
    oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=1 << 30)
    return cofpy.read_rows(tmp_path / "ref.tsv")


def _rows(s, tmp_path):
    s.search()
    s.write_tsv(tmp_path / "gpu.tsv")
    return cofpy.read_rows(tmp_path / "gpu.tsv")


def _bulge_search(fa, pats, flags=0, count_only=False, **tune):
    n = len(pats)
    s = cofpy.Search(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30, count_only=count_only,
                     flags=flags)
    if tune:
        s.set_tuning(**tune)
    s.load_file(fa)
    return s


@pytest.mark.parametrize("flags", [0, cofpy.COF_FLAG_NO_GRAPH, cofpy.COF_FLAG_SYNC_ROUNDS])
def test_bulge_schedules_match_oracle(tmp_path, flags):
    fa, pats = _bulge_genome(tmp_path)
    want = _oracle(tmp_path, fa, pats, 3, 1, 1)
    assert len(want) > 0
    s = _bulge_search(fa, pats, flags | cofpy.COF_FLAG_TILE_STATS, bulge_first_tiles=16, bulge_round_tiles=24)
    got = _rows(s, tmp_path)
    st = s.stats()
    assert got == want
    if flags & cofpy.COF_FLAG_SYNC_ROUNDS:
        assert st.n_pipeline_rounds == 0
    else:
        assert st.n_pipeline_rounds >= 10
    # per-tile accounting adds up to the totals
    ts = s.tile_stats()
    assert int(ts["candidates"].sum()) == st.n_candidates
    assert int(ts["dfs_steps"].sum()) == st.verify_steps
    assert np.all(np.diff(ts["first_end"].astype(np.int64)) > 0)
    # without the flag, candidates are still counted but DFS steps are not
    n = len(pats)
    s.set_query(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30, flags=flags)
    assert _rows(s, tmp_path) == want
    ts2 = s.tile_stats()
    assert np.array_equal(ts2["candidates"], ts["candidates"])
    assert int(ts2["dfs_steps"].sum()) == 0
    s.close()


def test_graph_replays_interleaved_with_eager(tmp_path):
    # each search captures the pipeline and replays the context's graph;
    # synchronous searches in between share the workspace, counters and queues
    fa, pats = _bulge_genome(tmp_path)
    want = _oracle(tmp_path, fa, pats, 3, 1, 1)
    n = len(pats)
    s = _bulge_search(fa, pats, 0, bulge_first_tiles=8, bulge_round_tiles=16)
    sums = set()
    for it in range(3):
        for flags in (0, cofpy.COF_FLAG_SYNC_ROUNDS):
            s.set_query(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30, flags=flags)
            assert _rows(s, tmp_path) == want, (it, flags)
            st = s.stats()
            sums.add((st.bulge_hits, st.bulge_checksum))
            s.set_query(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30, flags=flags,
                        count_only=True)
            s.search()
            st = s.stats()
            sums.add((st.bulge_hits, st.bulge_checksum))
    assert len(sums) == 1
    s.close()


def test_graph_updated_in_place_and_replayed(tmp_path):
    # the first graph search instantiates the pipeline graph; the next ones,
    # with the same pipeline shape, update it in place and replay it, and an
    # eager search in between leaves it valid
    fa, pats = _bulge_genome(tmp_path)
    want = _oracle(tmp_path, fa, pats, 3, 1, 1)
    n = len(pats)
    s = _bulge_search(fa, pats, 0, bulge_first_tiles=8, bulge_round_tiles=16)
    assert _rows(s, tmp_path) == want
    st = s.stats()
    assert st.graph_instantiations >= 1 and st.graph_updates == 0
    for flags in (0, cofpy.COF_FLAG_NO_GRAPH, 0):
        s.set_query(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30, flags=flags)
        assert _rows(s, tmp_path) == want, flags
        st = s.stats()
        if flags == 0:
            assert st.graph_instantiations == 0 and st.graph_updates >= 1
        else:
            assert st.graph_instantiations == 0 and st.graph_updates == 0
    s.close()


@pytest.mark.parametrize("flags", [0, cofpy.COF_FLAG_SYNC_ROUNDS])
def test_queue_overflow_redo(tmp_path, flags):
    fa, pats = _bulge_genome(tmp_path)
    want = _oracle(tmp_path, fa, pats, 3, 1, 1)
    s = _bulge_search(fa, pats, flags)
    assert _rows(s, tmp_path) == want
    cand = s.tile_stats()["candidates"].astype(np.int64)
    # a queue that holds the densest tile but not a whole round of them
    cap = max(64, int(cand.max()))
    # a round length at which every full round overflows, whichever tiles it
    # holds: the rt lightest tiles already exceed the queue
    light = np.cumsum(np.sort(cand))
    rt = int(np.argmax(light > cap)) + 1
    assert light[rt - 1] > cap and rt < len(cand)
    s.set_tuning(bulge_first_tiles=rt, bulge_round_tiles=rt, bulge_queue_cap=cap)
    assert _rows(s, tmp_path) == want
    assert s.stats().n_redo_rounds > 0
    s.close()


def _plain_case(tmp_path, n_guides=40):
    rng = random.Random(77)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(n_guides)]
    plants = [(g + "TGG", 0, 20, 4, 3) for g in guides]
    sg = cofpy.SynthGenome(91, [400_000, 250_007], n_frac=0.02, gap=(30, 900), iupac_frac=0.0005,
                           plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    pats = [g + "NNN" for g in guides]
    in_text = f"{fa}\n{FILT}\n" + "".join(f"{p} 4\n" for p in pats)
    oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=1 << 22)
    return fa, pats, cofpy.read_rows(tmp_path / "ref.tsv")


@pytest.mark.parametrize("threads,chunk", [(128, 0), (256, 16), (512, 7), (512, 0), (128, 64)])
def test_plain_launch_shapes(tmp_path, threads, chunk):
    fa, pats, want = _plain_case(tmp_path)
    assert len(want) > 0
    s = cofpy.Search(pats, [FILT] * len(pats), [4] * len(pats), batch_nucl=1 << 22)
    s.set_tuning(plain_threads=threads, plain_pat_chunk=chunk)
    s.load_file(fa)
    assert s.tuning()["plain_threads"] == threads
    assert _rows(s, tmp_path) == want
    s.close()


def test_plain_queue_kernel(tmp_path):
    # the persistent scan (two-wave workgroups draining per-XCD tile queues):
    # one matrix-core group per wave, and a query of four groups in one launch
    # (two filters of one length; the second set's guides hold an IUPAC letter,
    # so its groups compare bit4 words) that the waves take alternately; rows
    # equal the oracle's (the union of one run per filter for the mixed query)
    fa, pats, want = _plain_case(tmp_path)
    s = cofpy.Search(pats, [FILT] * len(pats), [4] * len(pats), batch_nucl=1 << 22)
    s.set_tuning(plain_queue=1)
    s.load_file(fa)
    assert s.tuning()["plain_queue"] == 1
    assert _rows(s, tmp_path) == want
    s.close()
    rng = random.Random(5)
    gfilt = "N" * 20 + "NGG"   # same length as FILT: one launch holds all four groups
    g1 = ["".join(rng.choice("ACGT") for _ in range(20)) + "NNN" for _ in range(12)]
    g2 = ["".join(rng.choice("ACGT") for _ in range(19)) + "RNNN" for _ in range(9)]
    want = set()
    for filt, gs, m in ((FILT, g1, 3), (gfilt, g2, 4)):
        in_text = f"{fa}\n{filt}\n" + "".join(f"{p} {m}\n" for p in gs)
        oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=1 << 22)
        want |= set(cofpy.read_rows(tmp_path / "ref.tsv"))
    # the same mixed launch on the persistent scan and on the default grid
    # scan (where a launch with a bit4 group runs every tile as a flagged
    # tile, so its plane groups compare bit4 words too)
    for queue in (1, 0):
        s = cofpy.Search(g1 + g2, [FILT] * len(g1) + [gfilt] * len(g2), [3] * len(g1) + [4] * len(g2),
                         batch_nucl=1 << 22)
        s.set_tuning(plain_queue=queue)
        s.load_file(fa)
        got = _rows(s, tmp_path)
        assert sorted(got) == sorted(want) and len(got) == len(want), queue
        s.close()


def test_autotune_keeps_results(tmp_path):
    fa, pats, want = _plain_case(tmp_path)
    n = len(pats)
    s = cofpy.Search(pats, [FILT] * n, [4] * n, batch_nucl=1 << 22)
    s.load_file(fa)
    t = s.autotune(1 << 20)
    # which shape wins is a timing outcome; only its validity is checked
    assert t["plain_threads"] in (64, 128, 256, 512)
    assert t["plain_pat_chunk"] in (0, 16, 64)
    assert t["plain_queue"] in (0, 1)
    assert s.tuning() == t
    assert _rows(s, tmp_path) == want
    s.close()
    fa, bpats = _bulge_genome(tmp_path)
    bwant = _oracle(tmp_path, fa, bpats, 3, 1, 1)
    s = _bulge_search(fa, bpats)
    t = s.autotune(1 << 20)
    assert t["verify_blocks_per_cu"] in (4, 8, 12, 16, 24)
    assert t["bulge_round_tiles"] in (0, 256, 1024, 4096)
    assert t["verify_refill"] in (8, 16, 32)
    assert s.tuning() == t
    assert _rows(s, tmp_path) == bwant
    s.close()


def test_invalid_tuning_rejected():
    s = cofpy.Search(["A" * 20 + "NNN"], [FILT], [2])
    for kw in ({"plain_threads": 100}, {"verify_blocks_per_cu": 65}, {"bulge_queue_cap": 10},
               {"plain_hit_cap": 100}, {"plain_queue": 2}, {"bulge_screen": 3}):
        with pytest.raises(cofpy.CofError) as e:
            s.set_tuning(**kw)
        assert e.value.status == cofpy.COF_E_INVALID
    s.close()
