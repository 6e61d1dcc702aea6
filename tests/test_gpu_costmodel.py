"""Cost model kernels on the GPU vs the CPU references, and the schedule it drives.

* forward == numpy float32 forward (rel 1e-5, abs 1e-5)
* loss/gradient == numpy analytic gradient (rel 1e-4) and == central finite
  differences of the float64 loss (rel 2e-3 on the checked entries)
* training lowers the loss; the captured-graph training loop and the eager
  loop give bit-identical parameters, and a repeat run is identical
* chunk features == the numpy restatement of their definitions
* a search ordered by a trained or an untrained model returns the oracle's rows
"""
import random

import numpy as np
import pytest

import cofpy
import costmodel
import costmodel_ref as ref
import oracle

pytestmark = pytest.mark.gpu


def _data(seed, n=300):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 12)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] + 0.1 * rng.normal(size=n)).astype(np.float32)
    return x, t


def test_forward_matches_numpy():
    m = cofpy.Mlp(seed=11)
    p = m.params()
    x, _ = _data(1, 1000)
    np.testing.assert_allclose(m.forward(x), ref.forward(p, x, np.float32), rtol=1e-5, atol=1e-5)
    # a row count that is not a multiple of the grid
    np.testing.assert_allclose(m.forward(x[:3]), ref.forward(p, x[:3], np.float32), rtol=1e-5, atol=1e-5)
    m.close()


def test_gradient_matches_analytic_and_finite_differences():
    m = cofpy.Mlp(seed=12)
    p = m.params()
    x, t = _data(2, 257)
    loss, g = m.loss_grad(x, t)
    l64, g64 = ref.loss_grad(p.astype(np.float64), x, t)
    assert abs(loss - l64) <= 1e-4 * abs(l64)
    np.testing.assert_allclose(g, g64, rtol=1e-4, atol=1e-5 * np.abs(g64).max())
    idx = list(range(0, ref.NP, 23)) + [ref.B3]
    fd = ref.finite_diff(p.astype(np.float64), x, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=2e-3, atol=1e-5 * np.abs(fd).max())
    m.close()


def test_training_graph_equals_eager_and_is_deterministic():
    x, t = _data(3, 500)
    outs = []
    for eager in (False, True, False):
        m = cofpy.Mlp(seed=5)
        l0, _ = m.loss_grad(x, t)
        loss = m.train(x, t, 200, 3e-3, eager=eager)
        assert loss < 0.5 * l0
        outs.append(m.params())
        m.close()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_training_resumes_across_calls():
    # Adam state persists across train() calls: 2 x 100 steps == 200 steps
    x, t = _data(4, 200)
    a = cofpy.Mlp(seed=8)
    a.train(x, t, 100, 2e-3)
    a.train(x, t, 100, 2e-3)
    b = cofpy.Mlp(seed=8)
    b.train(x, t, 200, 2e-3)
    assert np.array_equal(a.params(), b.params())
    a.close()
    b.close()


def _bulge_case(tmp_path, seed=31):
    rng = random.Random(seed)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(3)]
    plants = [(g + "TGG", 0, 20, 3, 10) for g in guides]
    sg = cofpy.SynthGenome(seed, [400_001, 300_000, 150_000], n_frac=0.03, gap=(50, 3000), iupac_frac=0.0005,
                           repeats=(0.1, 6, 300, 0.1), plants=plants)
    fa = tmp_path / "g.fa"
    sg.write_fasta(fa)
    return sg, fa, [g + "NNN" for g in guides]


FILT = "N" * 20 + "NRG"


def test_chunk_features_match_restatement(tmp_path):
    sg, fa, pats = _bulge_case(tmp_path)
    n = len(pats)
    s = cofpy.Search(pats, [FILT] * n, [3] * n, [1] * n, [1] * n)
    s.load_synth(sg)
    got = s.chunk_features()
    packed = sg.packed_bytes()
    _, _, chunks, nc, _ = sg.layout()
    pams, any_set, pam_set = ref.query_inputs(pats, [FILT] * n, [True] * n)
    assert len(got) == nc
    for c in list(range(0, nc, 3)) + [nc - 1]:
        ch = chunks[c]
        v = ref.genome_nibbles(packed, ch.byte_offset, ch.end - ch.start)
        want = ref.chunk_features(v, pams, any_set, pam_set)
        np.testing.assert_allclose(got[c], want, rtol=1e-4, atol=1e-6, err_msg=f"chunk {c}")
    s.close()


def test_model_order_keeps_rows(tmp_path):
    sg, fa, pats = _bulge_case(tmp_path, 32)
    n = len(pats)
    in_text = f"{fa}\n{FILT} 1 1\n" + "".join(f"{p} 3\n" for p in pats)
    oracle.run(in_text, tmp_path / "ref.tsv", batch_nucl=1 << 30)
    want = cofpy.read_rows(tmp_path / "ref.tsv")
    assert len(want) > 0
    s = cofpy.Search(pats, [FILT] * n, [3] * n, [1] * n, [1] * n, batch_nucl=1 << 30)
    s.set_tuning(bulge_first_tiles=8, bulge_round_tiles=16)
    s.load_file(fa)
    x, t = costmodel.collect(s)
    assert x.shape[1] == 12 and len(x) == len(t) and t.max() > 0
    trained, loss = costmodel.fit(x, t, steps=300)
    assert np.isfinite(loss)
    untrained = cofpy.Mlp(seed=99)
    sums = set()
    for m in (None, trained, untrained):
        s.set_cost_model(m)
        s.search()
        s.write_tsv(tmp_path / "gpu.tsv")
        assert cofpy.read_rows(tmp_path / "gpu.tsv") == want
        st = s.stats()
        sums.add((st.bulge_hits, st.bulge_checksum))
    assert len(sums) == 1
    trained.close()
    untrained.close()
    s.close()
