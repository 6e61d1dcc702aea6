"""CPU checks of the cost model's references and host logic.

The numpy analytic gradient (the reference the device backward is compared
with) is pinned by central finite differences in float64; the folding of
input/target standardisation into the first and last layer is checked
against the unfolded model; the feature restatement is checked against
closed-form values on hand-made chunks.
"""
import math

import numpy as np

import costmodel
import costmodel_ref as ref


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, ref.NP)
    return p


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    x = rng.normal(size=(37, 12))
    t = rng.normal(size=37)
    _, g = ref.loss_grad(p, x, t)
    idx = list(range(0, ref.NP, 17)) + [ref.B3, ref.W3, ref.B1, ref.B2 + 5]
    fd = ref.finite_diff(p, x, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_relu_kinks_and_zero_rows():
    # all-negative first layer: only b3 and w3 paths are live
    p = np.zeros(ref.NP)
    p[ref.B1:ref.W2] = -1.0
    p[ref.B3] = 0.5
    x = np.ones((4, 12))
    t = np.zeros(4)
    l, g = ref.loss_grad(p, x, t)
    assert math.isclose(l, 0.25)
    assert math.isclose(g[ref.B3], 1.0)
    assert np.count_nonzero(g) == 1


def test_fold_standardisation():
    rng = np.random.default_rng(5)
    p = _params(2)
    x = rng.normal(3.0, 2.0, size=(50, 12))
    mu, sd = x.mean(0), x.std(0)
    tmu, tsd = 7.0, 3.0
    want = ref.forward(p, (x - mu) / sd) * tsd + tmu
    got = ref.forward(costmodel.fold(p, mu, sd, tmu, tsd).astype(np.float64), x)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)


def test_feature_restatement_closed_forms():
    T, C, A, G, N = 1, 2, 4, 8, 15
    # 16 nt: GCGC AAAAAA N TTTTT
    v = np.array([G, C, G, C, A, A, A, A, A, A, N, T, T, T, T, T], dtype=np.uint8)
    f = ref.chunk_features(v, [((0, G),)], {0b1010101010101010}, set())
    assert math.isclose(f[0], 4 / 15)            # G,C,G,C of 15 single bases
    assert math.isclose(f[1], 1 / 16)
    assert math.isclose(f[4], 1 / 16)            # one C->G step (positions 1-2)
    assert math.isclose(f[5], 1 / 16)            # AAAAAA starts once
    assert math.isclose(f[6], 2 / 16)            # two G cells
    assert math.isclose(f[11], 16 / 65536)
    # 8-mers: starts 0..2 are valid (0..7, 1..8, 2..9); 3.. hit the N
    assert math.isclose(f[2], math.log2(3) / 16)
    assert f[3] == 0.0
    # AAAAAAAA = code 2 repeated -> 0xAAAA
    v2 = np.full(20, A, dtype=np.uint8)
    f2 = ref.chunk_features(v2, [], {0xAAAA}, {0xAAAA})
    assert f2[2] == 0.0 and f2[3] == 1.0
    assert math.isclose(f2[7], 13 / 20) and math.isclose(f2[8], 13 / 20)
    assert math.isclose(f2[9], math.log2(14) / 16)


def test_query_inputs_pams_and_seeds():
    pams, any_set, pam_set = ref.query_inputs(["ACGTACGTACNNN"], ["NNNNNNNNNNNGG"], [True])
    # forward NGG cells (rel 0: G, rel 1: G); reverse CCN -> cells C, C
    assert pams == [((0, 8), (1, 8)), ((0, 2), (1, 2))]
    # forward seed: the 8 nt before the trailing N run: positions 2..9 = GTACGTAC
    code = {"T": 0, "C": 1, "A": 2, "G": 3}
    km = 0
    for c in "GTACGTAC":
        km = (km << 2) | code[c]
    assert km in pam_set and km in any_set
