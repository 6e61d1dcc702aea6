"""Training driver of the bulge scheduler's cost model (host logic over the C ABI).

Rows are genome chunks: features from ``cof_chunk_features`` (computed on the
device), target log(1 + traceback DFS transitions of the chunk's filter tiles)
from ``cof_get_tile_stats`` of a search run with ``COF_FLAG_TILE_STATS``.
Training runs on the device (``cof_mlp_train``: full-batch Adam, one step
captured in a hipGraph).  Inputs and target are standardised for training;
the scaling is folded into the first and last layer afterwards, so the model
handed to ``cof_set_cost_model`` consumes raw features and predicts raw
log-work.
"""
import numpy as np

import cofpy

IN, H = cofpy.MLP_IN, 64
W1, B1, W2, B2, W3, B3 = 0, 768, 832, 4928, 4992, 5056


def chunk_targets(search):
    """Per-chunk log(1 + DFS transitions) of the last search (needs COF_FLAG_TILE_STATS)."""
    ts = search.tile_stats()
    n = len(search.chunk_features())
    work = np.zeros(n, dtype=np.float64)
    np.add.at(work, ts["chunk"].astype(np.int64), ts["dfs_steps"].astype(np.float64))
    return np.log1p(work).astype(np.float32)


def collect(search):
    """(features, targets) rows of the search's genome and query; runs one search with tile stats."""
    search.set_flags(search.flags | cofpy.COF_FLAG_TILE_STATS)
    search.search()
    x = search.chunk_features()
    t = chunk_targets(search)
    search.set_flags(search.flags & ~cofpy.COF_FLAG_TILE_STATS)
    return x, t


def fold(params, mu, sd, tmu, tsd):
    """Parameters of the model on raw inputs/targets from one trained on standardised ones."""
    p = np.asarray(params, dtype=np.float64).copy()
    w1 = p[W1:B1].reshape(H, IN) / sd
    p[B1:W2] = p[B1:W2] - w1 @ mu
    p[W1:B1] = w1.ravel()
    p[W3:B3] *= tsd
    p[B3] = p[B3] * tsd + tmu
    return p.astype(np.float32)


def fit(x, t, seed=1, steps=3000, lr=3e-3, device=0):
    """Trains a fresh model on rows (x, t); returns (Mlp on raw features, final standardised loss)."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    mu, sd = x.mean(0), x.std(0)
    sd = np.where(sd > 1e-12, sd, 1.0)
    tmu, tsd = t.mean(), max(t.std(), 1e-12)
    m = cofpy.Mlp(seed=seed, device=device)
    loss = m.train(((x - mu) / sd).astype(np.float32), ((t - tmu) / tsd).astype(np.float32), steps, lr)
    m.set_params(fold(m.params(), mu, sd, tmu, tsd))
    return m, loss
