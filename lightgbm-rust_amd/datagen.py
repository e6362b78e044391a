"""Deterministic synthetic inputs of the benchmark workloads.

One generator drives every workload: splitmix64 streams -> uniforms in [0, 1)
(53-bit) -> Box-Muller normals. Seeds per workload c: X from 1000 + c, the
weight vector from 2000 + c, labels from 3000 + c.

Workloads (rows x features):
  binary_100k   : 100,000 x 28 N(0,1); y ~ Bernoulli(sigmoid(w.x + 0.5 x0 x1))
  binary_10p5m  : 10,500,000 x 28 (20 N(0,1) + 8 lognormal(0,1)); same label
                  model, then 5% of labels flipped
  multiclass_2m : 2,000,000 x 100, 10 classes: 80 N(0,1) columns, 12 columns
                  of 2-16 levels, 8 categorical columns of 50 Zipf(1.2)
                  categories; y = argmax_c(x.W_c + cat embeddings + Gumbel)
  rank_2m       : ~20,000 queries (~2M documents) x 136, graded relevance 0-4
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, start, n):
    """Outputs start..start+n of the splitmix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniforms(seed, start, n):
    return (splitmix64(seed, start, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normals(seed, start, n):
    """Box-Muller on consecutive uniform pairs; output k uses pair k."""
    u = uniforms(seed, 2 * start, 2 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def _chunks(n, chunk, fn):
    """fn(r0, r1) over row chunks on a few threads (numpy releases the GIL in
    its array loops); every chunk depends only on its own stream offsets, so
    the result does not depend on the thread count."""
    ranges = [(r0, min(n, r0 + chunk)) for r0 in range(0, n, chunk)]
    workers = max(1, min(16, os.cpu_count() or 1, len(ranges)))
    if workers == 1:
        for r in ranges:
            fn(*r)
        return
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda r: fn(*r), ranges))


def _features(c, n, F, lognormal_cols=(), chunk=1 << 18):
    X = np.empty((n, F), np.float32)

    def fill(r0, r1):
        z = normals(1000 + c, r0 * F, (r1 - r0) * F).reshape(r1 - r0, F)
        for j in lognormal_cols:
            z[:, j] = np.exp(z[:, j])
        X[r0:r1] = z.astype(np.float32)

    _chunks(n, chunk, fill)
    return X


def _binary_labels(c, X, flip_rate=0.0, chunk=1 << 18):
    n, F = X.shape
    w = normals(2000 + c, 0, F)
    y = np.empty(n, np.float32)

    def fill(r0, r1):
        x = X[r0:r1].astype(np.float64)
        z = x @ w + 0.5 * x[:, 0] * x[:, 1]
        p = 1.0 / (1.0 + np.exp(-z))
        u = uniforms(3000 + c, r0, r1 - r0)
        lab = (u < p)
        if flip_rate > 0:
            f = uniforms(4000 + c, r0, r1 - r0) < flip_rate
            lab = lab ^ f
        y[r0:r1] = lab.astype(np.float32)

    _chunks(n, chunk, fill)
    return y


def binary_100k(n=100_000):
    X = _features(1, n, 28)
    return X, _binary_labels(1, X)


def binary_10p5m(n=10_500_000):
    X = _features(2, n, 28, lognormal_cols=range(20, 28))
    return X, _binary_labels(2, X, flip_rate=0.05)


def multiclass_2m(n=2_000_000, K=10, chunk=1 << 16):
    """The batched multiclass workload. Columns 80-91 hold floor(U k_j) with
    k_j = 2 + (j mod 15); columns 92-99 are categories 0-49 drawn from
    Zipf(1.2) by inverse CDF (pass them as categorical_feature). Labels are
    argmax_c(x.W_c + E_j[cat_j]_c + Gumbel) with W ~ N(0,1)[92 x K] and one
    N(0,1)[50 x K] embedding per categorical column."""
    c, F, nnum, nlev, ncat, ncats = 4, 100, 80, 12, 8, 50
    X = np.empty((n, F), np.float32)
    pz = 1.0 / np.arange(1, ncats + 1) ** 1.2
    cdf = np.cumsum(pz / pz.sum())
    W = normals(2000 + c, 0, (nnum + nlev) * K).reshape(nnum + nlev, K)
    E = normals(2100 + c, 0, ncat * ncats * K).reshape(ncat, ncats, K)
    y = np.empty(n, np.float32)

    def fill(r0, r1):
        m = r1 - r0
        X[r0:r1, :nnum] = normals(1000 + c, r0 * nnum, m * nnum).reshape(m, nnum)
        u = uniforms(1100 + c, r0 * nlev, m * nlev).reshape(m, nlev)
        lev = 2 + (np.arange(nlev) % 15)
        X[r0:r1, nnum:nnum + nlev] = np.floor(u * lev)
        u = uniforms(1200 + c, r0 * ncat, m * ncat).reshape(m, ncat)
        cats = np.minimum(np.searchsorted(cdf, u, side="right"), ncats - 1)
        X[r0:r1, nnum + nlev:] = cats
        z = X[r0:r1, :nnum + nlev].astype(np.float64) @ W
        for j in range(ncat):
            z += E[j][cats[:, j]]
        g = -np.log(-np.log(np.maximum(uniforms(3000 + c, r0 * K, m * K).reshape(m, K), 1e-300)))
        y[r0:r1] = np.argmax(z + g, axis=1).astype(np.float32)

    _chunks(n, chunk, fill)
    return X, y


MULTICLASS_CATEGORICAL = list(range(92, 100))


def rank_2m(n_queries=20_000, max_size=2000, F=136):
    """Lambdarank workload: queries of Geometric(p=1/100) size capped at 2000,
    half N(0,1) features, half N(0,1) x Bernoulli(0.3) (70% zeros); graded
    relevance 0-4 = clip(floor(5 Phi(z)), 0, 4) of a noisy linear score."""
    c = 5
    u = uniforms(5000 + c, 0, n_queries)
    sizes = np.minimum(np.floor(np.log1p(-u) / np.log1p(-0.01)).astype(np.int64) + 1, max_size)
    qb = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    n = int(qb[-1])
    X = _features(c, n, F)
    half = F // 2
    keep = uniforms(6000 + c, 0, n * (F - half)).reshape(n, F - half) < 0.3
    X[:, half:] = np.where(keep, X[:, half:], 0.0).astype(np.float32)
    v = normals(2000 + c, 0, F)
    z = X.astype(np.float64) @ v + 0.5 * normals(3000 + c, 0, n) * np.linalg.norm(v)
    z = z / np.sqrt(np.linalg.norm(v) ** 2 * 1.25)
    from math import erf, sqrt
    phi = 0.5 * (1.0 + np.vectorize(erf)(z / sqrt(2.0))) if n <= 200_000 else _phi(z)
    rel = np.clip(np.floor(5.0 * phi), 0, 4).astype(np.float32)
    return X, rel, qb


def _phi(z):
    # Abramowitz-Stegun 7.1.26 erf, vectorised (|error| < 1.5e-7)
    t = 1.0 / (1.0 + 0.3275911 * np.abs(z) / np.sqrt(2.0))
    y = 1.0 - (((((1.061405429 * t - 1.453152027) * t) + 1.421413741) * t - 0.284496736) * t + 0.254829592) * t * \
        np.exp(-(z * z) / 2.0)
    return 0.5 * (1.0 + np.sign(z) * y)


def holdout(workload, n):
    """A held-out sample of the same generator (different row range)."""
    if workload == "binary_10p5m":
        F, ln, flip, c = 28, range(20, 28), 0.05, 2
    else:
        F, ln, flip, c = 28, (), 0.0, 1
    X = np.empty((n, F), np.float32)
    off = 1 << 40
    z = normals(1000 + c, off, n * F).reshape(n, F)
    for j in ln:
        z[:, j] = np.exp(z[:, j])
    X[:] = z.astype(np.float32)
    w = normals(2000 + c, 0, F)
    x = X.astype(np.float64)
    p = 1.0 / (1.0 + np.exp(-(x @ w + 0.5 * x[:, 0] * x[:, 1])))
    lab = uniforms(3000 + c, off, n) < p
    if flip:
        lab = lab ^ (uniforms(4000 + c, off, n) < flip)
    return X, lab.astype(np.float32)
