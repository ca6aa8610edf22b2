"""Seeded synthetic workloads of the benchmark configurations.

Each generator returns numpy arrays only (no GPU work):
  train (rows, cols, vals), test (rows, cols, vals), shape, and a dict of the
  run settings of that configuration.

* ``c1`` -- 5,000 x 2,000 BPMF, 5% of cells observed uniformly at random,
  rank-10 truth with N(0,1) factors plus N(0, 0.5^2) noise, 20% held out;
  K = 10, fixed precision 5, burnin 20, nsamples 50.
* ``c2`` -- 138,000 x 27,000 recommender-scale BPMF with 20,000,000 observed
  cells.  Row and column degrees follow a rank-frequency power law
  d_r = clip(C r^-1.1, 20, cap) with C solved so the degrees sum to the
  target (row cap = min(30,000, number of columns)); cells are drawn by
  degree-weighted column sampling without duplicates.  Values come from a
  rank-32 truth with N(0,1) factors, standardised over the observed cells to
  mean 3.5 and sd 1.0, rounded to 0.5 steps and clipped to [0.5, 5]; 10% held
  out; K = 32, adaptive noise precision, burnin 50, nsamples 100.
* ``c3`` -- Macau with sparse binary row side information: 15,000 x 350
  with 60,000 observed cells (column degrees rank-frequency power law,
  exponent 1.3; rows uniform), values from a rank-32 truth rescaled to mean 6,
  sd 1.5 plus N(0, 0.6^2) noise; row features 15,000 x 100,000 binary with
  ~100 distinct features per row drawn with probability ~ rank^-1.5; 10% held
  out; K = 32, lambda_beta 5, CG tolerance 1e-6, fixed precision 5.0 (noise
  model and chain length are not fixed by the configuration), burnin 20,
  nsamples 50.
* ``c4`` -- 3-mode tensor 50,000 x 2,000 x 100 with 10,000,000 observed
  cells, each mode's index drawn from a rank-frequency power law with exponent
  1.2; values from a CP rank-64 truth with factor entries of variance 1/8
  (= 1/sqrt(64)) plus N(0, 0.3^2) noise; 10% held out; K = 64, fixed
  precision 1/0.09, burnin 30, nsamples 50.

* ``c5`` -- 64 independent BPMF instances of uneven size: row and column
  counts log-uniform in [1,000, 50,000] x [500, 10,000], density log-uniform
  in [0.1%, 5%] (cells uniformly at random), num_latent drawn from
  {8, 16, 32, 64, 128}; values u.v + N(0, 0.5^2) from a truth of the
  instance's rank whose factor entries have variance K^-1/2 (so u.v has unit
  variance at every rank); 10% held out; fixed precision 4 (= 1 / 0.5^2),
  burnin 20, nsamples 50 (chain lengths are not fixed by the configuration).
  ``scale`` < 1 shrinks every instance's row and column counts (for tests).

``c3`` and ``c4`` return the generic form used for N-mode data:
  (train_idx [nnz x nmodes] int32, train_vals), (test_idx, test_vals), dims,
  settings, side (one scipy COO matrix or None per mode).
"""
import numpy as np


def zipf_degrees(n, total, exponent, lo, hi):
    """d_r = clip(C r^-exponent, lo, hi), r = 1..n, with sum(d) ~= total."""
    r = np.arange(1, n + 1, dtype=np.float64)
    base = r ** (-exponent)

    def tot(C):
        return np.clip(C * base, lo, hi).sum()

    a, b = 0.0, float(total) * 10
    for _ in range(200):
        m = 0.5 * (a + b)
        if tot(m) < total:
            a = m
        else:
            b = m
    d = np.clip(np.round(b * base), lo, hi).astype(np.int64)
    return d


def _split(rng, r, c, v, frac):
    perm = rng.permutation(len(v))
    nte = int(round(len(v) * frac))
    te, tr = perm[:nte], perm[nte:]
    return (r[tr], c[tr], v[tr]), (r[te], c[te], v[te])


def c1(seed=0):
    rng = np.random.default_rng(seed)
    nrow, ncol, K = 5000, 2000, 10
    ncell = int(0.05 * nrow * ncol)
    cells = rng.choice(nrow * ncol, ncell, replace=False)
    r, c = (cells // ncol).astype(np.int32), (cells % ncol).astype(np.int32)
    U, V = rng.standard_normal((nrow, K)), rng.standard_normal((ncol, K))
    y = np.einsum("ij,ij->i", U[r], V[c]) + 0.5 * rng.standard_normal(ncell)
    train, test = _split(rng, r, c, y, 0.2)
    settings = dict(num_latent=10, precision=5.0, burnin=20, nsamples=50)
    return train, test, (nrow, ncol), settings


def _degree_sampled_cells(rng, deg_r, weight_c, ncol):
    """Unique (row, col) cells with row degrees ~deg_r and columns drawn with
    probability proportional to weight_c."""
    nrow = len(deg_r)
    p = weight_c / weight_c.sum()
    cdf = np.cumsum(p)
    cdf[-1] = 1.0
    need = deg_r.copy()
    keys = np.zeros(0, dtype=np.int64)
    for _ in range(8):
        if need.sum() == 0:
            break
        rows = np.repeat(np.arange(nrow, dtype=np.int64), need)
        cols = np.searchsorted(cdf, rng.random(len(rows)))
        keys = np.unique(np.concatenate([keys, rows * ncol + cols]))
        have = np.bincount((keys // ncol).astype(np.int64), minlength=nrow)
        need = np.maximum(deg_r - have, 0)
        # saturated rows take the remaining columns uniformly
        need = np.minimum(need, ncol - have)
    return keys


def c2(seed=0, nrow=138000, ncol=27000, nnz=20_000_000, K=32):
    rng = np.random.default_rng(seed)
    deg_r = zipf_degrees(nrow, nnz, 1.1, 20, min(30000, ncol))
    deg_c = zipf_degrees(ncol, nnz, 1.1, 20, min(30000, nrow))
    deg_r = deg_r[rng.permutation(nrow)]
    w_c = deg_c[rng.permutation(ncol)].astype(np.float64)
    keys = _degree_sampled_cells(rng, deg_r, w_c, ncol)
    if len(keys) > nnz:
        keys = np.sort(rng.choice(keys, nnz, replace=False))
    elif len(keys) < nnz:
        # top up with uniformly random new cells
        extra = np.unique(rng.integers(0, nrow * ncol, 2 * (nnz - len(keys)), dtype=np.int64))
        extra = np.setdiff1d(extra, keys)[: nnz - len(keys)]
        keys = np.sort(np.concatenate([keys, extra]))
    r = (keys // ncol).astype(np.int32)
    c = (keys % ncol).astype(np.int32)
    U = rng.standard_normal((nrow, K)).astype(np.float32)
    V = rng.standard_normal((ncol, K)).astype(np.float32)
    y = np.empty(len(r), dtype=np.float64)
    step = 1 << 20
    for s in range(0, len(r), step):
        y[s:s + step] = np.einsum("ij,ij->i", U[r[s:s + step]], V[c[s:s + step]])
    y = (y - y.mean()) / y.std() + 3.5
    y = np.clip(np.round(y * 2.0) / 2.0, 0.5, 5.0)
    train, test = _split(rng, r, c, y, 0.1)
    settings = dict(num_latent=K, precision="adaptive", burnin=50, nsamples=100)
    return train, test, (nrow, ncol), settings


def c3(seed=0, nrow=15000, ncol=350, ncell=60000, nfeat=100000, feat_per_row=100, K=32):
    import scipy.sparse
    rng = np.random.default_rng(seed)
    deg_c = zipf_degrees(ncol, ncell, 1.3, 1, nrow).astype(np.float64)
    keys = np.zeros(0, dtype=np.int64)
    cdf = np.cumsum(deg_c / deg_c.sum())
    cdf[-1] = 1.0
    while len(keys) < ncell:
        m = 2 * (ncell - len(keys))
        cols = np.searchsorted(cdf, rng.random(m))
        rows = rng.integers(0, nrow, m)
        keys = np.unique(np.concatenate([keys, rows.astype(np.int64) * ncol + cols]))
    keys = np.sort(rng.choice(keys, ncell, replace=False))
    r, c = (keys // ncol).astype(np.int32), (keys % ncol).astype(np.int32)
    U, V = rng.standard_normal((nrow, K)), rng.standard_normal((ncol, K))
    y = np.einsum("ij,ij->i", U[r], V[c])
    y = (y - y.mean()) / y.std() * 1.5 + 6.0 + 0.6 * rng.standard_normal(ncell)
    # binary features: ~feat_per_row distinct ids per row, P(id = k) ~ (k+1)^-1.5
    w = np.arange(1, nfeat + 1, dtype=np.float64) ** -1.5
    fcdf = np.cumsum(w / w.sum())
    fcdf[-1] = 1.0
    draws = np.searchsorted(fcdf, rng.random((nrow, 12 * feat_per_row))).astype(np.int64)
    fk = np.unique(np.arange(nrow, dtype=np.int64)[:, None] * nfeat + draws)
    fr, fc = (fk // nfeat).astype(np.int32), (fk % nfeat).astype(np.int32)
    # keep at most feat_per_row per row, in id order
    first = np.searchsorted(fr, np.arange(nrow))
    keep = (np.arange(len(fr)) - first[fr]) < feat_per_row
    F = scipy.sparse.coo_matrix((np.ones(keep.sum()), (fr[keep], fc[keep])), shape=(nrow, nfeat))
    idx = np.stack([r, c], 1)
    perm = rng.permutation(ncell)
    nte = ncell // 10
    te, tr = perm[:nte], perm[nte:]
    settings = dict(num_latent=K, precision=5.0, burnin=20, nsamples=50, lambda_beta=5.0, tol=1e-6)
    return (idx[tr], y[tr]), (idx[te], y[te]), (nrow, ncol), settings, [F, None]


def c4(seed=0, dims=(50000, 2000, 100), ncell=10_000_000, K=64, R=64):
    rng = np.random.default_rng(seed)
    cdfs = []
    for d in dims:
        w = zipf_degrees(d, ncell, 1.2, 1, ncell).astype(np.float64)[rng.permutation(d)]
        cdf = np.cumsum(w / w.sum())
        cdf[-1] = 1.0
        cdfs.append(cdf)
    mult = np.array([dims[1] * dims[2], dims[2], 1], dtype=np.int64)
    keys = np.zeros(0, dtype=np.int64)
    while len(keys) < ncell:
        m = int(1.1 * (ncell - len(keys))) + 1000
        k = np.zeros(m, dtype=np.int64)
        for q in range(3):
            k += np.searchsorted(cdfs[q], rng.random(m)).astype(np.int64) * mult[q]
        keys = np.unique(np.concatenate([keys, k]))
    keys = np.sort(rng.choice(keys, ncell, replace=False))
    idx = np.stack([keys // mult[0], (keys // mult[1]) % dims[1], keys % dims[2]], 1).astype(np.int32)
    sd = 64 ** -0.25  # entries of variance 1/sqrt(64)
    Fs = [(rng.standard_normal((d, R)) * sd).astype(np.float32) for d in dims]
    y = np.empty(ncell, dtype=np.float64)
    step = 1 << 20
    for s in range(0, ncell, step):
        i = idx[s:s + step]
        y[s:s + step] = (Fs[0][i[:, 0]] * Fs[1][i[:, 1]] * Fs[2][i[:, 2]]).sum(1)
    y += 0.3 * rng.standard_normal(ncell)
    perm = rng.permutation(ncell)
    nte = ncell // 10
    te, tr = perm[:nte], perm[nte:]
    settings = dict(num_latent=K, precision=1.0 / 0.09, burnin=30, nsamples=50)
    return (idx[tr], y[tr]), (idx[te], y[te]), tuple(dims), settings, [None, None, None]


def _uniform_cells(rng, nrow, ncol, ncell):
    """ncell distinct cells drawn uniformly from an nrow x ncol matrix."""
    total = nrow * ncol
    ncell = min(ncell, total)
    keys = np.zeros(0, dtype=np.int64)
    while len(keys) < ncell:
        m = int(1.2 * (ncell - len(keys))) + 16
        keys = np.unique(np.concatenate([keys, rng.integers(0, total, m, dtype=np.int64)]))
    if len(keys) > ncell:
        keys = np.sort(rng.choice(keys, ncell, replace=False))
    return (keys // ncol).astype(np.int32), (keys % ncol).astype(np.int32)


def c5(seed=0, n_instances=64, scale=1.0):
    """List of instances, each (train, test, shape, settings) like ``c1``."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n_instances):
        nrow = int(round(np.exp(rng.uniform(np.log(1000), np.log(50000))) * scale))
        ncol = int(round(np.exp(rng.uniform(np.log(500), np.log(10000))) * scale))
        nrow, ncol = max(nrow, 8), max(ncol, 8)
        dens = np.exp(rng.uniform(np.log(0.001), np.log(0.05)))
        K = int(rng.choice([8, 16, 32, 64, 128]))
        ncell = max(int(round(nrow * ncol * dens)), 20)
        r, c = _uniform_cells(rng, nrow, ncol, ncell)
        sd = K ** -0.25
        U = (rng.standard_normal((nrow, K)) * sd).astype(np.float32)
        V = (rng.standard_normal((ncol, K)) * sd).astype(np.float32)
        y = np.einsum("ij,ij->i", U[r], V[c]).astype(np.float64) + 0.5 * rng.standard_normal(len(r))
        train, test = _split(rng, r, c, y, 0.1)
        settings = dict(num_latent=K, precision=4.0, burnin=20, nsamples=50, density=float(dens))
        out.append((train, test, (nrow, ncol), settings))
    return out
