"""Pins the oracle's packing / sorting / tiling functions (CPU).

Checks: the reference's own Sparse known-answer test (golden/sparse_kat.json), the
notebook's tiling known answers (golden/pfc_tiling.json), scipy's coo_tocsr (the
routine the reference calls, sparse_for_schic.py:159-162) and dense brute force."""
import json
import math
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st
from scipy.sparse import coo_matrix
from scipy.sparse._sparsetools import coo_tocsr

from oracle import ref_np as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _dense2(idx, val, shape):
    return np.asarray(coo_matrix((val, (idx[0], idx[1])), shape=tuple(shape)).todense())


def test_sparse_kat_oracle():
    kat = _load("sparse_kat.json")
    idx = np.array(kat["indices"], dtype=np.int64)
    val = np.array(kat["values"], dtype=np.float32)
    shape = tuple(kat["shape"])
    sidx, sval, indptr = R.sort_indices(idx, val, shape)
    assert np.array_equal(_dense2(sidx, sval, shape), np.array(kat["expected"]["sort_indices"]))
    assert list(sidx[0]) == sorted(sidx[0])
    for key, s in [("slice_:2", slice(0, 2)), ("slice_:10", slice(0, 10)), ("slice_0:", slice(0, 3)),
                   ("slice_2:", slice(2, 3)), ("slice_1:3", slice(1, 3))]:
        i2, v2, sh = R.slice_rows(sidx, sval, indptr, shape, s.start, s.stop)
        i2 = i2.copy()
        i2[0] -= s.start
        assert np.array_equal(_dense2(i2, v2, sh), np.array(kat["expected"][key])), key


def test_sparse_kat_product_host_ops():
    """The product's host-side Sparse reshaping (no device work) on the same KAT."""
    from fasthigashi.sparse_for_schic import Sparse
    kat = _load("sparse_kat.json")

    def new():
        return Sparse(kat["indices"], kat["values"], tuple(kat["shape"]))

    exp = {k: np.array(v) for k, v in kat["expected"].items()}
    assert np.array_equal(np.asarray(new().permute(0, 1).to_scipy().todense()), exp["permute_0_1"])
    assert np.array_equal(np.asarray(new().permute(1, 0).to_scipy().todense()), exp["permute_1_0"])
    assert np.array_equal(np.asarray(new().reshape(1, 12).to_scipy().todense()), exp["reshape_1_12"])
    assert np.array_equal(np.asarray(new().reshape(4, 3).to_scipy().todense()), exp["reshape_4_3"])


@settings(max_examples=60, deadline=None)
@given(st.integers(1, 12), st.integers(1, 6), st.integers(1, 5), st.integers(0, 80), st.integers(0, 2 ** 31 - 1),
       st.booleans())
def test_sort_indices_matches_coo_tocsr(n_rows, n_cols, n_cells, nnz, seed, presorted):
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.integers(0, n_rows, nnz), rng.integers(0, n_cols, nnz),
                    rng.integers(0, n_cells, nnz)]).astype(np.int64)
    if presorted:
        idx = idx[:, np.argsort(idx[0], kind="stable")]
    val = rng.random(nnz).astype(np.float32)
    shape = (n_rows, n_cols, n_cells)
    sidx, sval, indptr = R.sort_indices(idx, val, shape)
    # scipy's coo_tocsr on (row, unfolded rest), as sparse_for_schic.py:153-182
    key = idx[1] * n_cells + idx[2]
    ip = np.empty(n_rows + 1, dtype=np.int64)
    cols = np.empty(nnz, dtype=np.int64)
    vals = np.empty(nnz, dtype=np.float32)
    coo_tocsr(n_rows, int(key.max()) + 1 if nnz else 1, nnz, idx[0], key, val, ip, cols, vals)
    assert np.array_equal(indptr, ip)
    assert np.array_equal(sidx[0], np.repeat(np.arange(n_rows), np.diff(ip)))
    assert np.array_equal(sidx[1] * n_cells + sidx[2], cols)
    assert np.array_equal(sval, vals)


def test_pfc_tiling_known_answers():
    from fasthigashi import tiling
    kat = _load("pfc_tiling.json")
    res, off = kat["resolution"], kat["off_diag"]
    assert tiling.compact_widths(kat["bins"], res, off) == kat["widths"]
    big = 288e9      # ample device memory: one cell chunk, as printed in the notebook
    for n, w in zip(kat["bins"], kat["widths"]):
        bs_bin, bs_cell = R.gpu_tiling(n, res, off, kat["num_cell"], big)
        assert bs_bin + 2 * off == w
        assert bs_cell == kat["bs_cell"]
        tiles = R.tile_plan(n, bs_bin, off)
        # every tile is at most w wide and the tiles cover the bins once
        assert max(t.width for t in tiles) <= w
        assert sum(t.rows for t in tiles) == n
    assert (kat["sparsity"] * (500000 / res) ** 2 <= 0.03) == kat["do_col"]


@pytest.mark.parametrize("n,bs,flank", [(40, 7, 5), (30, 30, 8), (25, 4, 0), (17, 5, 20)])
def test_tile_plan_brute_force(n, bs, flank):
    """Compact band tiles against their definition: row block [i, i+bs), columns within flank of it."""
    tiles = R.tile_plan(n, bs, flank)
    assert [t.start for t in tiles] == list(range(0, n, bs))
    for t in tiles:
        lo = max(t.start - flank, 0)
        hi = min(t.start + t.rows + flank, n)
        assert (t.col_start, t.width) == (lo, hi - lo)
        assert t.shift == t.col_start
        assert t.local_start == t.start - t.col_start
    cs = R.cell_slices(10, 13, 4)
    assert cs == [slice(0, 4), slice(4, 8), slice(8, 10), slice(10, 13)]


def test_tile_coo_and_densify_brute_force():
    rng = np.random.default_rng(4)
    n, N, flank = 23, 7, 4
    dense = np.zeros((n, n, N), np.float32)
    for _ in range(120):
        i, j, k = rng.integers(0, n), rng.integers(0, n), rng.integers(0, N)
        if abs(i - j) <= flank:
            dense[i, j, k] = rng.random() + 0.5
    idx = np.array(np.nonzero(dense))
    val = dense[tuple(idx)]
    perm = rng.permutation(idx.shape[1])
    sidx, sval, indptr = R.sort_indices(idx[:, perm], val[perm], dense.shape)
    for t in R.tile_plan(n, 6, flank):
        for cells in (slice(0, 3), slice(3, 7)):
            coo, v, ps = R.tile_coo(sidx, sval, indptr, dense.shape, t, cells)
            got = R.densify(coo, v, ps, transpose=False, dt=np.float32)
            want = np.maximum(dense[t.start:t.start + t.rows, t.col_start:t.col_start + t.width, cells],
                              np.float32(1e-8))
            assert np.array_equal(got, want)
            got_t = R.densify(coo, v, ps, transpose=True, dt=np.float32)
            assert np.array_equal(got_t, want.transpose(2, 0, 1))


def test_pack_chrom_dense_brute_force():
    """Oracle packing against a dense per-cell restatement of the same steps on a tiny input."""
    from scipy.sparse import csr_matrix
    rng = np.random.default_rng(7)
    n, N, off = 12, 5, 3
    mats, dens = [], []
    for c in range(N):
        d = np.zeros((n, n), np.float32)
        for _ in range(15):
            a, b = sorted(rng.integers(0, n - 3, 2))      # the last 3 bins never get contacts
            d[a, b] += 1
        d = d + d.T
        mats.append(csr_matrix(d))
        dens.append(d)
    idx, val, shape, mapping = R.pack_chrom(mats, off)
    band = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]) <= off
    bulk = sum(d * band for d in dens)
    valid = bulk.sum(1) > 0
    assert shape == (int(valid.sum()), int(valid.sum()), N)
    assert list(np.nonzero(mapping >= 0)[0]) == list(np.nonzero(valid)[0])
    raw = []
    for c, d in enumerate(dens):
        db = (d * band).astype(np.float64)
        db = db * n / (db.sum() + 1e-15)
        sub = db[np.ix_(valid, valid)]
        r, cc = np.nonzero(sub)
        raw.extend((c, i, j, sub[i, j]) for i, j in zip(r, cc))
    got = sorted(zip(idx[2], idx[0], idx[1], val))
    want_vals = np.log1p(np.array([x[3] for x in raw]))
    cap = want_vals.mean() + 15 * want_vals.std()
    want = sorted((c, i, j, min(v, cap)) for (c, i, j, _), v in zip(raw, want_vals))
    assert [g[:3] for g in got] == [w[:3] for w in want]
    assert np.allclose([g[3] for g in got], [w[3] for w in want], rtol=1e-6)
    assert not np.isnan(val).any()                       # FastHigashi_Wrapper.py:299
    assert idx.min() >= 0 and (idx.max(1) < np.array(shape)).all()   # :353-354
