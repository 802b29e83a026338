"""Packed scHi-C storage and band tiling (drop-in for sparse_for_schic.py).

* `Sparse` — n-d COO container with the reference's host API (permute, reshape,
  slicing, ...).  `sort_indices` (the CSR build) runs on the device: a stable counting
  sort by row whose output order is that of scipy's coo_tocsr (sparse_for_schic.py:133-182).
* `Chrom_Dataset` — bin tiles x cell chunks with the reference's slice lists
  (sparse_for_schic.py:356-510).  The packed COO lives in HBM; `fetch` scatters a tile
  into a dense device tensor with the 1e-8 floor of densify_jit (:279-320).
"""
from __future__ import annotations

import math
from typing import List, Optional

import numpy as np
import torch

from . import _lib as L
from . import ops
from .engine import TileSpec
from .packing import DeviceCOO, TileBuilder, coo_to_device, csr_sort_device
from .tiling import tile_plan


def calc_offset(dims):
    """Row-major strides of a shape (sparse_for_schic.py:15-18)."""
    dims = np.asarray(dims)
    offset = np.ones_like(dims)
    offset[:-1] = np.cumprod(dims[:0:-1])[::-1]
    return offset


def unfold(indices, shape):
    """Multi-index -> flat index (sparse_for_schic.py:28-36)."""
    shape = np.asarray(shape)
    flat = calc_offset(shape) @ np.asarray(indices).astype(shape.dtype, copy=False)
    assert (flat >= 0).all() and (flat < np.prod(shape)).all()
    return flat


def fold(indices, shape, out=None):
    """Flat index -> multi-index (sparse_for_schic.py:39-54)."""
    shape = np.asarray(shape)
    off = calc_offset(shape)
    res = np.empty([len(shape), len(indices)], dtype=np.asarray(indices).dtype) if out is None else out
    rem = np.asarray(indices)
    for i in range(len(shape) - 1):
        res[i], rem = np.divmod(rem, off[i])
    res[-1] = rem
    return res


def _arr(x, copy):
    return np.array(x, copy=True) if copy else np.asarray(x)


class Sparse:
    """n-d COO tensor; rows (dim 0) can be CSR-indexed after sort_indices()."""

    def __init__(self, indices, values, shape, indptr=None, copy=True, verbose=False):
        self.ndim = len(shape)
        self.indices = _arr(indices, copy)
        self.values = _arr(values, copy)
        self.shape = _arr(shape, copy)
        self.indptr = None if indptr is None else _arr(indptr, copy)
        self.verbose = verbose
        assert self.indices.shape == (self.ndim, len(self.values))
        assert (self.indices >= 0).all()
        assert (self.indices < self.shape[:, None]).all(), (self.indices.max(1), self.shape)

    def sort_indices(self, dim=0, force=False):
        """Stable sort by row + indptr on the device (sparse_for_schic.py:133-182)."""
        assert dim == 0
        if not force and self.indptr is not None:
            return
        n_rows = int(self.shape[0])
        nnz = len(self.values)
        if nnz == 0:
            self.indptr = np.zeros(n_rows + 1, dtype=np.int64)
            return
        if self.ndim != 3:
            # permuted / reshaped views: sort the (row, flattened rest) pairs as a 3-d problem
            rest = unfold(self.indices[1:], self.shape[1:])
            idx3 = np.stack([self.indices[0], rest, np.zeros_like(rest)])
        else:
            idx3 = self.indices
        if idx3.max() >= 2 ** 31:
            raise ValueError("sort_indices: indices exceed int32")
        i = torch.as_tensor(np.ascontiguousarray(idx3, dtype=np.int32), device="cuda")
        v = torch.zeros(nnz, dtype=torch.float32, device="cuda")
        _, _, ptr, perm = csr_sort_device(i, v, n_rows, want_perm=True)
        perm = perm.cpu().numpy()
        self.indices = np.ascontiguousarray(self.indices[:, perm])
        self.values = np.ascontiguousarray(self.values[perm])
        self.indptr = ptr.cpu().numpy()

    def get_slice_idx_value(self, idx, dim=0, device="cpu"):
        """Entries of rows [start, stop) as views, with the slice shape and start (:200-213)."""
        assert dim == 0 and isinstance(idx, slice) and (idx.step is None or idx.step == 1)
        if self.indptr is None:
            self.sort_indices(dim)
        start = idx.start if idx.start is not None else 0
        stop = min(idx.stop if idx.stop is not None else self.shape[dim], self.shape[dim])
        lo, hi = self.indptr[start], self.indptr[stop]
        return self.indices[:, lo:hi], self.values[lo:hi], (stop - start,) + tuple(self.shape[1:]), start

    def slicing(self, idx, dim=0):
        indices, values, shape, start = self.get_slice_idx_value(idx, dim)
        indices = indices.copy()
        indices[dim] -= start
        return Sparse(indices, values, shape)

    def indexing(self, idx, dim=0):
        assert dim == 0 and isinstance(idx, int) and 0 <= idx < self.shape[dim]
        if self.indptr is None:
            self.sort_indices(dim)
        sl = slice(self.indptr[idx], self.indptr[idx + 1])
        return Sparse(self.indices[1:, sl], self.values[sl], tuple(self.shape[1:]))

    def __getitem__(self, item):
        if isinstance(item, int):
            return self.indexing(item)
        if isinstance(item, slice):
            return self.slicing(item)
        raise NotImplementedError

    def permute(self, *dims, inplace=False):
        assert tuple(sorted(dims)) == tuple(range(self.ndim))
        dims = np.array(dims)
        indices, shape = self.indices[dims], self.shape[dims]
        indptr = self.indptr if dims[0] == 0 else None
        if inplace:
            self.indices, self.shape, self.indptr = indices, shape, indptr
            return self
        return Sparse(indices, self.values, shape, indptr, verbose=self.verbose)

    def reshape(self, *dims, inplace=False):
        dims = np.array(dims)
        if (dims == -1).any():
            assert (dims == -1).sum() <= 1
            dims[dims == -1] = self.shape.prod() // -dims.prod()
        assert self.shape.prod() == dims.prod()
        indices = fold(unfold(self.indices, self.shape), dims)
        if inplace:
            self.indices, self.shape, self.ndim = indices, dims, len(dims)
            return self
        return Sparse(indices, self.values, dims, verbose=self.verbose,
                      indptr=self.indptr if dims[0] == self.shape[0] else None)

    def filter_max_distance(self, max_distance=100):
        keep = np.abs(self.indices[1] - self.indices[0]) <= max_distance
        self.indices = np.asarray([a[keep] for a in self.indices])
        self.values = self.values[keep]
        self.indptr = None

    def to_dense(self):
        v = np.zeros(tuple(self.shape))
        v[tuple(self.indices)] = self.values
        return v

    def to_scipy(self):
        from scipy.sparse import coo_matrix
        return coo_matrix((self.values, tuple(self.indices)), tuple(self.shape))

    def to_csr(self):
        from scipy.sparse import csr_matrix
        return csr_matrix((self.values, tuple(self.indices)), tuple(self.shape))

    def to_pytorch(self, **context):
        return torch.sparse_coo_tensor(self.indices, self.values, self.shape.tolist(), **context)

    def __len__(self):
        return int(self.shape[0])

    def numel(self):
        return int(np.prod(self.shape))


def densify_jit(shape, indices, values, device, transpose: bool = False, do_conv: bool = False):
    """Dense tile of a +1-padded tile COO (sparse_for_schic.py:279-320): scatter, crop the
    padding, clamp to >= 1e-8.  shape = (rows + 2, w + 2, cells)."""
    if do_conv:
        raise NotImplementedError("densify_jit: the sparse 3x3 convolution path is unused by the reference "
                                  "(sparse_conv is False) and not implemented")
    rows, w, nc = int(shape[0]) - 2, int(shape[1]) - 2, int(shape[2])
    r = torch.as_tensor(indices[0], device=device).to(torch.int32) - 1
    c = torch.as_tensor(indices[1], device=device).to(torch.int32) - 1
    k = torch.as_tensor(indices[2], device=device).to(torch.int32)
    v = torch.as_tensor(values, device=device).to(torch.float32).contiguous()
    idx = torch.stack([r, c, k]).contiguous()
    wp = ops.rup(w, 8)
    Xc = torch.empty(nc, rows, wp, dtype=torch.float32, device=device)
    ops.fill_(Xc, 1e-8)
    bt = L.BandTile()
    bt.offset, bt.s_row, bt.s_col, bt.s_cell = 0, wp, 1, rows * wp
    bt.row_start, bt.rows, bt.shift, bt.width = 0, rows, 0, w
    tab = ops._upload((L.BandTile * 1)(bt), torch.device(device))
    tor = torch.zeros(max(rows, 1), dtype=torch.int32, device=device)
    if v.numel():
        L.call("fh_band_scatter", idx.data_ptr(), v.data_ptr(), v.numel(), tor.data_ptr(), tab.data_ptr(), 0, nc,
               Xc.data_ptr(), L.F32, None, ops.stream_handle())
    out = Xc[:, :, :w]
    return out if transpose else out.permute(1, 2, 0)


class Fake_Sparse:
    """+1-padded COO of one (bin tile, cell chunk) (sparse_for_schic.py:322-353), built on
    demand from the dataset's packed COO."""

    def __init__(self, dataset: "Chrom_Dataset", tile: TileSpec, cells: slice):
        self._ds, self.tile, self.cells = dataset, tile, cells
        self.slice_ = slice(tile.start, tile.start + tile.rows)
        self.shape = [tile.rows + 2, tile.width + 2, cells.stop - cells.start]

    @property
    def indices(self):
        idx, _ = self._ds._tile_coo(self.tile, self.cells)
        return [torch.as_tensor(idx[0] + 1), torch.as_tensor(idx[1] + 1), torch.as_tensor(idx[2])]

    @property
    def values(self):
        return torch.as_tensor(self._ds._tile_coo(self.tile, self.cells)[1])

    def pin_memory(self):
        return self

    def densify(self, save_context, transpose=False, do_conv=False, out=None):
        return self._ds._dense(self.tile, self.cells, transpose)


class Chrom_Dataset:
    """Bin tiles x cell chunks of one chromosome; the packed COO stays in HBM."""

    def __init__(self, tensor, bs_bin, bs_cell, good_qc_num=-1, kind="hic", upper_sim=False, compact=False,
                 flank=0, chrom="chr1", resolution=10000):
        if upper_sim:
            raise NotImplementedError("Chrom_Dataset: upper_sim=True is not used by FastHigashi.prep_dataset and "
                                      "is not implemented")
        if isinstance(tensor, torch.Tensor):
            raise NotImplementedError("Chrom_Dataset: dense tensor input is not implemented; pass a Sparse")
        self.resolution, self.chrom = resolution, chrom
        self.length = self.num_bin = int(tensor.shape[0])
        if good_qc_num == -1:
            good_qc_num = int(tensor.shape[-1])
        self.num_cell = int(good_qc_num)
        self.total_cell_num = int(tensor.shape[-1])
        self.bs_cell, self.bs_bin = int(bs_cell), int(bs_bin)
        self.upper_sim, self.compact, self.flank = upper_sim, compact, flank
        if compact:
            self.tiles = tile_plan(self.num_bin, self.bs_bin, flank)
        else:
            self.tiles = [TileSpec(i, min(i + self.bs_bin, self.num_bin) - i, self.num_bin, 0, i, 0)
                          for i in range(0, self.num_bin, self.bs_bin)]
        self.bin_slice_list = [slice(t.start, t.start + t.rows) for t in self.tiles]
        self.local_bin_slice_list = [slice(t.local_start, t.local_start + t.rows) for t in self.tiles]
        self.col_bin_slice_list = ([slice(t.col_start, t.col_start + t.width) for t in self.tiles] if compact
                                   else [slice(None) for _ in self.tiles])
        starts = list(range(0, self.num_cell, self.bs_cell)) + list(range(self.num_cell, self.total_cell_num,
                                                                         self.bs_cell))
        self.cell_slice_list = [slice(s, min(s + self.bs_cell, self.num_cell if s < self.num_cell
                                             else self.total_cell_num)) for s in starts]
        good = [c for c in self.cell_slice_list if c.start < self.num_cell]
        bad = [c for c in self.cell_slice_list if c.start >= self.num_cell]
        self.tensor_list = [[Fake_Sparse(self, t, c) for c in good] for t in self.tiles]
        self.bad_tensor_list = [[Fake_Sparse(self, t, c) for c in bad] for t in self.tiles]
        self.kind_list = [[kind] * len(good) for _ in self.tiles]
        self.bad_kind_list = [[kind] * len(bad) for _ in self.tiles]
        self.num_bin_batch = len(self.tensor_list)
        self.num_cell_batch = len(good)
        self.num_cell_batch_bad = len(bad)
        self.uniq_kind = np.unique(self.kind_list)
        self.shape = [self.num_bin, (self.bs_bin + 2 * flank) if compact else tensor.shape[1], self.num_cell]
        # packed COO (sorted by row) on the device, plus the host copy for Fake_Sparse views
        tensor.sort_indices()
        self._host = tensor
        self.coo: DeviceCOO = coo_to_device(tensor.indices, tensor.values, tensor.shape)
        self._builder = TileBuilder(self.coo)

    def __len__(self):
        return self.length

    def _tile_coo(self, tile, cells):
        idx, val, _, start = self._host.get_slice_idx_value(slice(tile.start, tile.start + tile.rows))
        m = (idx[2] >= cells.start) & (idx[2] < cells.stop)
        local = np.stack([idx[0][m] - start, idx[1][m] - tile.shift, idx[2][m] - cells.start])
        return local, val[m]

    def _dense(self, tile, cells, transpose):
        Xc = self._builder.dense(tile, cells)[:, :, :tile.width]
        return Xc if transpose else Xc.permute(1, 2, 0)

    def pin_memory(self):
        return self

    def fetch(self, bin_id, cell_id, save_context=None, transpose=False, good_qc=True, **kwargs):
        """((dense tile, [seconds]), kind): (cells, rows, w) if transpose else (rows, w, cells)."""
        import time
        t0 = time.perf_counter()
        if good_qc:
            fs, kind = self.tensor_list[bin_id][cell_id], self.kind_list[bin_id][cell_id]
        else:
            fs, kind = self.bad_tensor_list[bin_id][cell_id], self.bad_kind_list[bin_id][cell_id]
        a = fs.densify(save_context, transpose if kind == "hic" else False)
        return (a, [time.perf_counter() - t0]), kind

    def fetch_bad(self, bin_id, cell_id, **kwargs):
        return self.fetch(bin_id, cell_id, good_qc=False, **kwargs)

    def norm(self):
        idx = self._host.indices
        v = self._host.values[idx[2] < self.num_cell].astype(np.float64)
        return float(np.sqrt(np.square(v).sum()))

    def tile_specs(self) -> List[TileSpec]:
        return list(self.tiles)


def test():
    """The reference's Sparse self-test (sparse_for_schic.py:634-661)."""
    def new_obj():
        return Sparse([[0, 2, 1, 0], [0, 3, 2, 1]], [1, 2, 3, 4.], (3, 4))
    base = new_obj().to_scipy().todense()
    o = new_obj()
    o.sort_indices()
    assert (o.to_scipy().todense() == base).all()
    assert (new_obj().permute(0, 1).to_scipy().todense() == base).all()
    assert (new_obj().permute(1, 0).to_scipy().todense() == base.T).all()
    assert (new_obj().reshape(1, 12).to_scipy().todense() == base.reshape(1, 12)).all()
    assert (new_obj().reshape(4, 3).to_scipy().todense() == base.reshape(4, 3)).all()
    for s in [slice(2), slice(10), slice(0, None), slice(2, None), slice(1, 3)]:
        assert (new_obj().slicing(s).to_scipy().todense() == base[s]).all()
    o = new_obj().permute(1, 0).reshape(6, 2).slicing(slice(1, 3)).permute(1, 0)
    assert (o.to_scipy().todense() == base.T.reshape(6, 2)[1:3].T).all()
