"""Host packing (product) against the oracle's per-cell scipy restatement: bit-exact."""
import numpy as np
import pytest

from fasthigashi import synthetic as syn
from fasthigashi.packing import pack_chrom
from oracle import ref_np as R


@pytest.mark.parametrize("off_diag", [30, 5])
def test_pack_chrom_bit_exact(off_diag):
    wl = syn.scaled(syn.WORKLOADS["c1"], n_cells=48, bins=[90, 60], band=30, off_diag=off_diag, name="pk")
    tab = syn.generate(wl)
    for c, n in enumerate(wl.bins):
        mats = syn.per_cell_csr(tab, c, wl.n_cells)
        ridx, rval, rshape, rmap = R.pack_chrom(mats, wl.off_diag)
        m = tab["chrom"] == c
        idx, val, shape, mapping = pack_chrom(tab["cell"][m], tab["a"][m], tab["b"][m], tab["count"][m], n,
                                              wl.n_cells, wl.off_diag)
        assert shape == rshape
        assert np.array_equal(mapping, rmap)
        assert np.array_equal(idx, ridx)
        assert np.array_equal(val.view(np.uint32), rval.view(np.uint32))


def test_pack_chrom_drops_empty_bins():
    # bins 3, 6 and 7 never have a contact: they are remapped away and later bins shift down
    cell = np.array([0, 0, 1, 1, 1])
    a = np.array([0, 1, 2, 4, 8])
    b = np.array([1, 2, 4, 5, 9])
    cnt = np.array([1, 2, 3, 4, 5])
    idx, val, shape, mapping = pack_chrom(cell, a, b, cnt, 10, 2, 5)
    assert shape == (7, 7, 2)
    assert list(mapping) == [0, 1, 2, -1, 3, 4, -1, -1, 5, 6]
    assert idx.min() >= 0 and (idx.max(1) < np.array(shape)).all()
