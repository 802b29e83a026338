"""GPU packing: CSR sort, band scatter, partial RWR and the X layout against the oracle."""
import numpy as np
import pytest
import torch

from fasthigashi import synthetic as syn
from oracle import ref_np as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def packed():
    wl = syn.scaled(syn.WORKLOADS["c1"], n_cells=96, bins=[110, 70], band=30, off_diag=30, name="gp")
    tab = syn.generate(wl)
    out = []
    for c, n in enumerate(wl.bins):
        mats = syn.per_cell_csr(tab, c, wl.n_cells)
        idx, val, shape, _ = R.pack_chrom(mats, wl.off_diag)
        sidx, sval, indptr = R.sort_indices(idx, val, shape)
        out.append((idx, val, shape, sidx, sval, indptr))
    return wl, out


def test_device_coo_sorted_like_oracle(cuda, packed):
    from fasthigashi.packing import coo_to_device
    wl, chroms = packed
    for idx, val, shape, sidx, sval, indptr in chroms:
        d = coo_to_device(idx, val, shape, cuda)
        assert np.array_equal(d.idx.cpu().numpy(), sidx)
        assert np.array_equal(d.val.cpu().numpy(), sval)
        assert np.array_equal(d.indptr_host, indptr)


@pytest.mark.parametrize("bs_bin", [40, 1000])
def test_dense_and_rwr_match_oracle(cuda, packed, bs_bin):
    from fasthigashi.packing import TileBuilder, coo_to_device
    wl, chroms = packed
    for idx, val, shape, sidx, sval, indptr in chroms:
        d = coo_to_device(idx, val, shape, cuda)
        tb = TileBuilder(d, cuda)
        n = shape[0]
        tiles = R.tile_plan(n, min(bs_bin, n), wl.off_diag)
        for tl in tiles:
            cells = slice(10, 80)
            coo, v, ps = R.tile_coo(sidx, sval, indptr, shape, tl, cells)
            ref = R.densify(coo, v, ps, transpose=True)
            Xc = tb.dense(tl, cells)
            got = Xc[:, :, :tl.width].cpu().numpy()
            assert np.array_equal(got, ref)
            assert (Xc[:, :, tl.width:] == 0).all()
            for epochs in (3, -1):
                rx, rn = R.partial_rwr(ref.copy(), tl.local_start, tl.local_start + tl.rows, False, True, False,
                                       force_rwr_epochs=epochs, final_transpose=False)
                out, cnt = tb.rwr(tb.dense(tl, cells), tl, epochs)
                assert cnt == rn
                g = out[:, :, :tl.width].cpu().numpy()
                assert np.abs(g - rx).max() <= 2e-5 * np.abs(rx).max()


def test_build_x_layout(cuda, packed):
    from fasthigashi.engine import SweepEngine, TileSpec
    from fasthigashi.packing import build_X, coo_to_device
    wl, chroms = packed
    tiles, refs, coos, sizes = [], [], [], []
    for idx, val, shape, sidx, sval, indptr in chroms:
        n = shape[0]
        tl_list = R.tile_plan(n, 40, wl.off_diag)
        tiles.append([TileSpec(t.start, t.rows, t.width, t.col_start, t.local_start, t.shift) for t in tl_list])
        refs.append([R.partial_rwr(R.densify(*R.tile_coo(sidx, sval, indptr, shape, t, slice(0, shape[2]))[:3],
                                              transpose=True), t.local_start, t.local_start + t.rows, False, True,
                                   False, force_rwr_epochs=4)[0] for t in tl_list])
        coos.append(coo_to_device(idx, val, shape, cuda))
        sizes.append(8)
    eng = SweepEngine(tiles, [c[2][0] for c in chroms], sizes, wl.n_cells, 16, "fp32")
    build_X(eng, coos, [4, 4], do_rwr=True, budget_bytes=1 << 20)
    for c in range(len(chroms)):
        for t, ref in enumerate(refs[c]):
            got = eng.X_tile(c, t)[:, :, :wl.n_cells].cpu().numpy()
            assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()
            assert (eng.X_tile(c, t)[:, :, wl.n_cells:] == 0).all()
    tn = eng.tnorm.cpu().numpy()
    ref_tn = [sum(float(np.square(x.astype(np.float64)).sum()) for x in refs[c]) for c in range(len(chroms))]
    assert np.allclose(tn, ref_tn, rtol=1e-5)
