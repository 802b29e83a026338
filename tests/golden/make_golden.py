"""Writes the golden fixtures of the CPU oracle tests (run once; outputs are committed).

* sparse_kat.json — the reference's only unit test, `sparse_for_schic.test()`
  (sparse_for_schic.py:634-661): the input COO and, for every operation it checks,
  the dense matrix scipy produces (scipy is the independent library the reference
  itself compares against).
* pfc_tiling.json — known answers printed in the reference notebook
  (PFC tutorial.ipynb:1559-1581): per-chromosome bin counts of the PFC dataset at
  500 kb with off_diag = 100 and the compact tile width bs_bin + 2 * off_diag.
"""
import json
import os

import numpy as np
from scipy.sparse import coo_matrix

HERE = os.path.dirname(os.path.abspath(__file__))


def sparse_kat():
    rows, cols, vals, shape = [0, 2, 1, 0], [0, 3, 2, 1], [1.0, 2.0, 3.0, 4.0], (3, 4)
    base = np.asarray(coo_matrix((vals, (rows, cols)), shape=shape).todense())
    cases = {
        "sort_indices": base,
        "permute_0_1": base,
        "permute_1_0": base.T,
        "reshape_1_12": base.reshape(1, 12),
        "reshape_4_3": base.reshape(4, 3),
        "chain": base.T.reshape(6, 2)[1:3].T,
    }
    slices = {"slice_:2": base[:2], "slice_:10": base[:10], "slice_0:": base[0:], "slice_2:": base[2:],
              "slice_1:3": base[1:3]}
    cases.update(slices)
    return dict(indices=[rows, cols], values=vals, shape=list(shape),
                expected={k: np.asarray(v).tolist() for k, v in cases.items()})


def pfc_tiling():
    # PFC tutorial.ipynb:1560-1581 (bins per chromosome, then "shape" lines of each Chrom_Dataset)
    bins = [457, 481, 391, 378, 357, 338, 314, 288, 248, 267, 265, 263, 193, 177, 166, 160, 158, 151, 114, 121,
            74, 71]
    widths = [315, 321, 298, 326, 319, 313, 305, 296, 324, 289, 289, 288, 297, 289, 283, 280, 279, 276, 314, 321,
              274, 271]
    # same cell: 4238 cells, recommend_bs_cell = 4238 for every chromosome, sparsity and the do_col decision
    return dict(resolution=500000, off_diag=100, bins=bins, widths=widths, num_cell=4238, bs_cell=4238,
                sparsity=0.05111497517410393, do_col=False)


if __name__ == "__main__":
    with open(os.path.join(HERE, "sparse_kat.json"), "w") as f:
        json.dump(sparse_kat(), f, indent=1)
    with open(os.path.join(HERE, "pfc_tiling.json"), "w") as f:
        json.dump(pfc_tiling(), f, indent=1)
