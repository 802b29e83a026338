"""Seeded synthetic single-cell Hi-C contact maps for the decomposition benchmarks.

Model (one rule for every workload):

* every cell type t and chromosome c own a +-1 compartment vector u_{t,c}; it is
  piecewise constant over blocks whose lengths are drawn uniformly from
  [BLOCK_MIN, BLOCK_MAX] bins, the sign of the first block is random and the sign
  alternates from block to block;
* a cell of type t draws its total contact count L ~ round(lognormal(ln mu, sigma));
* the L contacts are spread multinomially over every intra-chromosome bin pair
  (a <= b, b - a <= band) of every chromosome, with weight
  (1 + |a - b|)^-1.1 * (1 + 0.3 * u_a * u_b), normalised over all pairs of all
  chromosomes;
* the result is an int32 COO table (cell, chrom, a, b, count).

The table enters the packing step the way per-cell contact matrices do in
Fast-Higashi: csr(a, b) + csr(a, b)^T, so the diagonal is doubled
(Fast_process.py:135-136).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence

import numpy as np

BLOCK_MIN, BLOCK_MAX = 4, 24


@dataclasses.dataclass
class Workload:
    name: str
    n_cells: int
    n_types: int
    bins: List[int]
    resolution: int
    band: int
    mu: float
    sigma: float
    rank: int
    n_iter: int
    off_diag: int
    do_rwr: bool = True
    do_col: bool = False
    dim1: float = 0.6
    seed: int = 0

    @property
    def chrom_list(self):
        return ["chr%d" % (i + 1) for i in range(len(self.bins))]


def _linspace_bins(hi, lo, k):
    return [int(v) for v in np.rint(np.linspace(hi, lo, k))]


# The five workloads of the decomposition benchmarks.  c1 is the small case the
# CPU restatement runs in full; c2 is the throughput workload.
WORKLOADS = {
    "c1": Workload("c1", 256, 4, [240, 200, 160], 500_000, 100, 2e4, 0.5,
                   rank=64, n_iter=30, off_diag=100, seed=1),
    "c2": Workload("c2", 10_000, 8, _linspace_bins(500, 95, 22), 500_000, 200, 5e4, 0.6,
                   rank=256, n_iter=50, off_diag=200, seed=2),
    "c3": Workload("c3", 5_000, 8, _linspace_bins(2490, 467, 23), 100_000, 100, 5e4, 0.6,
                   rank=256, n_iter=50, off_diag=100, do_rwr=False, seed=3),
    "c4": Workload("c4", 2_000, 8, [500, 480, 400, 380, 360], 500_000, 200, 5e4, 0.6,
                   rank=128, n_iter=50, off_diag=200, do_rwr=False, seed=4),
    "c5": Workload("c5", 20_000, 12, _linspace_bins(500, 95, 22), 500_000, 200, 5e4, 0.6,
                   rank=256, n_iter=50, off_diag=200, do_rwr=False, seed=5),
}


def scaled(wl: Workload, n_cells=None, bins=None, rank=None, name=None, **kw) -> Workload:
    """A copy of a workload with fewer cells / other bin counts (parity-test sizes)."""
    d = dataclasses.asdict(wl)
    if n_cells is not None:
        d["n_cells"] = n_cells
    if bins is not None:
        d["bins"] = list(bins)
    if rank is not None:
        d["rank"] = rank
    d["name"] = name or (wl.name + "_small")
    d.update(kw)
    return Workload(**d)


def compartments(rng: np.random.Generator, n_bins: int) -> np.ndarray:
    u = np.empty(n_bins, dtype=np.float64)
    sign = 1.0 if rng.random() < 0.5 else -1.0
    pos = 0
    while pos < n_bins:
        length = int(rng.integers(BLOCK_MIN, BLOCK_MAX + 1))
        u[pos:pos + length] = sign
        sign = -sign
        pos += length
    return u


def pair_table(n_bins: int, band: int):
    """All (a, b) with a <= b < n_bins and b - a <= band, ordered by a then b."""
    a_list, b_list = [], []
    for d in range(0, min(band, n_bins - 1) + 1):
        a = np.arange(0, n_bins - d, dtype=np.int32)
        a_list.append(a)
        b_list.append(a + d)
    a = np.concatenate(a_list)
    b = np.concatenate(b_list)
    order = np.lexsort((b, a))
    return a[order], b[order]


def generate(wl: Workload):
    """Return the COO contact table of a workload.

    Output: dict(cell, chrom, a, b, count) of int32 arrays, sorted by
    (cell, chrom, a, b), plus 'cell_type' (n_cells,) and 'bins'.
    """
    rng = np.random.default_rng(wl.seed)
    n_chrom = len(wl.bins)
    pairs = [pair_table(n, wl.band) for n in wl.bins]
    pair_off = np.concatenate([[0], np.cumsum([len(p[0]) for p in pairs])])
    base_w = [np.power(1.0 + (b - a), -1.1) for a, b in pairs]
    comp = [[compartments(rng, n) for n in wl.bins] for _ in range(wl.n_types)]
    cdfs = []
    for t in range(wl.n_types):
        w = np.concatenate([bw * (1.0 + 0.3 * comp[t][c][pairs[c][0]] * comp[t][c][pairs[c][1]])
                            for c, bw in enumerate(base_w)])
        cdf = np.cumsum(w)
        cdfs.append(cdf / cdf[-1])
    cell_type = np.arange(wl.n_cells) % wl.n_types
    rng.shuffle(cell_type)
    totals = np.rint(rng.lognormal(math.log(wl.mu), wl.sigma, size=wl.n_cells)).astype(np.int64)
    totals = np.maximum(totals, 1)
    out_cell, out_pair, out_cnt = [], [], []
    for i in range(wl.n_cells):
        u = rng.random(int(totals[i]))
        idx = np.searchsorted(cdfs[cell_type[i]], u, side="right")
        idx = np.minimum(idx, pair_off[-1] - 1)
        uniq, cnt = np.unique(idx, return_counts=True)
        out_cell.append(np.full(len(uniq), i, dtype=np.int32))
        out_pair.append(uniq.astype(np.int64))
        out_cnt.append(cnt.astype(np.int32))
    cell = np.concatenate(out_cell)
    pidx = np.concatenate(out_pair)
    cnt = np.concatenate(out_cnt)
    chrom = (np.searchsorted(pair_off, pidx, side="right") - 1).astype(np.int32)
    local = pidx - pair_off[chrom]
    a = np.empty(len(pidx), dtype=np.int32)
    b = np.empty(len(pidx), dtype=np.int32)
    for c in range(n_chrom):
        m = chrom == c
        a[m] = pairs[c][0][local[m]]
        b[m] = pairs[c][1][local[m]]
    return dict(cell=cell, chrom=chrom, a=a, b=b, count=cnt,
                cell_type=cell_type.astype(np.int32), bins=list(wl.bins))


def per_cell_csr(table, chrom_index: int, n_cells: int):
    """Per-cell symmetric float32 CSR matrices of one chromosome (csr + csr.T)."""
    from scipy.sparse import csr_matrix
    n = table["bins"][chrom_index]
    m = table["chrom"] == chrom_index
    cell, a, b, cnt = table["cell"][m], table["a"][m], table["b"][m], table["count"][m]
    bounds = np.searchsorted(cell, np.arange(n_cells + 1))
    out = []
    for i in range(n_cells):
        s = slice(bounds[i], bounds[i + 1])
        m1 = csr_matrix((cnt[s].astype(np.float32), (a[s], b[s])), shape=(n, n), dtype="float32")
        out.append(m1 + m1.T)
    return out
