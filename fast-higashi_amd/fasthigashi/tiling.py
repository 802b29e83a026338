"""Bin tiles x cell chunks of a chromosome (Chrom_Dataset slicing and prep_dataset sizes).

Restates sparse_for_schic.py:401-510 (compact=True) and FastHigashi_Wrapper.py:495-519.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .engine import TileSpec


def recommend_bs_bin(resolution: int) -> int:
    """FastHigashi_Wrapper.py:501."""
    return min(max(int(15000000 / resolution), 128), 256)


def bs_bin_for(num_bin: int, resolution: int, on_device: bool = True) -> int:
    """Bin tile height: the whole chromosome on the CPU path, else ceil(n / ceil(n / rec))
    (FastHigashi_Wrapper.py:507-512)."""
    if not on_device:
        return num_bin
    rec = recommend_bs_bin(resolution)
    n_batch = max(math.ceil(num_bin / rec), 1)
    return math.ceil(num_bin / n_batch)


def bs_cell_for(bs_bin: int, off_diag: int, num_cell: int, avail_mem: float) -> int:
    """Cell chunk of the reference's device branch (FastHigashi_Wrapper.py:497-517)."""
    max_tensor_size = avail_mem / (4 * 12)
    bs_cell = int(max_tensor_size / (bs_bin * (bs_bin + 2 * off_diag)))
    bs_cell = max(bs_cell, 1)
    n_batch = int(math.ceil(num_cell / bs_cell))
    bs_cell = int(math.ceil(num_cell / n_batch))
    return min(bs_cell, num_cell)


def tile_plan(num_bin: int, bs_bin: int, flank: int) -> List[TileSpec]:
    """Compact band tiles (sparse_for_schic.py:457-492, 509-510): rows [i, i+bs), columns
    [i - left, i + rows + right) with left = flank if i > flank else i."""
    tiles = []
    for i in range(0, num_bin, bs_bin):
        rows = min(i + bs_bin, num_bin) - i
        right = flank if num_bin - i - rows - flank > 0 else num_bin - i - rows
        if i > flank:
            left, local_start, col_start, shift = flank, flank, i - flank, i - flank
        else:
            left, local_start, col_start, shift = i, i, 0, 0
        tiles.append(TileSpec(i, rows, rows + left + right, col_start, local_start, shift))
    return tiles


def compact_widths(bins, resolution: int, off_diag: int) -> List[int]:
    """bs_bin + 2 * flank per chromosome, the Chrom_Dataset.shape[1] of each dataset."""
    return [bs_bin_for(n, resolution) + 2 * off_diag for n in bins]


def size_of(num_bin: int, size_ratio: float, resolution: int, rank: int) -> int:
    """Per-chromosome CP rank (parafac2_intergrative.py:558-560)."""
    return min(int(num_bin * size_ratio * resolution / 1000000), rank)


def structured_random_init(n_bins, sizes, n_cells: int, rank: int, seed: int = 0):
    """Seeded factors with the reference's init structure (parafac2_intergrative.py:71-80,
    283-290): A ~ 1 + 1e-2 randn, B = I + 1e-2 randn(s) added to every row, and (M, D) from
    the SVD of a cell x sum(s) matrix (here Gaussian instead of the RWR-feature SVD)."""
    rng = np.random.default_rng(seed)
    A = [(rng.standard_normal((n, s)) * 1e-2 + 1).astype(np.float32) for n, s in zip(n_bins, sizes)]
    B = [(np.eye(s) + 1e-2 * rng.standard_normal(s)[None, :]).astype(np.float32) for s in sizes]
    C = rng.standard_normal((n_cells, sum(sizes))).astype(np.float32)
    U, S, Vh = np.linalg.svd(C, full_matrices=False)
    M = U[:, :rank].astype(np.float32)
    SVh = (S[:rank, None] * Vh[:rank]).astype(np.float32)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    D = [SVh[:, offs[i]:offs[i + 1]].copy() for i in range(len(sizes))]
    return A, B, D, M
