"""Polar factor of batched matrices (drop-in for project2orthogonal.py:6-55).

The reference computes U[..., :rank] Vh[..., :rank, :] from an FP32 SVD.  Here the
polar factor comes from the FP64 Gram G = Y^T Y and G^{-1/2} by a Newton-Schulz
iteration on the matrix cores (libfh_hip), P = Y G^{-1/2}; for full-column-rank Y this
is the same matrix.  Singular values are the square roots of the eigenvalues of G.
"""
from __future__ import annotations

import torch

from . import ops


def project2orthogonal(matrix: torch.Tensor, rank: int = None, compute_device=None):
    """Returns (polar factor with the shape of `matrix`, singular values (..., rank)).

    `rank` must be None or min(m, k): a truncated product U_r Vh_r of a rank-deficient
    factorisation has no Newton-Schulz counterpart and raises instead of falling back."""
    m, k = matrix.shape[-2], matrix.shape[-1]
    full = min(m, k)
    if rank is not None and rank != full:
        raise NotImplementedError("project2orthogonal: only the full polar factor (rank = min(m, k)) is "
                                  "implemented; got rank=%d for a %dx%d matrix" % (rank, m, k))
    dev = matrix.device if matrix.is_cuda else torch.device("cuda")
    batch_shape = matrix.shape[:-2]
    Y = matrix.to(dev, torch.float32).reshape(-1, m, k)
    wide = k > m
    if wide:
        Y = Y.transpose(1, 2)
    Y = Y.contiguous()
    P, nuc, G = ops.polar(Y, return_gram=True)
    # singular values (descending) from the FP64 Gram, for API parity only
    ev = torch.linalg.eigvalsh(G[:, :, :G.shape[1]])
    S = torch.sqrt(torch.clamp(ev, min=0)).flip(-1).to(torch.float32)
    if wide:
        P = P.transpose(1, 2)
    out_dev = compute_device if compute_device is not None else matrix.device
    return P.reshape(*batch_shape, m, k).to(out_dev), S.reshape(*batch_shape, full).to(out_dev)
