"""Partial random walk with restart on a dense band tile (drop-in for partial_rwr.py:45-175).

x: (cells, rows, w) on the device.  Same arguments and return values as the reference;
the work runs in libfh_hip (3x3 average pool, A A^T, transition matrix, power iteration
Q <- 0.5 Q P + 0.5 I on the matrix cores, optional do_col, Q A).
"""
from __future__ import annotations

import torch

from . import _lib as L
from . import ops
from .ops import rup
from .packing import rwr_padded


@torch.no_grad()
def partial_rwr(x: torch.Tensor, slice_start: int, slice_end: int, do_conv: bool, do_rwr: bool, do_col: bool,
                bin_cov: torch.Tensor = torch.ones(1), bin_cov_row: torch.Tensor = torch.ones(1),
                return_rwr_iter: bool = False, force_rwr_epochs: int = -1, final_transpose: bool = True, **kw):
    n_iter = 0
    if do_conv or do_rwr:
        dev = x.device if x.is_cuda else torch.device("cuda")
        nc, rows, w = x.shape
        wp = rup(w, 8)
        Xc = torch.zeros(nc, rows, wp, dtype=torch.float32, device=dev)
        Xc[:, :, :w] = x.to(dev, torch.float32)
        if do_conv and rows > 1:                                              # :77-81
            pooled = torch.zeros_like(Xc)
            L.call("fh_avgpool3", Xc.data_ptr(), wp, rows * wp, rows, w, nc, pooled.data_ptr(), wp, rows * wp,
                   ops.stream_handle())
            Xc = pooled
        if do_rwr:
            if slice_end - slice_start != rows:
                raise ValueError("partial_rwr: the diagonal block [%d, %d) must span the %d tile rows"
                                 % (slice_start, slice_end, rows))
            cov = None
            if do_col:
                cov = bin_cov.to(dev, torch.float32).reshape(nc, w).contiguous()
            out, n_iter = rwr_padded(Xc, w, slice_start, force_rwr_epochs, do_col, cov)
        else:
            out = Xc
        x = out[:, :, :w]
        if final_transpose:                                                  # :171-172
            x = x.permute(1, 2, 0)
    if return_rwr_iter:
        return x, n_iter
    return x, 0
