"""Dense 3-way CP-ALS (drop-in for parafac_integrative.py:12-112).

parafac(X, rank, n_iter_max, init) fits X (I x rank x K) ~ [[A, B, D]] with the
reference's schedule: balance_norm, then per iteration the A, B, D updates
F = MTTKRP (Hadamard of Grams + 1e-10 I)^{-1}, the relative-loss stop (< 1e-5) and a
balance_norm when not stopped.  Runs on libfh_hip (grouped GEMMs, FP64 Gram inverses).
"""
from __future__ import annotations

import torch

from .engine import CpAls
from .ops import rup


def update_factor(f, A, B):
    """f[:] = solve(A + 1e-10 I, B)^T (parafac_integrative.py:12-15); A (r x r), B (r x n)."""
    A = A.double().clone()
    A.diagonal().add_(1e-10)
    f[:] = torch.linalg.solve(A, B.double()).T.to(f.dtype)


def balance_norm(factors):
    """Columns of every factor to unit norm, the norm product moved to the last factor
    (parafac_integrative.py:19-26)."""
    norm = 1
    for f in factors:
        norm = norm * torch.norm(f, dim=0)
    even = norm + 1e-15
    for i in range(len(factors)):
        factors[i] = factors[i] / (torch.norm(factors[i], dim=0) + 1e-15)
    factors[-1] = factors[-1] * even


@torch.no_grad()
def parafac(X, rank, n_iter_max=100, init=None, verbose=False, common_factor=None):
    """Returns (factors [A, B, D], ||Xhat||^2, <Xhat, X>) like the reference."""
    if init is None:
        raise ValueError("parafac: init factors are required (the reference has no default init either)")
    I, J, K = X.shape
    if J != rank:
        raise NotImplementedError("parafac: the second mode must have size rank (got %d vs %d)" % (J, rank))
    dev = torch.device("cuda")
    s, sp, Kp = rank, rup(rank, 8), rup(K, 8)
    T = torch.zeros(I, sp, Kp, dtype=torch.float32, device=dev)
    T[:, :s, :K] = torch.as_tensor(X).to(dev, torch.float32)
    A = torch.zeros(I, sp, dtype=torch.float32, device=dev)
    B = torch.zeros(s, sp, dtype=torch.float32, device=dev)
    D = torch.zeros(Kp, sp, dtype=torch.float32, device=dev)
    A[:, :s] = torch.as_tensor(init[0]).to(dev, torch.float32)
    B[:, :s] = torch.as_tensor(init[1]).to(dev, torch.float32)
    D[:K, :s] = torch.as_tensor(init[2]).to(dev, torch.float32)
    cp = CpAls([dict(T=T.data_ptr(), A=A.data_ptr(), B=B.data_ptr(), D=D.data_ptr(), n=I, s=s, sp=sp, Dld=sp)],
               Kp, dev)
    cp.run(int(n_iter_max))
    _, inner, xhat2 = cp.loss_terms()[0]
    return [A[:, :s].clone(), B[:, :s].clone(), D[:K, :s].clone()], float(xhat2), float(inner)
