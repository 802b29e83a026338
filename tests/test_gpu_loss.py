"""Differentiable PARAFAC2 reconstruction loss (fh_loss_fwd / fh_loss_bwd behind a
torch.autograd.Function): value and gradient against the FP64 oracle at a small size, and at the
2,000-cell, 5-chromosome (500/480/400/380/360 bins), band-200, rank-128 workload against FP64
autograd on a plain-torch restatement of the loss and FP64 central differences (step 1e-6) on 64
random factor entries; then 200 Adam steps (lr 1e-3) from the ALS solution."""
import math
import sys
import time

import numpy as np
import pytest
import torch

from fasthigashi import synthetic as syn
from tests.helpers import build_dataset, random_init

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(data, F0, sizes, rank):
    from fasthigashi.engine import SweepEngine, TileSpec
    tiles = [[TileSpec(t.start, t.rows, t.width, t.col_start, t.local_start, t.shift) for t in cd.tiles]
             for cd in data]
    eng = SweepEngine(tiles, [cd.num_bin for cd in data], sizes, data[0].X[0].shape[2], rank, dtype="fp32")
    eng.set_X_tiles([cd.X for cd in data])
    eng.set_factors(F0.A, F0.B, F0.D, F0.M)
    return eng


def _unpack(eng, gA, gB, gD, gM):
    """packed device gradients -> per-chromosome numpy blocks (A n x s, B s x s, D R x s), M N x R"""
    A, B, D = [], [], []
    for cl in eng.chroms:
        A.append(gA[cl.aoff: cl.aoff + cl.n * cl.sp].view(cl.n, cl.sp)[:, :cl.s].double().cpu().numpy())
        B.append(gB[cl.boff: cl.boff + cl.s * cl.sp].view(cl.s, cl.sp)[:, :cl.s].double().cpu().numpy())
        D.append(gD[:, cl.doff: cl.doff + cl.s].double().cpu().numpy())
    return A, B, D, gM[:eng.N].double().cpu().numpy()


def _flat(blocks):
    return np.concatenate([x.ravel() for x in blocks]) if isinstance(blocks, list) else blocks.ravel()


@pytest.fixture(scope="module")
def small():
    wl = syn.scaled(syn.WORKLOADS["c1"], n_cells=160, bins=[120, 100, 90], rank=32, off_diag=30, band=30,
                    name="c1_tiny")
    data, _, _ = build_dataset(wl, bs_bin=50)
    F0, sizes = random_init(data, wl, wl.rank, seed=5)
    eng = _engine(data, F0, sizes, wl.rank)
    for _ in range(3):
        eng.sweep()
    torch.cuda.synchronize()
    return data, eng


@pytest.mark.parametrize("forward", ["rows", "gemm"])
def test_forward_kernels_match_oracle_small(cuda, small, forward):
    """Both forward kernels (row-stationary and GEMM-epilogue) against the FP64 oracle."""
    from oracle import ref_np as R
    from fasthigashi.loss import LossPlan
    data, eng = small
    plan = LossPlan(eng, forward=forward)
    plan.load(eng.A_all, eng.B_all, eng.D_all, eng.M)
    got = float(plan.forward().item())
    per = plan.loss_c.cpu().numpy()
    F = R.Factors(*eng.get_factors())
    proj = eng.projections()
    ref = R.parafac2_loss(data, F, proj)
    _say("small %s forward: %.9e oracle %.9e" % (forward, got, ref))
    assert abs(got - ref) < 1e-6 * ref
    assert abs(per.sum() - got) < 1e-12 * got
    for c in range(len(data)):
        sub = R.parafac2_loss([data[c]], R.Factors([F.A[c]], [F.B[c]], [F.D[c]], F.M), [proj[c]])
        assert abs(per[c] - sub) < 1e-6 * sub, c


@pytest.mark.parametrize("point", ["als", "perturbed"])
def test_loss_and_gradient_match_oracle_small(cuda, small, point):
    """At the ALS solution the A, B, D gradients nearly vanish (the CP-ALS step just solved for
    them), so their error is measured against the factor-Gram term they are the difference of
    (the gradient of sum ||Q C^T||^2, the loss at X = 0); at a perturbed point against the
    gradient itself."""
    from oracle import ref_np as R
    from fasthigashi.loss import LossPlan, parafac2_loss
    data, eng = small
    plan = LossPlan(eng)
    params = plan.params_from_engine()
    if point == "perturbed":
        g = torch.Generator(device=eng.device).manual_seed(11)
        with torch.no_grad():
            for p in params:
                p.mul_(1 + 0.05 * torch.randn(p.shape, generator=g, device=p.device))
    L = parafac2_loss(plan, *params)
    L.backward()
    torch.cuda.synchronize()
    A, B, D, M = _unpack(eng, *[p.detach() for p in params])
    F = R.Factors(A, B, D, M)
    proj = eng.projections()
    ref = R.parafac2_loss(data, F, proj)
    g_ref = R.parafac2_loss_grad(data, F, proj)
    zero = [R.ChromData(cd.chrom, cd.num_bin, cd.resolution, cd.tiles, [np.zeros_like(x) for x in cd.X],
                        cd.bs_bin, cd.cell_chunks) for cd in data]
    g_gram = R.parafac2_loss_grad(zero, F, proj)
    rel = abs(float(L.detach()) - ref) / ref
    _say("small %s: loss %.9e oracle %.9e rel %.2e" % (point, float(L.detach()), ref, rel))
    got = _unpack(eng, *[p.grad for p in params])
    errs = {}
    for name, g, r, q in zip("ABDM", got, (g_ref.A, g_ref.B, g_ref.D, g_ref.M),
                             (g_gram.A, g_gram.B, g_gram.D, g_gram.M)):
        g, r, q = _flat(g), _flat(r), _flat(q)
        d = np.linalg.norm(g - r)
        errs[name] = (d / np.linalg.norm(r), d / np.linalg.norm(q))
        _say("small %s: grad %s |g| %.3e |gram term| %.3e  err/|g| %.2e  err/|gram term| %.2e"
             % (point, name, np.linalg.norm(r), np.linalg.norm(q), *errs[name]))
    assert rel < 1e-6
    for name, (e_g, e_q) in errs.items():
        if point == "perturbed":
            assert e_g < 1e-4, (name, e_g)
        assert e_q < 1e-5, (name, e_q)


# ---------------------------------------------------------------------------- 2,000-cell workload

class TorchLoss:
    """Plain-torch FP64 restatement of the loss over the engine's tiles (the checker)."""

    def __init__(self, eng):
        self.eng = eng

    def factors(self, A, B, D, M, c):
        cl = self.eng.chroms[c]
        a = A[cl.aoff: cl.aoff + cl.n * cl.sp].view(cl.n, cl.sp)[:, :cl.s]
        b = B[cl.boff: cl.boff + cl.s * cl.sp].view(cl.s, cl.sp)[:, :cl.s]
        d = D[:, cl.doff: cl.doff + cl.s]
        return a, b, d

    def resid(self, A, B, D, M, c, t, bins=None, cells=None, with_x=True):
        """E = X - Q C^T of tile t of chromosome c (optionally some of its bins / cells)."""
        eng = self.eng
        cl = eng.chroms[c]
        tl = cl.tiles[t]
        a, b, d = self.factors(A, B, D, M, c)
        N = eng.N
        Mc = M[:N] if cells is None else M[cells]
        C = Mc @ d
        X = eng.X_tile(c, t)[:, :, :N] if cells is None else eng.X_tile(c, t)[:, :, cells]
        P = eng.P_tile(c, t)[:, :, :cl.s]
        rows = slice(tl.start, tl.start + tl.rows)
        ar = a[rows]
        if bins is not None:
            X, P, ar = X[bins], P[bins], ar[bins]
        Q = (P.double() @ b) * ar[:, None, :]
        return X.double() - Q @ C.T if with_x else Q @ C.T

    def __call__(self, A, B, D, M, with_x=True):
        tot = 0.0
        for c, cl in enumerate(self.eng.chroms):
            for t in range(len(cl.tiles)):
                tot = tot + self.resid(A, B, D, M, c, t, with_x=with_x).square().sum()
        return tot

    def fd(self, params, which, index, h=1e-6):
        """Central difference (L(+h) - L(-h)) / 2h, evaluated as sum (E+ - E-)(E+ + E-) / 2h over the
        residuals the entry touches (the same number without cancelling the two totals)."""
        eng = self.eng
        plus = [p.clone() for p in params]
        minus = [p.clone() for p in params]
        plus[which].view(-1)[index] += h
        minus[which].view(-1)[index] -= h
        parts = []
        if which == 3:                                                    # M: one cell, all chromosomes
            cell = index // eng.R
            for c, cl in enumerate(eng.chroms):
                for t in range(len(cl.tiles)):
                    parts.append((c, t, None, [cell]))
        else:
            c = self.chrom_of(which, index)
            cl = eng.chroms[c]
            for t, tl in enumerate(cl.tiles):
                if which == 0:                                            # A: one bin
                    i = (index - cl.aoff) // cl.sp
                    if tl.start <= i < tl.start + tl.rows:
                        parts.append((c, t, [i - tl.start], None))
                else:
                    parts.append((c, t, None, None))
        tot = 0.0
        for c, t, bins, cells in parts:
            ep = self.resid(*plus, c, t, bins, cells)
            em = self.resid(*minus, c, t, bins, cells)
            tot += float(((ep - em) * (ep + em)).sum())
        return tot / (2 * h)

    def chrom_of(self, which, index):
        for c, cl in enumerate(self.eng.chroms):
            if which == 0 and cl.aoff <= index < cl.aoff + cl.n * cl.sp:
                return c
            if which == 1 and cl.boff <= index < cl.boff + cl.s * cl.sp:
                return c
            if which == 2 and cl.doff <= index % self.eng.Dld < cl.doff + cl.sp:
                return c
        raise ValueError(index)


def _valid_entries(eng, rng, per_block=16):
    """random in-range (not padding) flat indices into A_all, B_all, D_all, M"""
    out = []
    for which in range(4):
        for _ in range(per_block):
            if which == 3:
                out.append((3, int(rng.integers(eng.N)) * eng.R + int(rng.integers(eng.R))))
                continue
            c = int(rng.integers(len(eng.chroms)))
            cl = eng.chroms[c]
            r = int(rng.integers(cl.s))
            if which == 0:
                out.append((0, cl.aoff + int(rng.integers(cl.n)) * cl.sp + r))
            elif which == 1:
                out.append((1, cl.boff + int(rng.integers(cl.s)) * cl.sp + r))
            else:
                out.append((2, int(rng.integers(eng.R)) * eng.Dld + cl.doff + r))
    return out


@pytest.fixture(scope="module")
def c4():
    from fasthigashi import tiling
    from fasthigashi.engine import SweepEngine
    from fasthigashi.packing import build_X, synth_device
    wl = syn.WORKLOADS["c4"]
    _say("c4: generating %d cells, bins %s" % (wl.n_cells, wl.bins))
    coos = synth_device(wl)
    tiles = [tiling.tile_plan(c.n, tiling.bs_bin_for(c.n, wl.resolution), wl.off_diag) for c in coos]
    sizes = [tiling.size_of(c.n, wl.dim1, wl.resolution, wl.rank) for c in coos]
    n_bins = [c.n for c in coos]
    eng = SweepEngine(tiles, n_bins, sizes, wl.n_cells, wl.rank, dtype="fp32")
    build_X(eng, coos, [0] * len(coos), do_rwr=wl.do_rwr)
    eng.set_factors(*tiling.structured_random_init(n_bins, sizes, wl.n_cells, wl.rank, seed=wl.seed))
    for _ in range(10):
        eng.sweep()
    torch.cuda.synchronize()
    _say("c4: 10 ALS sweeps, rec_error %s" % eng.history()["rec_error"][-3:])
    return eng


def test_loss_c4_against_fp64_autograd_and_differences(cuda, c4):
    """Errors of the A, B, D, M gradients at the ALS solution, measured against the factor-Gram term
    each is the difference of (the gradient of sum ||Q C^T||^2)."""
    from fasthigashi.loss import LossPlan, parafac2_loss
    eng = c4
    plan = LossPlan(eng)
    params = plan.params_from_engine()
    L = parafac2_loss(plan, *params)
    L.backward()
    torch.cuda.synchronize()
    tl = TorchLoss(eng)
    p64 = [p.detach().double().requires_grad_(True) for p in params]
    L64 = tl(*p64)
    L64.backward()
    q64 = [p.detach().double().requires_grad_(True) for p in params]
    tl(*q64, with_x=False).backward()
    L64 = float(L64.detach())
    rel = abs(float(L.detach()) - L64) / L64
    _say("c4: loss %.12e fp64 autograd %.12e rel %.2e" % (float(L.detach()), L64, rel))
    assert rel < 1e-6
    for name, p, q, r in zip("ABDM", params, p64, q64):
        d = float((p.grad.double() - q.grad).norm())
        _say("c4: grad %s |g| %.3e |gram term| %.3e err/|g| %.2e err/|gram term| %.2e"
             % (name, float(q.grad.norm()), float(r.grad.norm()), d / float(q.grad.norm()), d / float(r.grad.norm())))
        assert d < 1e-5 * float(r.grad.norm()), name
    rng = np.random.default_rng(0)
    with torch.no_grad():
        base = [p.detach().double() for p in params]
        worst = {}
        for which, index in _valid_entries(eng, rng):
            fd = tl.fd(base, which, index)
            ad = float(p64[which].grad.view(-1)[index])
            gpu = float(params[which].grad.view(-1)[index])
            scale = float(q64[which].grad.abs().max())
            w = worst.setdefault("ABDM"[which], [0.0, 0.0])
            w[0], w[1] = max(w[0], abs(ad - fd) / scale), max(w[1], abs(gpu - fd) / scale)
        for k, (e_fd, e_gpu) in worst.items():
            _say("c4: 16 entries of %s: |autograd - FD|, |HIP - FD| over max |gram term|: %.2e, %.2e"
                 % (k, e_fd, e_gpu))
            assert e_fd < 1e-8, (k, e_fd)
            assert e_gpu < 1e-5, (k, e_gpu)


def test_adam_refinement_c4(cuda, c4):
    from fasthigashi.loss import LossPlan, refine_adam
    plan = LossPlan(c4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    params, hist, final = refine_adam(plan, steps=200, lr=1e-3)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _say("c4 adam: 200 steps in %.2f s (%.1f ms/step: loss, gradient, Adam update)" % (dt, 1e3 * dt / 200))
    _say("c4 adam: start %.6e min %.6e final %.6e (steps 0, 1, 10, 50, 100, 199: %s)"
         % (hist[0], hist.min(), final, ", ".join("%.6e" % hist[k] for k in (0, 1, 10, 50, 100, 199))))
    assert np.all(np.isfinite(hist)) and math.isfinite(final)
    assert final < hist[0], "200 Adam steps did not lower the loss of the ALS solution"
    # padding entries of the packed layouts get zero gradient and stay zero
    A, B, D, M = [p.detach() for p in params]
    for cl in c4.chroms:
        assert float(A[cl.aoff: cl.aoff + cl.n * cl.sp].view(cl.n, cl.sp)[:, cl.s:].abs().sum()) == 0.0
        assert float(B[cl.boff: cl.boff + cl.s * cl.sp].view(cl.s, cl.sp)[:, cl.s:].abs().sum()) == 0.0
        assert float(D[:, cl.doff + cl.s: cl.doff + cl.sp].abs().sum()) == 0.0
    assert float(M[c4.N:].abs().sum()) == 0.0
