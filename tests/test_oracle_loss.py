"""The oracle's reconstruction loss and its gradient, pinned without the product:
central finite differences on every factor entry, an exactly representable model, and the
reference's norm decomposition of the loss."""
import numpy as np
import pytest

from oracle import ref_np as R


def _case(seed, orthonormal, n_cells=9, rank=5, bins=(7, 5), sizes=(3, 4), bs=3, flank=2):
    rng = np.random.default_rng(seed)
    data, proj, A, B, D = [], [], [], [], []
    for c, (n, s) in enumerate(zip(bins, sizes)):
        tiles = R.tile_plan(n, bs, flank)
        X = [rng.standard_normal((t.rows, t.width, n_cells)) for t in tiles]
        P = []
        for t in tiles:
            p = rng.standard_normal((t.rows, t.width, s))
            if orthonormal:
                p = np.stack([np.linalg.qr(q)[0] for q in p])
            P.append(p)
        data.append(R.ChromData("chr%d" % (c + 1), n, 1, tiles, X, bs, [slice(0, n_cells)]))
        proj.append(P)
        A.append(rng.standard_normal((n, s)))
        B.append(rng.standard_normal((s, s)))
        D.append(rng.standard_normal((rank, s)))
    M = rng.standard_normal((n_cells, rank))
    if orthonormal:
        M = np.linalg.qr(M)[0]
    return data, proj, R.Factors(A, B, D, M)


def _entries(F):
    """(name, chrom, index) of every factor entry."""
    out = []
    for name in ("A", "B", "D"):
        for c, a in enumerate(getattr(F, name)):
            out += [(name, c, ix) for ix in np.ndindex(a.shape)]
    out += [("M", None, ix) for ix in np.ndindex(F.M.shape)]
    return out


def _get(F, name, c):
    return F.M if name == "M" else getattr(F, name)[c]


@pytest.mark.parametrize("orthonormal", [False, True])
def test_gradient_matches_central_differences(orthonormal):
    data, proj, F = _case(1, orthonormal)
    g = R.parafac2_loss_grad(data, F, proj)
    h = 1e-6
    worst = 0.0
    for name, c, ix in _entries(F):
        Fp, Fm = F.copy(), F.copy()
        _get(Fp, name, c)[ix] += h
        _get(Fm, name, c)[ix] -= h
        fd = (R.parafac2_loss(data, Fp, proj) - R.parafac2_loss(data, Fm, proj)) / (2 * h)
        an = _get(g, name, c)[ix]
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    assert worst < 1e-6, worst


def test_exact_model_has_zero_loss_and_gradient():
    data, proj, F = _case(2, True)
    for c, cd in enumerate(data):
        C = F.M @ F.D[c]
        for t, tile in enumerate(cd.tiles):
            rows = slice(tile.start, tile.start + tile.rows)
            W = F.B[c][None] * F.A[c][rows][:, None, :]
            cd.X[t] = proj[c][t] @ W @ C.T
    scale = sum(float(np.square(x).sum()) for cd in data for x in cd.X)
    assert R.parafac2_loss(data, F, proj) < 1e-24 * scale
    g = R.parafac2_loss_grad(data, F, proj)
    for name, c, ix in _entries(F)[::7]:
        assert abs(_get(g, name, c)[ix]) < 1e-10 * np.sqrt(scale)


def test_loss_is_the_reference_norm_decomposition():
    """Orthonormal P_i and M: ||X||^2 - 2 <P^T X M, [[A, B, D]]> + ||[[A, B, D]]||^2, the terms the
    reference combines into its reconstruction error (parafac2_intergrative.py:621-632, 667-668)."""
    data, proj, F = _case(3, True)
    tnorm = sum(float(np.square(x).sum()) for cd in data for x in cd.X)
    cross = 0.0
    core = 0.0
    for c, cd in enumerate(data):
        core += R.core_norm(F.A[c], F.B[c], F.D[c], cd.bs_bin)
        for t, tile in enumerate(cd.tiles):
            rows = slice(tile.start, tile.start + tile.rows)
            T = np.swapaxes(proj[c][t], 1, 2) @ cd.X[t] @ F.M                   # (rows, s, R)
            model = np.einsum("ir,jr,kr->ijk", F.A[c][rows], F.B[c], F.D[c])
            cross += float((T * model).sum())
    expect = tnorm - 2 * cross + core
    got = R.parafac2_loss(data, F, proj)
    assert abs(got - expect) < 1e-10 * tnorm, (got, expect)
