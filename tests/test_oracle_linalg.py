"""Pins the oracle's polar factor (project2orthogonal.py:6-55) and CP-ALS
(parafac_integrative.py:12-112) by invariants and independent library routines (CPU)."""
import numpy as np
import pytest
import scipy.linalg

from oracle import ref_np as R


@pytest.mark.parametrize("m,k", [(40, 12), (12, 12), (300, 64), (9, 20)])
def test_polar_invariants(m, k):
    rng = np.random.default_rng(m * 100 + k)
    Y = rng.standard_normal((3, m, k))
    U, S = R.project2orthogonal(Y)
    r = min(m, k)
    for b in range(3):
        u = U[b]
        # orthonormal columns (tall) / rows (wide)
        g = u.T @ u if m >= k else u @ u.T
        assert np.allclose(g, np.eye(r), atol=1e-10)
        # Y^T U symmetric PSD and <Y, U> = sum of singular values = tr((Y^T Y)^{1/2})
        h = Y[b].T @ u if m >= k else u @ Y[b].T
        assert np.allclose(h, h.T, atol=1e-10)
        assert np.linalg.eigvalsh((h + h.T) / 2).min() > -1e-10
        assert np.isclose((Y[b] * u).sum(), S[b].sum(), rtol=1e-12)
        ev = np.linalg.eigvalsh(Y[b].T @ Y[b] if m >= k else Y[b] @ Y[b].T)
        assert np.isclose(np.sqrt(np.clip(ev, 0, None)).sum(), S[b].sum(), rtol=1e-10)
        # the unique polar factor of a full-rank matrix: scipy's polar decomposition
        up, _ = scipy.linalg.polar(Y[b], side="right")
        assert np.allclose(u, up, atol=1e-10)


def test_polar_eigh_fallback_agrees():
    rng = np.random.default_rng(9)
    Y = rng.standard_normal((2, 30, 8))
    U, S = R.project2orthogonal(Y)
    U2, _ = R.project2orthogonal_eigh(Y)
    assert np.allclose(U, U2, atol=1e-6)


def _rand_factors(rng, n, s, K):
    return [rng.random((n, s)) + 0.2, rng.random((s, s)) + 0.2, rng.random((K, s)) + 0.2]


def test_mttkrp_unfolding():
    """The MTTKRP of each mode against the explicit unfolding times the Khatri-Rao product."""
    rng = np.random.default_rng(1)
    n, s, K = 6, 4, 5
    X = rng.standard_normal((n, s, K))
    A, B, D = _rand_factors(rng, n, s, K)
    kr = lambda P, Q: np.stack([np.kron(P[:, r], Q[:, r]) for r in range(P.shape[1])], axis=1)
    assert np.allclose(R.mttkrp(X, [A, B, D], 0), X.reshape(n, s * K) @ kr(B, D))
    assert np.allclose(R.mttkrp(X, [A, B, D], 1), X.transpose(1, 0, 2).reshape(s, n * K) @ kr(A, D))
    assert np.allclose(R.mttkrp(X, [A, B, D], 2), X.transpose(2, 0, 1).reshape(K, n * s) @ kr(A, B))


def test_update_factor_normal_equations():
    rng = np.random.default_rng(2)
    s = 6
    F = rng.standard_normal((20, s))
    G = F.T @ F
    Mt = rng.standard_normal((11, s))
    out = R.update_factor(G, Mt)
    assert np.allclose(out @ (G + 1e-10 * np.eye(s)), Mt, atol=1e-10)


def test_balance_norm_invariants():
    rng = np.random.default_rng(3)
    f = _rand_factors(rng, 9, 5, 7)
    before = np.einsum("ir,jr,kr->ijk", *f)
    g = [x.copy() for x in f]
    R.balance_norm(g)
    assert np.allclose(np.linalg.norm(g[0], axis=0), 1.0)
    assert np.allclose(np.linalg.norm(g[1], axis=0), 1.0)
    assert np.allclose(np.einsum("ir,jr,kr->ijk", *g), before, rtol=1e-10)


def test_parafac_exact_fixed_point():
    """An exact rank-s tensor started at its own factors stays put (to rounding) and the loss
    terms are ||X||^2 (norm) and ||X||^2 (inner product)."""
    rng = np.random.default_rng(4)
    n, s, K = 12, 5, 9
    f = _rand_factors(rng, n, s, K)
    X = np.einsum("ir,jr,kr->ijk", *f)
    fac, norm, inner, hist = R.parafac(X, s, 5, [x.copy() for x in f])
    assert np.allclose(np.einsum("ir,jr,kr->ijk", *fac), X, rtol=1e-8)
    assert np.isclose(norm, (X ** 2).sum(), rtol=1e-10)
    assert np.isclose(inner, (X ** 2).sum(), rtol=1e-10)
    assert len(hist) <= 2          # wdiff < 1e-5 stops the loop (parafac_integrative.py:107)


def test_parafac_monotone_and_recovers():
    """From a perturbed start ALS decreases the loss monotonically and recovers a planted tensor."""
    rng = np.random.default_rng(5)
    n, s, K = 30, 4, 15
    f = _rand_factors(rng, n, s, K)
    X = np.einsum("ir,jr,kr->ijk", *f)
    init = [x + 0.05 * rng.standard_normal(x.shape) for x in f]
    fac, _, _, hist = R.parafac(X, s, 60, init)
    losses = [h["loss"] for h in hist]
    assert all(b <= a * (1 + 1e-9) for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 1e-4 * (X ** 2).sum()      # the wdiff < 1e-5 stop ends it early


def test_core_norm_closed_form():
    """Sum over tiles of ||[[A_t, B, D]]||^2 = 1^T (A^T A o B^T B o D^T D) 1 for any tiling."""
    rng = np.random.default_rng(6)
    A, B, D = _rand_factors(rng, 23, 5, 8)
    cf = float(np.ones(5) @ ((A.T @ A) * (B.T @ B) * (D.T @ D)) @ np.ones(5))
    for bs in (23, 7, 1):
        assert np.isclose(R.core_norm(A, B, D, bs), cf, rtol=1e-12)
