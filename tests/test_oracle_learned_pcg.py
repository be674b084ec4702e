"""CPU pins of the learned-preconditioner oracle (oracle/learned_pcg.py):
finite differences for the MLP gradients and for the unrolled-PCG gradient, numpy
linear algebra for the block inverses and PCG, and the BF16 rounding against exact
bit patterns."""

import numpy as np
import pytest

from oracle import learned_pcg as lp


def test_bf16_rounding_known_bits():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 3.0e38, np.inf], np.float32)
    r = lp.to_bf16(x)
    # ties round to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0 + 2 ** -7)  # above halfway -> rounds up
    assert r[3] == -2.5 and r[4] == 0.0 and np.isinf(r[6])
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    B, d0, h, d3 = 5, 7, 6, 4
    p = {k: v.astype(np.float64) for k, v in lp.init_params(d0, h, d3, seed=1).items()}
    p["b1"] += 0.1
    X = rng.standard_normal((B, d0))
    G = rng.standard_normal((B, d3))

    def loss(pp):
        Y, _ = lp.mlp_forward(X, pp, bf16=False, dtype=np.float64)
        return float(np.sum(Y * G))

    Y, cache = lp.mlp_forward(X, p, bf16=False, dtype=np.float64)
    g = lp.mlp_backward(cache, G, p, bf16=False, dtype=np.float64)
    eps = 1e-6
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        flat = p[name].reshape(-1)
        for i in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp_ = loss(p)
            flat[i] = old - eps
            lm_ = loss(p)
            flat[i] = old
            fd = (lp_ - lm_) / (2 * eps)
            assert abs(fd - g[name].reshape(-1)[i]) < 1e-6 * max(1.0, abs(fd)), name


def _small_system(seed=0, nm=4, ma=3):
    rng = np.random.default_rng(seed)
    N = np.array([ma, ma - 1, ma, 1][:nm], np.int32)
    n = nm + 3 * int(N.sum())
    M = rng.standard_normal((n, n))
    K = M @ M.T / n + 0.05 * np.eye(n)
    y = rng.standard_normal(n)
    return K, y, N, ma


def test_block_inverses_and_apply():
    K, y, N, ma = _small_system()
    blocks = lp.diag_blocks(K, N, ma)
    shift = np.abs(np.random.default_rng(1).standard_normal(blocks.shape[:2])) * 0.1
    inv = lp.block_inverses(blocks, N, ma, 1e-3, shift)
    idx = lp.block_index(N, ma)
    z = lp.apply_blocks(inv, N, ma, y)
    for b in range(len(N)):
        live = idx[b][idx[b] >= 0]
        m = len(live)
        Bm = K[np.ix_(live, live)] + 1e-3 * np.eye(m) + np.diag(shift[b, :m])
        np.testing.assert_allclose(z[live], np.linalg.solve(Bm, y[live]), rtol=1e-10)
    # every index covered exactly once
    assert sorted(idx[idx >= 0].tolist()) == list(range(len(y)))


def test_pcg_solves_spd_system():
    K, y, N, ma = _small_system(seed=2)
    inv = lp.block_inverses(lp.diag_blocks(K, N, ma), N, ma, 0.0)
    x, it, hist = lp.pcg(K, y, lambda r: lp.apply_blocks(inv, N, ma, r), tol=1e-12,
                         max_iter=200)
    np.testing.assert_allclose(x, np.linalg.solve(K, y), rtol=1e-8, atol=1e-10)
    assert hist[-1] <= 1e-12 and it <= len(y) + 5


@pytest.mark.parametrize("steps", [1, 3, 6])
def test_unrolled_grad_matches_finite_differences(steps):
    K, y, N, ma = _small_system(seed=3)
    blocks = lp.diag_blocks(K, N, ma)
    rng = np.random.default_rng(4)
    shift = np.abs(rng.standard_normal(blocks.shape[:2])) * 0.3
    lam = 1e-3

    def loss(s):
        return lp.unrolled_loss(K + lam * np.eye(len(y)), y,
                                lp.block_inverses(blocks, N, ma, lam, s), N, ma, steps)

    L, g = lp.unrolled_loss_grad(K + lam * np.eye(len(y)), y,
                                 lp.block_inverses(blocks, N, ma, lam, shift), N, ma, steps)
    assert abs(L - loss(shift)) < 1e-12
    n_live = 1 + 3 * np.clip(N, 0, ma)
    eps = 1e-6
    for b in range(len(N)):
        for l in range(n_live[b]):
            s1, s2 = shift.copy(), shift.copy()
            s1[b, l] += eps
            s2[b, l] -= eps
            fd = (loss(s1) - loss(s2)) / (2 * eps)
            assert abs(fd - g[b, l]) < 1e-6 * max(1.0, abs(fd)), (b, l, fd, g[b, l])
        assert np.all(g[b, n_live[b]:] == 0.0)


def test_features_shape_and_padding():
    X = np.random.default_rng(0).standard_normal((3, 4, 5))
    N = np.array([4, 2, 0], np.int32)
    f = lp.conformer_features(X, N, 2.0, 1e-10)
    assert f.shape == (3, 7)
    np.testing.assert_allclose(f[1, :5], X[1, :2].mean(0))
    np.testing.assert_allclose(f[2, :5], 0.0)
    np.testing.assert_allclose(f[:, 5], np.log(2.0))
    np.testing.assert_allclose(f[:, 6], -10.0)
