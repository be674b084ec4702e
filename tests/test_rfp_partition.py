"""CPU mirror of the GPU RFP Cholesky's panel partition and tile addressing.

kernelforge_amd/csrc/rfp_chol.hip factors an RFP (TRANSR='N', UPLO='U') array in
place with a right-looking blocked algorithm whose 64-wide panels are aligned at
k = n // 2, so each tile is a plain strided matrix (column-major for rows < k,
row-major for rows >= k).  This test replays exactly that tile schedule in NumPy
(same partition, same per-tile base/stride formulas) and checks
  * every address touched lies inside [0, n(n+1)/2) and matches LAPACK's own RFP
    layout (scipy dtrttf), and
  * the replayed algorithm reproduces numpy.linalg.cholesky and dpftrf's factor.
"""

import numpy as np
import pytest
import scipy.linalg as sl

NB = 64


def part(n):
    k = n // 2
    nlo = -(-k // NB)
    nblk = nlo + -(-(n - k) // NB)

    def start(J):
        return J * NB if J < nlo else k + (J - nlo) * NB

    def width(J):
        end = k if J < nlo else n
        return min(NB, end - start(J))

    return k, nblk, start, width


def tile_addr(n, r0, c0, r, c):
    k = n // 2
    s = n if n % 2 else n + 1
    if r0 < k:
        return (k + 1 + r0) + c0 * s + r * 1 + c * s
    return (r0 - k) * s + c0 + r * s + c


def rfp_lower(n, r, c):
    k = n // 2
    s = n if n % 2 else n + 1
    return (k + 1 + r) + c * s if r < k else (r - k) * s + c


NBO = 4  # 64-wide panels per outer trailing update


def range_decode(t, nblk, Kbeg):
    """Tile index t -> (K, I) exactly as syrk_range decodes it."""
    import math
    r = nblk - Kbeg
    q = int((2.0 * r + 1.0 - math.sqrt((2.0 * r + 1.0) ** 2 - 8.0 * t)) * 0.5)
    q = max(q, 0)
    while q > 0 and q * r - q * (q - 1) // 2 > t:
        q -= 1
    while (q + 1) * r - (q + 1) * q // 2 <= t:
        q += 1
    K = Kbeg + q
    return K, K + (t - (q * r - q * (q - 1) // 2))


def blocked_potrf(arf, n):
    """Replay of potrf_impl: per 256-column outer panel, 64-wide inner steps
    (diag POTRF, panel TRSM, syrk_range over the rest of the outer panel), then one
    256-deep syrk_range over everything to its right."""
    k, nblk, start, width = part(n)
    a = arf.copy()
    size = n * (n + 1) // 2

    def load(r0, rn, c0, cn, lower):
        T = np.zeros((NB, NB))
        for r in range(rn):
            for c in range(cn):
                if lower and r < c:
                    continue
                ad = tile_addr(n, r0, c0, r, c)
                assert 0 <= ad < size
                assert ad == rfp_lower(n, r0 + r, c0 + c)
                T[r, c] = a[ad]
        return T

    def store(T, r0, rn, c0, cn, lower):
        for r in range(rn):
            for c in range(cn):
                if lower and r < c:
                    continue
                a[tile_addr(n, r0, c0, r, c)] = T[r, c]

    def syrk_range(Jlo, Jhi, Kbeg, Kend):
        nt = sum(nblk - K for K in range(Kbeg, Kend))
        seen = set()
        for t in range(nt):
            K, I = range_decode(t, nblk, Kbeg)
            assert Kbeg <= K < Kend and K <= I < nblk
            seen.add((I, K))
            ri, rni, rk, rnk = start(I), width(I), start(K), width(K)
            C = load(ri, rni, rk, rnk, I == K)
            for J in range(Jlo, Jhi):
                Ai = load(ri, rni, start(J), width(J), False)
                Ak = load(rk, rnk, start(J), width(J), False)
                C -= Ai @ Ak.T
            store(C, ri, rni, rk, rnk, I == K)
        assert len(seen) == nt

    for J0 in range(0, nblk, NBO):
        J1 = min(J0 + NBO, nblk)
        for J in range(J0, J1):
            s0, w = start(J), width(J)
            L = load(s0, w, s0, w, True)[:w, :w]
            L = np.linalg.cholesky(np.tril(L) + np.tril(L, -1).T)
            inv = np.linalg.inv(L)
            T = np.zeros((NB, NB))
            T[:w, :w] = L
            store(T, s0, w, s0, w, True)
            for I in range(J + 1, nblk):
                r0, rn = start(I), width(I)
                Pn = load(r0, rn, s0, w, False)
                Pn[:rn, :w] = Pn[:rn, :w] @ inv.T
                store(Pn, r0, rn, s0, w, False)
            if J + 1 < J1:
                syrk_range(J, J + 1, J + 1, J1)
        if J1 < nblk:
            syrk_range(J0, J1, J1, nblk)
    return a


@pytest.mark.parametrize("n", [1, 2, 5, 64, 65, 127, 128, 129, 200, 257])
def test_partition_covers_and_aligns(n):
    k, nblk, start, width = part(n)
    cols = []
    for J in range(nblk):
        s, w = start(J), width(J)
        assert w >= 1
        assert (s < k) == (s + w - 1 < k) or k == 0  # no panel straddles k
        cols.extend(range(s, s + w))
    assert cols == list(range(n))


@pytest.mark.parametrize("n", [3, 64, 130, 203, 390])
def test_blocked_rfp_cholesky_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    K = A @ A.T + n * np.eye(n)
    trttf = sl.get_lapack_funcs("trttf", (K,))
    arf, info = trttf(np.asfortranarray(K), transr="N", uplo="U")
    assert info == 0
    for r in range(n):
        for c in range(r + 1):
            assert arf[rfp_lower(n, r, c)] == K[r, c]
    got = blocked_potrf(arf, n)
    pftrf = sl.get_lapack_funcs("pftrf", (K,))
    ref, info = pftrf(n, arf.copy(), transr="N", uplo="U")
    assert info == 0
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
    L = np.linalg.cholesky(K)
    for r in range(n):
        for c in range(r + 1):
            assert abs(got[rfp_lower(n, r, c)] - L[r, c]) < 1e-10 * max(1.0, abs(L[r, c]))
