"""CPU restatement (NumPy, FP64/FP32) of the learned block-Jacobi PCG pieces.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU baseline as the checker of the GPU path, never by the product.

None of this has a counterpart in the reference (it solves energy+force systems by
RFP Cholesky only, src/math.cpp:134-171); the pieces are pinned by mathematics in
tests/test_oracle_learned_pcg.py:
  * mlp_forward / mlp_backward  -- central finite differences of the loss in FP64
    (with the BF16 rounding switched off) and an independent autograd-free check of
    the ReLU masks;
  * block_inverses               -- numpy.linalg.inv of each shifted block;
  * pcg                          -- numpy.linalg.solve on the same system;
  * unrolled_loss_grad           -- central finite differences of the unrolled loss
    with respect to every shift entry.

Layout of the energy+force system: [E_0..E_{nm-1}, F_0 (3 N_0), F_1, ...]; block b is
(E_b, F_b) with local size 1 + 3 N_b (padded to bs = 1 + 3 max_atoms).
"""

import numpy as np


def to_bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(finite, r, u >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32)


def init_params(d0, h, d3, seed=0, out_bias=0.0):
    """He-initialised weights (fan_in x fan_out, float32) and biases."""
    rng = np.random.default_rng(seed)
    p = {
        "W1": (rng.standard_normal((d0, h)) * np.sqrt(2.0 / d0)).astype(np.float32),
        "b1": np.zeros(h, np.float32),
        "W2": (rng.standard_normal((h, h)) * np.sqrt(2.0 / h)).astype(np.float32),
        "b2": np.zeros(h, np.float32),
        "W3": (rng.standard_normal((h, d3)) * np.sqrt(1.0 / h) * 0.1).astype(np.float32),
        "b3": np.full(d3, out_bias, np.float32),
    }
    return p


def mlp_forward(X, p, bf16=True, dtype=np.float32):
    q = (lambda w: to_bf16(w)) if bf16 else (lambda w: w)
    X = X.astype(dtype)
    Z1 = X @ q(p["W1"]).astype(dtype) + p["b1"].astype(dtype)
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ q(p["W2"]).astype(dtype) + p["b2"].astype(dtype)
    A2 = np.maximum(Z2, 0)
    Y = A2 @ q(p["W3"]).astype(dtype) + p["b3"].astype(dtype)
    return Y, (X, Z1, Z2)


def mlp_backward(cache, dY, p, bf16=True, dtype=np.float32):
    q = (lambda w: to_bf16(w)) if bf16 else (lambda w: w)
    X, Z1, Z2 = cache
    dY = dY.astype(dtype)
    A1, A2 = np.maximum(Z1, 0), np.maximum(Z2, 0)
    g = {"W3": A2.T @ dY, "b3": dY.sum(0)}
    dZ2 = (dY @ q(p["W3"]).astype(dtype).T) * (Z2 > 0)
    g["W2"], g["b2"] = A1.T @ dZ2, dZ2.sum(0)
    dZ1 = (dZ2 @ q(p["W2"]).astype(dtype).T) * (Z1 > 0)
    g["W1"], g["b1"] = X.T @ dZ1, dZ1.sum(0)
    return g


def block_index(N, max_atoms):
    """Global indices (nb, bs) of each block's local positions (-1 for padding)."""
    nm = len(N)
    bs = 1 + 3 * max_atoms
    n = np.clip(N, 0, max_atoms)
    offs = np.concatenate([[0], np.cumsum(3 * n)])
    idx = -np.ones((nm, bs), np.int64)
    for b in range(nm):
        idx[b, 0] = b
        idx[b, 1:1 + 3 * n[b]] = nm + offs[b] + np.arange(3 * n[b])
    return idx


def diag_blocks(K, N, max_atoms):
    """(nb, bs, bs) diagonal blocks of a dense E+F matrix (zeros in padding)."""
    idx = block_index(N, max_atoms)
    nb, bs = idx.shape
    out = np.zeros((nb, bs, bs))
    for b in range(nb):
        live = idx[b][idx[b] >= 0]
        m = len(live)
        out[b, :m, :m] = K[np.ix_(live, live)]
    return out


def block_inverses(blocks, N, max_atoms, lam, shift=None):
    nb, bs, _ = blocks.shape
    n = 1 + 3 * np.clip(N, 0, max_atoms)
    inv = np.zeros_like(blocks)
    for b in range(nb):
        m = n[b]
        B = blocks[b, :m, :m] + lam * np.eye(m)
        if shift is not None:
            B = B + np.diag(shift[b, :m])
        inv[b] = np.eye(bs)
        inv[b, :m, :m] = np.linalg.inv(B)
    return inv


def apply_blocks(inv, N, max_atoms, r):
    idx = block_index(N, max_atoms)
    z = np.zeros_like(r)
    for b in range(idx.shape[0]):
        live = idx[b][idx[b] >= 0]
        m = len(live)
        z[live] = inv[b, :m, :m] @ r[live]
    return z


def pcg(A, y, prec, tol=1e-8, max_iter=1000):
    x = np.zeros_like(y)
    r = y.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    ny = np.linalg.norm(y)
    hist = [1.0]
    k = 0
    while k < max_iter and np.linalg.norm(r) / ny > tol:
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        hist.append(np.linalg.norm(r) / ny)
        z = prec(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
        k += 1
    return x, k, np.array(hist)


def unrolled_loss(A, y, inv, N, max_atoms, steps):
    """log(||r_steps|| / ||y||) of `steps` PCG iterations from x = 0."""
    prec = lambda r: apply_blocks(inv, N, max_atoms, r)
    r = y.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    for k in range(steps):
        q = A @ p
        al = rz / (p @ q)
        r = r - al * q
        if k + 1 < steps:
            z = prec(r)
            rz2 = r @ z
            p = z + (rz2 / rz) * p
            rz = rz2
    return 0.5 * np.log((r @ r) / (y @ y))


def unrolled_loss_grad(A, y, inv, N, max_atoms, steps):
    """Loss and d loss / d shift (nb, bs) where inv = (K_bb + lambda + diag(shift))^-1,
    by reverse-mode differentiation of the PCG recurrences."""
    idx = block_index(N, max_atoms)
    prec = lambda r: apply_blocks(inv, N, max_atoms, r)
    K = steps
    R, Z, P, Q = [y.copy()], [], [], []
    rho, sig, alp, bet = [], [], [], []
    Z.append(prec(R[0]))
    P.append(Z[0].copy())
    rho.append(R[0] @ Z[0])
    for k in range(K):
        Q.append(A @ P[k])
        sig.append(P[k] @ Q[k])
        alp.append(rho[k] / sig[k])
        R.append(R[k] - alp[k] * Q[k])
        if k + 1 < K:
            Z.append(prec(R[k + 1]))
            rho.append(R[k + 1] @ Z[k + 1])
            bet.append(rho[k + 1] / rho[k])
            P.append(Z[k + 1] + bet[k] * P[k])
    rr = R[K] @ R[K]
    loss = 0.5 * np.log(rr / (y @ y))
    grad = np.zeros(idx.shape)

    def prec_adj(zb, k, rb):
        t = prec(zb)
        for b in range(idx.shape[0]):
            live = idx[b][idx[b] >= 0]
            grad[b, :len(live)] -= t[live] * Z[k][live]
        return rb + t

    rb = R[K] / rr
    pb = np.zeros_like(y)
    rho_b = np.zeros(K + 1)
    for k in range(K - 1, -1, -1):
        if k + 1 < K:
            zb = pb.copy()
            betb = pb @ P[k]
            pb = bet[k] * pb
            rho_b[k + 1] += betb / rho[k]
            rho_b[k] += -betb * rho[k + 1] / rho[k] ** 2
            rb = rb + rho_b[k + 1] * Z[k + 1]
            zb = zb + rho_b[k + 1] * R[k + 1]
            rb = prec_adj(zb, k + 1, rb)
        alpb = -(rb @ Q[k])
        qb = -alp[k] * rb
        rho_b[k] += alpb / sig[k]
        sigb = -alpb * rho[k] / sig[k] ** 2
        pb = pb + sigb * Q[k]
        qb = qb + sigb * P[k]
        pb = pb + A @ qb
    zb = pb + rho_b[0] * R[0]
    prec_adj(zb, 0, rb)
    return loss, grad


def conformer_features(X, N, sigma, lam):
    """Per-conformer network input: mean descriptor over the live atoms, log sigma,
    log10 lambda (float64; the network sees it after standardisation)."""
    nm, ma, rep = X.shape
    n = np.clip(N, 1, ma)
    mask = (np.arange(ma)[None, :] < np.clip(N, 0, ma)[:, None]).astype(np.float64)
    mean = (X * mask[..., None]).sum(1) / n[:, None]
    extra = np.tile([np.log(sigma), np.log10(lam)], (nm, 1))
    return np.concatenate([mean, extra], axis=1)
