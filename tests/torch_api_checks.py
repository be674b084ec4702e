"""Checks of the torch-tensor interface (kernelforge_amd.cuda_local_kernels and
kernelforge_amd.cuda_fchl19_repr) against the oracle.

Run as a process of its own (tests/test_gpu_torch_api.py starts it): torch has to load
its HIP runtime before libkf_hip.so, and the other GPU tests load the library without
torch.  Prints one line per check and exits non-zero on the first failure.
"""

import os
import sys

import torch  # noqa: I001 -- before any kernelforge_amd module

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import scipy.linalg as sl  # noqa: E402

import oracle  # noqa: E402
from kernelforge_amd import cuda_fchl19_repr as cfr  # noqa: E402
from kernelforge_amd import cuda_local_kernels as clk  # noqa: E402
from kernelforge_amd import synthetic  # noqa: E402

DEV = torch.device("cuda:0")
ELEMENTS = (1, 6, 7, 8, 16)
SIGMA = 2.0


def say(msg):
    print(msg, flush=True)


def rel(a, b):
    a = a.detach().cpu().numpy().astype(np.float64) if isinstance(a, torch.Tensor) else a
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) if b.size else 0.0


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t.to(dtype) if dtype is not None else t


def element_index(Q):
    lut = np.zeros(max(ELEMENTS) + 1, np.int32)
    for i, z in enumerate(ELEMENTS):
        lut[z] = i
    return lut[Q].astype(np.int32)


def unpack_lower(arf, n):
    arf = arf.detach().cpu().numpy().astype(np.float64)
    A = sl.get_lapack_funcs("tfttr", (arf,))(n, arf, transr="N", uplo="L")[0]
    return np.tril(A) + np.tril(A, -1).T


def check_representation():
    coords, zs, _ = synthetic.random_molecules(24, 0)
    C, Q, N = synthetic.pad_batch(coords, zs)
    Xo, dXo = oracle.fchl19_batch(C, Q, N, grad=True)
    Xo_rep = oracle.fchl19_batch(C, Q, N)
    Qi = element_index(Q)
    X = cfr.generate_fchl_acsf(dev(C), dev(Qi), dev(N), 5)
    assert X.dtype == torch.float64 and X.device.type == "cuda"
    e1 = rel(X, Xo_rep)
    X32 = cfr.generate_fchl_acsf(dev(C, torch.float32), dev(Qi), dev(N), 5)
    assert X32.dtype == torch.float32
    e2 = rel(X32, Xo_rep)
    Xg, dXg = cfr.generate_fchl_acsf_and_gradients(dev(C), dev(Qi), dev(N), 5)
    nm, ma = Q.shape
    assert tuple(dXg.shape) == (nm, ma, 720, ma, 3)
    e3 = rel(Xg, Xo)
    e4 = rel(dXg.reshape(nm, ma, 720, 3 * ma), dXo)
    Xs, dXs = cfr.generate_fchl_acsf_and_gradients_md_single(dev(C[:1]), dev(Qi[:1]), dev(N[:1]), 5)
    e5 = rel(dXs.reshape(1, ma, 720, 3 * ma), dXo[:1])
    say(f"representation: fp64 {e1:.1e}, fp32 in/out {e2:.1e}, with Jacobians {e3:.1e} / "
        f"{e4:.1e}, md_single {e5:.1e}")
    assert e1 < 1e-12 and e3 < 1e-12 and e4 < 1e-12 and e5 < 1e-12 and e2 < 1e-5


def check_scalar_kernels():
    coords, zs, _ = synthetic.random_molecules(40, 1)
    C, Q, N = synthetic.pad_batch(coords, zs)
    X = oracle.fchl19_batch(C, Q, N)
    Ko = oracle.kernel_symm(X, Q, N, SIGMA)
    K = clk.kernel_gaussian_symm(dev(X), dev(Q), dev(N), SIGMA)
    assert K.dtype == torch.float64
    e1 = rel(K, Ko)
    Kr = clk.kernel_gaussian_rect(dev(X[:15]), dev(Q[:15]), dev(N[:15]), dev(X), dev(Q), dev(N),
                                  SIGMA)
    e2 = rel(Kr, Ko[:15])
    arf = clk.kernel_gaussian_symm_rfp(dev(X), dev(Q), dev(N), SIGMA)
    e3 = rel(unpack_lower(arf, 40), Ko)
    arf32 = clk.kernel_gaussian_symm_rfp(dev(X, torch.float32), dev(Q), dev(N), SIGMA)
    assert arf32.dtype == torch.float32 and arf.dtype == torch.float64
    e4 = rel(unpack_lower(arf32, 40), Ko)
    # the reference's unpacking: kernelmath.rfp_to_full(K_rfp, n, uplo='U', transr='N')
    from kernelforge_amd import kernelmath

    e5 = rel(kernelmath.rfp_to_full(arf.cpu().numpy(), 40, uplo="U", transr="N"), Ko)
    say(f"scalar kernels: symm {e1:.1e}, rect {e2:.1e}, symm_rfp (UPLO=L) {e3:.1e}, "
        f"float32 in/out {e4:.1e}, kernelmath.rfp_to_full(uplo='U') {e5:.1e}")
    assert max(e1, e2, e3, e5) < 1e-12 and e4 < 1e-6


def conformer_set(n, seed):
    base, z = synthetic.aspirin_base()
    C, Z, E, F = synthetic.conformers(base, z, n, seed)
    Q = np.tile(Z, (n, 1)).astype(np.int32)
    N = np.full(n, len(Z), np.int32)
    return C, Z, E, F, Q, N


def check_energy_force_forms():
    C, Z, E, F, Q, N = conformer_set(5, 2)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    Ko = oracle.full_symm(X, dX, Q, N, SIGMA)
    K = clk.kernel_gaussian_full_symm(dev(X), dev(dX), dev(Q), dev(N), SIGMA)
    e1 = rel(K, Ko)
    arf = clk.kernel_gaussian_full_symm_rfp(dev(X), dev(dX), dev(Q), dev(N), SIGMA)
    big = Ko.shape[0]
    e2 = rel(unpack_lower(arf, big), Ko)
    alpha = np.random.default_rng(3).standard_normal(big)
    ad = clk.compute_alpha_desc(dev(dX), dev(N), dev(alpha[5:]))
    ado = oracle.alpha_desc(dX, Q, N, alpha[5:])
    e3 = rel(ad, ado)
    Eq, Fq = clk.kernel_gaussian_full_matvec(dev(X[:2]), dev(dX[:2]), dev(Q[:2]), dev(N[:2]),
                                             dev(X), dev(Q), dev(N), dev(alpha[:5]), ad, SIGMA)
    Eo, Fo = oracle.full_matvec(X[:2], dX[:2], X, ado, alpha[:5], Q[:2], Q, N[:2], N, SIGMA)
    e4, e5 = rel(Eq, Eo), rel(Fq, Fo)
    say(f"E+F forms: full_symm {e1:.1e}, full_symm_rfp (UPLO=L) {e2:.1e}, alpha_desc {e3:.1e}, "
        f"full_matvec E {e4:.1e} F {e5:.1e}")
    assert max(e1, e2, e3, e4, e5) < 1e-10


def check_md_cached_matvec(steps=100):
    """100 MD-style steps: the query molecule moves, its descriptors and Jacobians come from
    the md_single path and E, F from the cached matvec; every step against the oracle's
    full_matvec on the oracle's own descriptors."""
    C, Z, E, F, Q, N = conformer_set(12, 1)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    rng = np.random.default_rng(7)
    alpha = rng.standard_normal(12 + 12 * 63) * 1e-2
    aE = alpha[:12]
    dXt, Xt, Qt, Nt = dev(dX), dev(X), dev(Q), dev(N)
    ad = clk.compute_alpha_desc(dXt, Nt, dev(alpha[12:]))
    ado = oracle.alpha_desc(dX, Q, N, alpha[12:])
    pre = clk.precompute_train(Xt, Qt, Nt, dev(aE), ad)
    live = np.arange(21)[None, :] < N[:, None]
    assert rel(pre[0], (X * X).sum(-1).reshape(-1)) < 1e-13
    assert rel(pre[1], ((X * ado).sum(-1) * live).reshape(-1)) < 1e-13
    assert rel(pre[3], ((ado + aE[:, None, None] * X) * live[..., None]).reshape(-1, 720)) < 1e-13
    base, _ = synthetic.aspirin_base()
    pos = base[None].copy()
    vel = rng.normal(0.0, 0.01, size=pos.shape)
    qi = dev(element_index(Q[:1]))
    qz, n1 = dev(Q[:1]), dev(N[:1])
    aEd = dev(aE)
    worst = 0.0
    for step in range(steps):
        pos = pos + vel
        vel = vel - 0.05 * (pos - base[None])  # a spring back to the base geometry
        Xq, dXq = cfr.generate_fchl_acsf_and_gradients_md_single(dev(pos), qi, n1, 5)
        Eq, Fq = clk.kernel_gaussian_full_matvec_cached(Xq, dXq, qz, n1, Xt, Qt, Nt, aEd, ad,
                                                        *pre, SIGMA)
        Xo, dXo = oracle.fchl19_batch(pos, Q[:1], N[:1], grad=True)
        Eo, Fo = oracle.full_matvec(Xo, dXo, X, ado, aE, Q[:1], Q, N[:1], N, SIGMA)
        worst = max(worst, rel(Eq, Eo), rel(Fq, Fo))
        if step in (0, steps - 1):
            say(f"md step {step + 1}/{steps}: E {float(Eq[0]):.6f} vs {float(Eo[0]):.6f}, "
                f"max rel err so far {worst:.1e}")
    assert worst < 1e-10, worst
    return Xt, Qt, Nt, aEd, ad, pre, qz, n1, qi, pos, X, ado, aE, Q, N


def check_cached_graph(state, replays=5):
    """The cached matvec into fixed output buffers, captured into a graph after a warm-up
    call and replayed for new query geometries copied into the captured inputs; each
    replay against the eager call, with eager work and synchronisations in between."""
    Xt, Qt, Nt, aEd, ad, pre, qz, n1, qi, pos, X, ado, aE, Q, N = state
    Xq, dXq = cfr.generate_fchl_acsf_and_gradients_md_single(dev(pos), qi, n1, 5)
    Xs, dXs = Xq.clone(), dXq.clone()
    Eb = torch.empty(1, dtype=torch.float64, device=DEV)
    Fb = torch.empty(63, dtype=torch.float64, device=DEV)
    args = (Xt, Qt, Nt, aEd, ad, *pre, [], [], Eb, Fb, SIGMA)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        clk.kernel_gaussian_full_matvec_cached_graph(Xs, dXs, qz, n1, *args)  # warm-up: plans
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        clk.kernel_gaussian_full_matvec_cached_graph(Xs, dXs, qz, n1, *args)
    rng = np.random.default_rng(9)
    worst = 0.0
    for r in range(replays):
        p = pos + rng.normal(0.0, 0.02, size=pos.shape)
        Xq, dXq = cfr.generate_fchl_acsf_and_gradients_md_single(dev(p), qi, n1, 5)
        Xs.copy_(Xq)
        dXs.copy_(dXq)
        g.replay()
        torch.cuda.synchronize()
        Ee, Fe = clk.kernel_gaussian_full_matvec_cached(Xq, dXq, qz, n1, Xt, Qt, Nt, aEd, ad,
                                                        *pre, SIGMA)
        torch.cuda.synchronize()
        worst = max(worst, rel(Eb, Ee.cpu().numpy()), rel(Fb, Fe.cpu().numpy()))
    say(f"cached matvec in a captured graph: {replays} replays, max rel diff vs eager "
        f"{worst:.1e}")
    assert worst < 1e-13


def main():
    say(f"torch {torch.__version__}, device {torch.cuda.get_device_name(0)}")
    check_representation()
    check_scalar_kernels()
    check_energy_force_forms()
    state = check_md_cached_matvec()
    check_cached_graph(state)
    say("TORCH_API_OK")
    return 0


if __name__ == "__main__":
    sys.exit(main())
