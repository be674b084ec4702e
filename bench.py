#!/usr/bin/env python3
"""Benchmark: FCHL19 local-kernel KRR energy fit on one MI355X.

Workload (one "step" = one full fit, inputs resident in HBM before the clock starts):
  20,000 synthetic molecules (seed 1, 5-23 atoms, H/C/N/O/S) -> FCHL19 descriptors
  (720-dim, FP64) -> element-matched local Gaussian kernel (sigma = 2) written as a
  20,000 x 20,000 FP64 RFP matrix -> RFP Cholesky with lambda = 1e-8 -> solve for alpha.
Targets: per-molecule energies minus an element-count baseline (least squares), as the
reference's LocalKRRModel fits them.  A 2,000-molecule held-out set (the next draws of
the same stream) gives the test MAE, computed after the timed region.

Prints ONE JSON line (rank 0).  `roofline` is for the dominant kernel (the symmetric
local kernel), timed with HIP events recorded by libkf_hip on the launch stream;
`cpu_baseline` is the oracle (oracle/, the reference's algorithm on OpenBLAS/OpenMP)
on a bounded sample of the same workload, scaled to the full size.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--cpu-sample 10000]

The default run also measures the 500-conformer energy+force fit and the 2,000-conformer
learned-preconditioner energy+force fit (reported beside the headline).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("KRR solve s @≤1e-6 rel err vs FP64: plain vs mixed+learned PCG; kernel TFLOP/s")
PEAK_F64_MFMA_TFLOPS = 78.6
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_INT8_MFMA_TOPS = 5000.0
SIGMA, LAMBDA = 2.0, 1e-8
ENERGY_REL_BAR = 1e-6
# a reduced-precision path replaces the FP64 fit as the head only when it is clearly faster:
# fit times repeat to about 1% between runs, and on a tie the reference precision stands
HEAD_MARGIN = 0.98
ELEMENTS = (1, 6, 7, 8, 16)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def element_baseline(z_list, energies):
    cnt = np.array([[np.sum(z == e) for e in ELEMENTS] for z in z_list], dtype=np.float64)
    coef, *_ = np.linalg.lstsq(cnt, energies, rcond=None)
    return cnt, coef


def label_rows(Q, N):
    ma = Q.shape[1]
    mask = np.arange(ma)[None, :] < np.clip(N, 0, ma)[:, None]
    labs, counts = np.unique(Q[mask], return_counts=True)
    return dict(zip(labs.tolist(), counts.tolist()))


def cpu_threads():
    try:
        return int(os.environ.get("OMP_NUM_THREADS", "")) or len(os.sched_getaffinity(0))
    except ValueError:
        return len(os.sched_getaffinity(0))


def host_cpu():
    """Model name and thread counts of the host the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count(),
            "affinity_threads": len(os.sched_getaffinity(0)), "threads_used": cpu_threads()}


def cpu_baseline(C, Q, N, y, sample, full_rows_sq, n_full):
    """Oracle pipeline on the first `sample` training molecules, scaled to the full fit:
    descriptors x n, kernel x sum_e R_e^2, Cholesky x n^3."""
    import oracle
    from threadpoolctl import threadpool_limits

    cores = cpu_threads()
    with threadpool_limits(limits=cores):
        Cs, Qs, Ns, ys = C[:sample], Q[:sample], N[:sample], y[:sample]
        oracle.fchl19_batch(Cs[:64], Qs[:64], Ns[:64])  # warm-up (thread pool spin-up)
        t0 = time.perf_counter()
        Xs = oracle.fchl19_batch(Cs, Qs, Ns)
        t1 = time.perf_counter()
        arf = oracle.kernel_symm_rfp(Xs, Qs, Ns, SIGMA)
        t2 = time.perf_counter()
        oracle.cho_solve_rfp(arf, ys, LAMBDA)
        t3 = time.perf_counter()
    rs_sample = sum(v * v for v in label_rows(Qs, Ns).values())
    s_desc = (t1 - t0) * n_full / sample
    s_kern = (t2 - t1) * full_rows_sq / rs_sample
    s_chol = (t3 - t2) * (n_full / sample) ** 3
    return {
        "value": round(s_desc + s_kern + s_chol, 3),
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "host": host_cpu(),
        "measured_s": round(t3 - t0, 2),
        "note": ("the oracle's exp-sum stops a molecule pair's bucket scan at its last "
                 "matching row, which the reference's bucketed scan does not; the CPU time "
                 "is therefore if anything lower than the reference's"),
        "sample": (f"oracle (C/OpenMP + scipy OpenBLAS) on the first {sample} of the 20,000 "
                   f"training molecules: descriptors {t1 - t0:.2f} s, FP64 symmetric RFP "
                   f"kernel {t2 - t1:.2f} s, dpftrf+dpftrs {t3 - t2:.3f} s; scaled to 20,000 "
                   f"by n (descriptors), sum_e R_e^2 (kernel) and n^3 (Cholesky): "
                   f"{s_desc:.1f} + {s_kern:.1f} + {s_chol:.1f} s"),
    }


EF_LAMBDA = 1e-10


def ef_pair_flop(na=21, matched=161, rep=720):
    """Algorithmic FLOP of one molecule pair of the full E+F kernel (EE, EF, FE, FF):
    three projections and a rank-1 update per matched atom pair plus the T_i contraction,
    over `rep` descriptor columns."""
    nc = 3 * na
    return matched * (6 * rep * nc + 2 * nc * nc) + na * 2 * nc * nc * rep


def active_columns(X, dX):
    """Descriptor columns that are non-zero for some atom, in X or in the Jacobian dX (the
    columns every E+F product has to run over; the others are exact zeros)."""
    nz = np.any(X.reshape(-1, X.shape[-1]) != 0, axis=0)
    nz |= np.any(dX.reshape(-1, dX.shape[-2], dX.shape[-1]) != 0, axis=(0, 2))
    return int(nz.sum())


def window_gram_flop(X, Q, N, width=16):
    """FLOP of the symmetric Gram products when each pair of same-label rows multiplies only
    the `width`-column windows both rows use: width * sum_e sum_w c_ew^2, with c_ew the rows
    of label e that are non-zero somewhere in window w (the symmetric half of
    2 * width * c_ew^2).  This is the work of an ideal window-skipping kernel; the dense
    count is rep * sum_e R_e^2."""
    nm, ma, rep = X.shape
    nw = -(-rep // width)
    valid = np.arange(ma)[None, :] < np.clip(N, 0, ma)[:, None]
    tot = 0.0
    for e in np.unique(Q[valid]):
        rows = X[valid & (Q == e)]
        if nw * width != rep:
            rows = np.pad(rows, ((0, 0), (0, nw * width - rep)))
        c = (rows.reshape(len(rows), nw, width) != 0).any(-1).sum(0).astype(np.float64)
        tot += width * float(np.sum(c * c))
    return tot


def ef_cpu_baseline(C, Z, E, F, sample, n_full):
    """Oracle E+F fit (descriptors + Jacobians, full RFP kernel, dpftrf/dpftrs) on
    `sample` conformers, scaled to n_full by n (descriptors), pairs (kernel), n^3 (solve)."""
    import oracle
    from threadpoolctl import threadpool_limits

    cores = cpu_threads()
    Q = np.tile(Z, (sample, 1)).astype(np.int32)
    Nn = np.full(sample, len(Z), np.int32)
    with threadpool_limits(limits=cores):
        oracle.fchl19_batch(C[:2], Q[:2], Nn[:2], grad=True)
        t0 = time.perf_counter()
        X, dX = oracle.fchl19_batch(C[:sample], Q, Nn, grad=True)
        t1 = time.perf_counter()
        arf = oracle.full_symm_rfp(X, dX, Q, Nn, SIGMA)
        t2 = time.perf_counter()
        y = np.concatenate([E[:sample], -F[:sample].reshape(-1)])
        oracle.cho_solve_rfp(arf, y, EF_LAMBDA)
        t3 = time.perf_counter()
    big_s, big_f = sample * (1 + 3 * len(Z)), n_full * (1 + 3 * len(Z))
    s_desc = (t1 - t0) * n_full / sample
    s_kern = (t2 - t1) * (n_full * (n_full + 1)) / (sample * (sample + 1))
    s_chol = (t3 - t2) * (big_f / big_s) ** 3
    return {
        "value": round(s_desc + s_kern + s_chol, 2), "unit": "s", "cores": cores, "kind": "port",
        "host": host_cpu(), "measured_s": round(t3 - t0, 2),
        "sample": (f"oracle on {sample} of {n_full} conformers: descriptors+Jacobians "
                   f"{t1 - t0:.2f} s, full E+F RFP kernel {t2 - t1:.2f} s, dpftrf+dpftrs "
                   f"{t3 - t2:.3f} s (order {big_s}); scaled by n, pairs and n^3: "
                   f"{s_desc:.1f} + {s_kern:.1f} + {s_chol:.1f} s"),
    }


def ef_leg(args, L):
    """Plain energy+force fit: 500 C9H8O4 conformers (seed 2), descriptors + Jacobians,
    full E+F kernel (order 32,000) in FP64 RFP, FP64 RFP Cholesky (lambda 1e-10), solve;
    held-out conformers predicted through alpha_desc + the full matvec."""
    from kernelforge_amd import capi, synthetic

    nconf, ntest = args.ef_conformers, 100
    base, z = synthetic.aspirin_base()
    C, Z, E, F = synthetic.conformers(base, z, nconf, 2)
    Ct, _, Et, Ft = synthetic.conformers(base, z, ntest, 102)
    ma, rs = len(Z), capi.rep_size()
    big = nconf * (1 + 3 * ma)
    DA = capi.DeviceArray
    Q = np.tile(Z, (nconf, 1)).astype(np.int32)
    Nn = np.full(nconf, ma, np.int32)
    dC, dQ, dN = DA.from_numpy(C), DA.from_numpy(Q), DA.from_numpy(Nn)
    X, dX = DA((nconf, ma, rs)), DA((nconf, ma, rs, 3 * ma))
    arf = DA(big * (big + 1) // 2)
    y = np.concatenate([E, -F.reshape(-1)])
    dy, alpha = DA.from_numpy(y), DA(big)
    log(f"E+F fit: {nconf} conformers, order {big}, RFP {arf.nbytes / 1e9:.2f} GB, "
        f"dX {dX.nbytes / 1e9:.2f} GB")

    def step():
        capi.fchl19(dC.ptr, dQ.ptr, dN.ptr, nconf, ma, X.ptr, dX.ptr)
        capi.local_kernel_full_symm_rfp(X.ptr, dX.ptr, dQ.ptr, dN.ptr, nconf, ma, rs, SIGMA,
                                        arf.ptr)
        capi.rfp_potrf(capi.KF_PREC_F64, arf.ptr, big, EF_LAMBDA)
        alpha.copy_from(dy)
        capi.rfp_potrs(capi.KF_PREC_F64, arf.ptr, big, alpha.ptr, 1)

    for _ in range(args.warmup):
        step()
        capi.synchronize()
    L.kf_timer_reset()
    L.kf_timer_enable(1)
    steps = max(1, min(args.steps, 2))
    capi.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        capi.synchronize()
        log(f"E+F fit: step {k + 1}/{steps} done")
    sec = (time.perf_counter() - t0) / steps
    L.kf_timer_enable(0)
    tm = {k: capi.timer_read(v) for k, v in (("fchl19", capi.TM_FCHL19), ("kernel", capi.TM_FULL),
                                             ("potrf", capi.TM_POTRF), ("potrs", capi.TM_POTRS))}
    ms = {k: (t / c if c else None) for k, (t, c) in tm.items()}
    # held-out predictions (after the timed region)
    a = alpha.numpy()
    Qt = np.tile(Z, (ntest, 1)).astype(np.int32)
    Nt = np.full(ntest, ma, np.int32)
    dCt, dQt, dNt = DA.from_numpy(Ct), DA.from_numpy(Qt), DA.from_numpy(Nt)
    Xt, dXt = DA((ntest, ma, rs)), DA((ntest, ma, rs, 3 * ma))
    capi.fchl19(dCt.ptr, dQt.ptr, dNt.ptr, ntest, ma, Xt.ptr, dXt.ptr)
    ad, aF, aE = DA((nconf, ma, rs)), DA.from_numpy(a[nconf:]), DA.from_numpy(a[:nconf])
    capi.local_alpha_desc(dX.ptr, dQ.ptr, dN.ptr, nconf, ma, rs, aF.ptr, ad.ptr)
    Ep, Fp = DA(ntest), DA(ntest * 3 * ma)
    capi.local_full_matvec(Xt.ptr, dXt.ptr, X.ptr, ad.ptr, aE.ptr, dQt.ptr, dQ.ptr, dNt.ptr,
                           dN.ptr, ntest, nconf, ma, ma, rs, SIGMA, Ep.ptr, Fp.ptr)
    capi.synchronize()
    e_mae = float(np.mean(np.abs(Ep.numpy() - Et)))
    f_mae = float(np.mean(np.abs(-Fp.numpy() - Ft.reshape(-1))))
    # columns of the conformer set that are non-zero somewhere (X of every conformer, the
    # Jacobians of the first 8): the kernels drop the others before the products
    nf = active_columns(X.numpy(), dX.numpy()[:8])
    flop = ef_pair_flop(ma, rep=nf) * nconf * (nconf + 1) / 2
    flop_dense = ef_pair_flop(ma) * nconf * (nconf + 1) / 2
    kt = flop / (ms["kernel"] * 1e-3) / 1e12
    out = {
        "workload": (f"energy+force KRR fit, {nconf} synthetic C9H8O4 conformers (seed 2): "
                     "descriptors + Jacobians, full E+F kernel in FP64 RFP (order "
                     f"{big}), FP64 RFP Cholesky lambda 1e-10, solve"),
        "s_per_fit": round(sec, 4),
        "breakdown_ms": {k: (round(v, 3) if v is not None else None) for k, v in ms.items()},
        "roofline": {"bound": "mfma", "achieved": round(kt, 3), "peak": PEAK_F64_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(kt / PEAK_F64_MFMA_TFLOPS, 4),
                     "traffic": None, "kernel": "deriv_vproj_kernel + deriv_ff_kernel "
                     "<FULL_SYMM_RFP>", "flop_per_launch": flop, "active_columns": nf,
                     "flop_per_launch_all_720_columns": flop_dense,
                     "tflops_all_720_columns": round(flop_dense / (ms["kernel"] * 1e-3) / 1e12, 3)},
        "test_mae": {"energy_ha": round(e_mae, 6), "force_ha_per_a": round(f_mae, 6),
                     "n_test_conformers": ntest},
    }
    log(f"E+F fit {sec:.3f} s/fit, breakdown {out['breakdown_ms']}, kernel {kt:.1f} TFLOP/s, "
        f"test MAE E {e_mae:.4g} F {f_mae:.4g}")
    if args.ef_cpu_sample > 0:
        log(f"E+F CPU baseline on {args.ef_cpu_sample} conformers ...")
        out["cpu_baseline"] = ef_cpu_baseline(C, Z, E, F, args.ef_cpu_sample, nconf)
        log(f"E+F CPU baseline {out['cpu_baseline']['value']} s (scaled)")
    return out


def learned_pcg_leg(args, L):
    """Energy+force fit of 2,000 C9H8O4 conformers (seed 3, order 128,000, lambda 1e-10):
    the shift network is trained on kernel systems of random 250-conformer subsets, then
    PCG with the matrix-free full matvec runs with the learned block-Jacobi correction and
    with plain block-Jacobi (fallback), both capped; FP64 RFP Cholesky on the full system
    is the direct solve and the reference the PCG energies are measured against."""
    from kernelforge_amd import capi, synthetic
    from kernelforge_amd import learned_pcg as LP

    lam, nconf, ntest = EF_LAMBDA, args.pcg_conformers, 100
    base, z = synthetic.aspirin_base()
    C, Z, E, F = synthetic.conformers(base, z, nconf, 3)
    Ct, _, Et, Ft = synthetic.conformers(base, z, ntest, 103)
    t0 = time.perf_counter()
    full = LP.DeviceSystem(C, Z, E, F)
    capi.synchronize()
    log(f"learned PCG: {nconf} conformers, order {full.n}; descriptors + Jacobians + "
        f"diagonal blocks {time.perf_counter() - t0:.2f} s")
    # ---- training systems: random subsets of the training conformers ----
    rng = np.random.default_rng(3)
    systems = []
    t0 = time.perf_counter()
    for k in range(args.pcg_systems):
        idx = np.sort(rng.choice(nconf, args.pcg_subset, replace=False))
        s_ = LP.DeviceSystem(C[idx], Z, E[idx], F[idx])
        s_.build_rfp()
        s_.drop_jacobians()
        systems.append(s_)
    capi.synchronize()
    t_build = time.perf_counter() - t0
    log(f"learned PCG: {len(systems)} training systems of {args.pcg_subset} conformers built "
        f"in {t_build:.1f} s")
    net = LP.ShiftNet(722, full.bs, seed=0)
    t0 = time.perf_counter()
    hist = LP.train(net, systems, lam, epochs=args.pcg_epochs, steps=10, log=log)
    capi.synchronize()
    t_train = time.perf_counter() - t0
    zero_loss = float(np.mean([LP.unrolled_loss_grad(s_, lam, LP.factor(s_, lam), 10)[0]
                               for s_ in systems[:8]]))
    trained_loss = float(np.mean([LP.unrolled_loss_grad(
        s_, lam, LP.factor(s_, lam, LP.predicted_shift(net, s_, lam)), 10)[0]
        for s_ in systems[:8]]))
    systems = None  # release the training kernels before the full-size solves
    # ---- PCG on the full system, matrix-free ----
    res = {}
    for name, shift in (("learned", LP.predicted_shift(net, full, lam)), ("block_jacobi", None)):
        inv = LP.factor(full, lam, shift)
        capi.synchronize()
        t0 = time.perf_counter()
        x, it, rel, h = LP.pcg(full, lam, inv, tol=1e-8, max_iter=args.pcg_iters,
                               matrix_free=True)
        capi.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"alpha": x.numpy(), "iters": it, "rel_residual": rel, "seconds": dt,
                     "history_every_10": [float(v) for v in h[::10]]}
        log(f"learned PCG: {name}: {it} iterations, true residual {rel:.3e}, {dt:.2f} s "
            f"({dt / max(it, 1) * 1e3:.1f} ms/iteration)")
    # ---- direct FP64 RFP Cholesky on the full system (the reference solution) ----
    alpha = capi.DeviceArray(full.n)
    t0 = time.perf_counter()
    full.build_rfp()
    capi.rfp_potrf(capi.KF_PREC_F64, full.arf.ptr, full.n, lam)
    alpha.copy_from(full.y)
    capi.rfp_potrs(capi.KF_PREC_F64, full.arf.ptr, full.n, alpha.ptr, 1)
    capi.synchronize()
    t_chol = time.perf_counter() - t0
    a_chol = alpha.numpy()
    full.free_rfp()
    log(f"learned PCG: FP64 RFP Cholesky (kernel build + potrf + potrs) {t_chol:.2f} s")
    # ---- held-out predictions of each solution ----
    test = LP.DeviceSystem(Ct, Z, Et, Ft)

    def predict(a):
        DA = capi.DeviceArray
        ad, aF, aE = DA((nconf, full.ma, full.rep)), DA.from_numpy(a[nconf:]), DA.from_numpy(a[:nconf])
        capi.local_alpha_desc(full.dX.ptr, full.dQ.ptr, full.dN.ptr, nconf, full.ma, full.rep,
                              aF.ptr, ad.ptr)
        Ep, Fp = DA(ntest), DA(ntest * 3 * full.ma)
        capi.local_full_matvec(test.X.ptr, test.dX.ptr, full.X.ptr, ad.ptr, aE.ptr, test.dQ.ptr,
                               full.dQ.ptr, test.dN.ptr, full.dN.ptr, ntest, nconf, full.ma,
                               full.ma, full.rep, SIGMA, Ep.ptr, Fp.ptr)
        capi.synchronize()
        return Ep.numpy(), -Fp.numpy()

    Ec, Fc = predict(a_chol)
    out_sol = {"cholesky_fp64": {"seconds": round(t_chol, 3),
                                 "test_mae_energy": float(np.mean(np.abs(Ec - Et))),
                                 "test_mae_force": float(np.mean(np.abs(Fc - Ft.reshape(-1))))}}
    for name, r in res.items():
        Ep, Fp = predict(r["alpha"])
        out_sol[name + "_pcg"] = {
            "seconds": round(r["seconds"], 3), "iterations": r["iters"],
            "rel_residual": r["rel_residual"],
            "energy_rel_err_vs_fp64": float(np.max(np.abs(Ep - Ec) / np.abs(Ec))),
            "force_rel_err_vs_fp64": float(np.max(np.abs(Fp - Fc)) / np.max(np.abs(Fc))),
            "test_mae_energy": float(np.mean(np.abs(Ep - Et))),
            "test_mae_force": float(np.mean(np.abs(Fp - Ft.reshape(-1)))),
            "residual_history_every_10": r["history_every_10"],
        }
    meets = [k for k, v in out_sol.items() if k == "cholesky_fp64"
             or v["energy_rel_err_vs_fp64"] <= ENERGY_REL_BAR]
    best = min(meets, key=lambda k: out_sol[k]["seconds"])
    return {
        "metric": METRIC, "value": out_sol[best]["seconds"], "unit": "s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(out_sol[best]["seconds"] * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"energy+force fit of {nconf} C9H8O4 conformers (seed 3), "
                                f"order {full.n}, lambda 1e-10: learned block-Jacobi PCG "
                                "(matrix-free full matvec) vs block-Jacobi PCG vs FP64 RFP "
                                "Cholesky; value = fastest solve within 1e-6 relative energy "
                                "error of the FP64 Cholesky solution"),
                   "selected": best, "pcg_tol": 1e-8, "pcg_iteration_cap": args.pcg_iters},
        "solves": out_sol,
        "training": {"systems": args.pcg_systems, "conformers_per_system": args.pcg_subset,
                     "epochs": args.pcg_epochs, "build_s": round(t_build, 2),
                     "train_s": round(t_train, 2), "epoch_mean_unrolled_loss": hist,
                     "unrolled_loss_first8_zero_shift": zero_loss,
                     "unrolled_loss_first8_trained": trained_loss},
    }


def traffic_probe(args, L):
    """Dispatches for the HBM-traffic counters (run under rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE): a 1 GiB device-to-device copy, whose counters calibrate the scale, then
    one FP64 symmetric RFP kernel launch on the n-train energy-fit molecules."""
    from kernelforge_amd import capi, synthetic

    DA = capi.DeviceArray
    a, b = DA(1 << 27), DA(1 << 27)
    a.zero()
    capi.synchronize()
    capi.check(L.kf_memcpy(b.ptr, a.ptr, a.nbytes, 2, None), "kf_memcpy")
    capi.synchronize()
    log(f"traffic probe: copied {a.nbytes} bytes")
    del a, b
    nm = args.n_train
    coords, zs, _ = synthetic.random_molecules(nm, 1)
    C, Q, N = synthetic.pad_batch(coords, zs)
    ma, rs = Q.shape[1], capi.rep_size()
    dC, dQ, dN = DA.from_numpy(C), DA.from_numpy(Q), DA.from_numpy(N)
    X = DA((nm, ma, rs))
    capi.fchl19(dC.ptr, dQ.ptr, dN.ptr, nm, ma, X.ptr)
    arf = DA(nm * (nm + 1) // 2)
    capi.synchronize()
    t0 = time.perf_counter()
    capi.local_kernel_symm_rfp(capi.KF_PREC_F64, X.ptr, dQ.ptr, dN.ptr, nm, ma, rs, SIGMA, arf.ptr)
    capi.synchronize()
    log(f"traffic probe: FP64 symmetric RFP kernel, {nm} molecules, "
        f"{time.perf_counter() - t0:.3f} s")
    return 0


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            return float(json.load(fh)[kernel]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-train", type=int, default=20000)
    ap.add_argument("--n-test", type=int, default=2000)
    ap.add_argument("--cpu-sample", type=int, default=10000,
                    help="molecules in the CPU-baseline sample (0 skips it)")
    ap.add_argument("--ef-conformers", type=int, default=500,
                    help="conformers of the energy+force fit leg (0 skips it)")
    ap.add_argument("--ef-cpu-sample", type=int, default=60,
                    help="conformers in the energy+force CPU-baseline sample (0 skips it)")
    ap.add_argument("--learned-pcg-leg", type=int, default=1,
                    help="1: the default energy workload also runs the 2,000-conformer "
                         "learned-PCG energy+force fit and reports it under learned_pcg_fit")
    ap.add_argument("--workload", choices=["energy", "learned-pcg", "traffic-probe"],
                    default="energy",
                    help="energy: the 20k energy fit (+ the 500-conformer E+F fit); "
                         "learned-pcg: the 2,000-conformer learned-preconditioner E+F fit")
    ap.add_argument("--pcg-conformers", type=int, default=2000)
    ap.add_argument("--pcg-systems", type=int, default=64)
    ap.add_argument("--pcg-subset", type=int, default=250)
    ap.add_argument("--pcg-epochs", type=int, default=6)
    ap.add_argument("--pcg-iters", type=int, default=100)
    ap.add_argument("--precision", choices=["fp64", "mixed", "i8x6", "both", "all"],
                    default="all",
                    help="fp64: plain path; i8x6: FP64-accurate int8-split kernel + FP64 RFP "
                         "Cholesky; mixed: BF16x3 kernel + FP32 factor + PCG; both: fp64 and "
                         "mixed; all: the three, each reported with its error vs the FP64 fit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the kernel "
              "would shard by blocks of molecule pairs, one block range per GPU, with the "
              "Cholesky distributed by block columns)", file=sys.stderr)
        return 2

    from kernelforge_amd import capi, synthetic

    L = capi.lib()
    name = ctypes_name(L)
    log(f"device {name}")
    if args.workload == "traffic-probe":
        return traffic_probe(args, L)
    if args.workload == "learned-pcg":
        out = learned_pcg_leg(args, L)
        out["device"] = name
        print(json.dumps(out), flush=True)
        return 0

    ntr, nte = args.n_train, args.n_test
    t0 = time.perf_counter()
    coords, zs, energies = synthetic.random_molecules(ntr + nte, 1)
    C, Q, N = synthetic.pad_batch(coords, zs)
    cnt, coef = element_baseline(zs[:ntr], energies[:ntr])
    y = energies[:ntr] - cnt @ coef
    log(f"generated {ntr + nte} molecules in {time.perf_counter() - t0:.1f} s")
    ma = Q.shape[1]
    rs = capi.rep_size()

    DA = capi.DeviceArray
    Ctr, Qtr, Ntr = DA.from_numpy(C[:ntr]), DA.from_numpy(Q[:ntr]), DA.from_numpy(N[:ntr])
    ytr = DA.from_numpy(y)
    Xtr = DA((ntr, ma, rs))
    arf = DA(ntr * (ntr + 1) // 2)
    alpha = DA(ntr)
    stream = 0  # the null stream; every call below is ordered on it

    run_fp64 = args.precision in ("fp64", "both", "all")
    run_mixed = args.precision in ("mixed", "both", "all")
    run_i8 = args.precision in ("i8x6", "all")
    arf32 = DA(ntr * (ntr + 1) // 2, np.float32) if run_mixed else None
    lwork = DA(ntr * (ntr + 1) // 2, np.float32) if run_mixed else None
    alpha_m = DA(ntr) if run_mixed else None
    alpha_8 = DA(ntr) if run_i8 else None
    pcg_info = {}

    def step_fp64():
        capi.fchl19(Ctr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, Xtr.ptr, 0, stream)
        capi.local_kernel_symm_rfp(capi.KF_PREC_F64, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                   SIGMA, arf.ptr, stream)
        capi.rfp_potrf(capi.KF_PREC_F64, arf.ptr, ntr, LAMBDA, stream)
        alpha.copy_from(ytr, stream)
        capi.rfp_potrs(capi.KF_PREC_F64, arf.ptr, ntr, alpha.ptr, 1, stream)

    def step_i8():
        # FP64-accurate int8-split kernel (six 7-bit slices per row, exact int32 MFMA
        # products, FP64 epilogue) -> the FP64 RFP Cholesky at the requested lambda
        capi.fchl19(Ctr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, Xtr.ptr, 0, stream)
        capi.local_kernel_symm_rfp(capi.KF_PREC_I8X6, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                   SIGMA, arf.ptr, stream)
        capi.rfp_potrf(capi.KF_PREC_F64, arf.ptr, ntr, LAMBDA, stream)
        alpha_8.copy_from(ytr, stream)
        capi.rfp_potrs(capi.KF_PREC_F64, arf.ptr, ntr, alpha_8.ptr, 1, stream)

    def step_mixed():
        # BF16x3 kernel -> FP32 factor -> FP64-residual PCG at the requested lambda; when
        # PCG stalls (K32 + lambda I is indefinite in FP32 storage at this size) the fit
        # falls back to the FP64 kernel and RFP Cholesky, inside the timed step.
        capi.fchl19(Ctr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, Xtr.ptr, 0, stream)
        capi.local_kernel_symm_rfp(capi.KF_PREC_BF16X3, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                   SIGMA, arf32.ptr, stream)
        try:
            it, rel, mu = capi.krr_solve_mixed(arf32.ptr, lwork.ptr, ntr, LAMBDA, ytr.ptr,
                                               alpha_m.ptr, tol=1e-10, max_iter=300,
                                               stream=stream)
            pcg_info.update(iters=it, rel_residual=rel, factor_shift=mu, fp64_fallback=False)
        except capi.NoConvergence as exc:
            pcg_info.update(pcg_error=str(exc), fp64_fallback=True)
            capi.local_kernel_symm_rfp(capi.KF_PREC_F64, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                       SIGMA, arf.ptr, stream)
            capi.rfp_potrf(capi.KF_PREC_F64, arf.ptr, ntr, LAMBDA, stream)
            alpha_m.copy_from(ytr, stream)
            capi.rfp_potrs(capi.KF_PREC_F64, arf.ptr, ntr, alpha_m.ptr, 1, stream)

    def timed(step, label):
        for w in range(args.warmup):
            step()
            capi.synchronize(stream)
            log(f"{label}: warmup {w + 1}/{args.warmup} done")
        L.kf_timer_reset()
        L.kf_timer_enable(1)
        capi.synchronize(stream)
        t0 = time.perf_counter()
        for s_ in range(args.steps):
            step()
            capi.synchronize(stream)
            log(f"{label}: step {s_ + 1}/{args.steps} done")
        capi.synchronize(stream)
        el = time.perf_counter() - t0
        L.kf_timer_enable(0)
        tm = {k: capi.timer_read(v) for k, v in
              (("fchl19", capi.TM_FCHL19), ("kernel", capi.TM_LOCAL_SYMM),
               ("potrf", capi.TM_POTRF), ("potrs", capi.TM_POTRS), ("pcg", capi.TM_MATVEC))}
        return el / args.steps, {k: (ms / c if c else None) for k, (ms, c) in tm.items()}

    res = {}
    if run_fp64:
        res["fp64"] = timed(step_fp64, "fp64")
    if run_i8:
        res["i8x6"] = timed(step_i8, "i8x6")
    if run_mixed:
        try:
            res["mixed"] = timed(step_mixed, "mixed")
        except RuntimeError as exc:  # reported, the plain path still stands
            log(f"mixed path failed: {exc}")
            if not res:
                raise
    rows = label_rows(Q[:ntr], N[:ntr])
    rows_sq = sum(v * v for v in rows.values())
    kernel_flop = rs * rows_sq  # symmetric half of 2 * R_e^2 * rep, summed over labels
    # the FP64 kernel multiplies only the 16-column windows both rows of a tile pair use
    win_flop = window_gram_flop(Xtr.numpy(), Q[:ntr], N[:ntr])

    # test predictions and errors (after the timed regions; not part of a step)
    Xte = DA((nte, ma, rs))
    Cte, Qte, Nte = DA.from_numpy(C[ntr:]), DA.from_numpy(Q[ntr:]), DA.from_numpy(N[ntr:])
    capi.fchl19(Cte.ptr, Qte.ptr, Nte.ptr, nte, ma, Xte.ptr, 0, stream)
    cnt_te = np.array([[np.sum(z == e) for e in ELEMENTS] for z in zs[ntr:]], dtype=np.float64)
    Kte = DA((nte, ntr))
    capi.local_kernel(capi.KF_PREC_F64, Xte.ptr, Xtr.ptr, Qte.ptr, Qtr.ptr, Nte.ptr, Ntr.ptr, nte,
                      ntr, ma, ma, rs, SIGMA, Kte.ptr, stream)
    capi.synchronize(stream)
    Kte_h = Kte.numpy()
    accuracy = {}
    acc_i8 = {}
    mae = None
    if "fp64" in res:
        a64 = alpha.numpy()
        pred64 = Kte_h @ a64 + cnt_te @ coef
        mae = float(np.mean(np.abs(pred64 - energies[ntr:])))
    if "i8x6" in res:
        Kte8 = DA((nte, ntr))
        capi.local_kernel(capi.KF_PREC_I8X6, Xte.ptr, Xtr.ptr, Qte.ptr, Qtr.ptr, Nte.ptr,
                          Ntr.ptr, nte, ntr, ma, ma, rs, SIGMA, Kte8.ptr, stream)
        capi.synchronize(stream)
        Kte8_h = Kte8.numpy()
        a8 = alpha_8.numpy()
        pred8 = Kte8_h @ a8 + cnt_te @ coef
        acc_i8.update(test_mae_ha=float(np.mean(np.abs(pred8 - energies[ntr:]))))
        if "fp64" in res:
            big = np.abs(Kte_h) > 1e-200
            acc_i8.update(
                energy_rel_err_max=float(np.max(np.abs(pred8 - pred64) / np.abs(pred64))),
                energy_abs_err_max_ha=float(np.max(np.abs(pred8 - pred64))),
                alpha_rel_err_inf=float(np.max(np.abs(a8 - a64)) / np.max(np.abs(a64))),
                kernel_rel_err_test_block=float(np.max(np.abs(Kte8_h - Kte_h)) /
                                                np.max(np.abs(Kte_h))),
                kernel_entry_rel_err_max_test_block=float(
                    np.max(np.abs(Kte8_h - Kte_h)[big] / np.abs(Kte_h)[big])),
                speedup_vs_fp64=res["fp64"][0] / res["i8x6"][0])
            acc_i8["meets_1e-6_bar"] = bool(acc_i8["energy_rel_err_max"] <= ENERGY_REL_BAR)
    if "mixed" in res:
        Kte32 = DA((nte, ntr), np.float32)
        capi.local_kernel(capi.KF_PREC_BF16X3, Xte.ptr, Xtr.ptr, Qte.ptr, Qtr.ptr, Nte.ptr,
                          Ntr.ptr, nte, ntr, ma, ma, rs, SIGMA, Kte32.ptr, stream)
        capi.synchronize(stream)
        am = alpha_m.numpy()
        Kpred = Kte_h if pcg_info.get("fp64_fallback") else Kte32.numpy().astype(np.float64)
        predm = Kpred @ am + cnt_te @ coef
        accuracy.update(pcg_iters=pcg_info.get("iters"),
                        pcg_rel_residual=pcg_info.get("rel_residual"),
                        fp32_factor_shift=pcg_info.get("factor_shift"),
                        fp64_fallback=pcg_info.get("fp64_fallback"),
                        pcg_error=pcg_info.get("pcg_error"),
                        test_mae_mixed_ha=float(np.mean(np.abs(predm - energies[ntr:]))))
        if "fp64" in res:
            accuracy.update(
                energy_rel_err_max=float(np.max(np.abs(predm - pred64) / np.abs(pred64))),
                energy_abs_err_max_ha=float(np.max(np.abs(predm - pred64))),
                alpha_rel_err_inf=float(np.max(np.abs(am - a64)) / np.max(np.abs(a64))),
                kernel_rel_err_test_block=float(np.max(np.abs(Kte32.numpy() - Kte_h)) /
                                                np.max(np.abs(Kte_h))),
                speedup_vs_fp64=res["fp64"][0] / res["mixed"][0])
    # The metric is the fit time at <= 1e-6 relative energy error vs FP64: a reduced-
    # precision path is the head only when its test energies meet that bar and it is
    # faster than every other path that does, by more than the run-to-run spread.
    head = "fp64" if "fp64" in res else ("i8x6" if "i8x6" in res else "mixed")
    if "fp64" in res:
        for cand, acc in (("i8x6", acc_i8), ("mixed", accuracy)):
            if (cand in res and acc.get("energy_rel_err_max", 1.0) <= ENERGY_REL_BAR
                    and res[cand][0] < HEAD_MARGIN * res[head][0]):
                head = cand
    if accuracy:
        accuracy["meets_1e-6_bar"] = bool(accuracy.get("energy_rel_err_max", 1.0) <= ENERGY_REL_BAR)
    sec_per_fit, avg_ms = res[head]
    if head == "mixed":  # three BF16 products per contraction term
        roof_flop, peak, kname = 3 * kernel_flop, PEAK_BF16_MFMA_TFLOPS, "local_bf16x3_kernel<SYMM_RFP>"
    elif head == "i8x6":  # 21 int8 slice products per contraction term
        roof_flop, peak, kname = 21 * win_flop, PEAK_INT8_MFMA_TOPS, "local_scalar_f64_kernel<0, 6>"
    else:
        roof_flop, peak, kname = win_flop, PEAK_F64_MFMA_TFLOPS, "local_scalar_f64_kernel<SYMM_RFP>"
    achieved = roof_flop / (avg_ms["kernel"] * 1e-3) / 1e12
    log(f"{head} fit {sec_per_fit:.3f} s/step, kernel {avg_ms['kernel']:.1f} ms "
        f"({achieved:.1f} TFLOP/s), breakdown {avg_ms}")
    if accuracy:
        log(f"accuracy {accuracy}")

    cpu = None
    if args.cpu_sample > 0:
        log(f"CPU baseline on {args.cpu_sample} molecules ...")
        cpu = cpu_baseline(C[:ntr], Q[:ntr], N[:ntr], y, min(args.cpu_sample, ntr), rows_sq, ntr)
        log(f"CPU baseline {cpu['value']} s (scaled)")

    out = {
        "metric": METRIC,
        "value": round(sec_per_fit, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_fit * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "bf16x3/f32/f64", "i8x6": "i8x6/i32/f64"}.get(head, "f64"),
        "data": "synthetic",
        "config": {
            "workload": (f"FCHL19 local-kernel KRR energy fit, {ntr} synthetic molecules "
                         f"(seed 1): descriptors + symmetric RFP kernel + solve; "
                         + {"mixed": "mixed: BF16x3 MFMA kernel (FP32 RFP), FP32 RFP Cholesky, "
                                     "FP64-residual PCG to 1e-10",
                            "i8x6": "FP64-accurate int8-split kernel (six 7-bit slices, int32 "
                                    "MFMA) and FP64 RFP Cholesky"}.get(
                                head, "plain FP64 kernel and RFP Cholesky")),
            "n_train": ntr, "n_test": nte, "max_atoms": ma, "rep": rs, "sigma": SIGMA,
            "lambda": LAMBDA, "precision": head,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TOP/s" if head == "i8x6" else "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": measured_traffic(kname) if head == "fp64" else None,
            "kernel": kname,
            "flop_per_launch": roof_flop,
            "flop_per_launch_dense": kernel_flop,
            "achieved_dense_equivalent": round(kernel_flop / (avg_ms["kernel"] * 1e-3) / 1e12, 3),
            "frac_dense_equivalent": round(kernel_flop / (avg_ms["kernel"] * 1e-3) / 1e12 / peak, 4),
            "flop_note": ("achieved/frac count the Gram FLOP of the 16-column descriptor "
                          "windows both rows of a pair use (the work of an exact "
                          "zero-window-skipping kernel); the dense-equivalent figures count "
                          "rep * sum_e R_e^2 (every column), which can exceed the peak because "
                          "the kernel skips windows that are exactly zero"),
            "avg_launch_ms": round(avg_ms["kernel"], 3),
        },
        "cpu_baseline": cpu,
        "breakdown_ms": {k: (round(v, 3) if v is not None else None) for k, v in avg_ms.items()},
        "device": name,
    }
    if "fp64" in res:
        out["plain_fp64"] = {"s_per_fit": round(res["fp64"][0], 4),
                             "breakdown_ms": {k: (round(v, 3) if v is not None else None)
                                              for k, v in res["fp64"][1].items()},
                             "kernel_tflops": round(win_flop / (res["fp64"][1]["kernel"] * 1e-3)
                                                    / 1e12, 3),
                             "kernel_tflops_dense_equivalent": round(
                                 kernel_flop / (res["fp64"][1]["kernel"] * 1e-3) / 1e12, 3),
                             "test_mae_ha": round(mae, 5)}
    if "mixed" in res:
        out["mixed"] = {"s_per_fit": round(res["mixed"][0], 4),
                        "breakdown_ms": {k: (round(v, 3) if v is not None else None)
                                         for k, v in res["mixed"][1].items()},
                        "kernel_tflops_bf16x3": round(3 * kernel_flop /
                                                      (res["mixed"][1]["kernel"] * 1e-3) / 1e12, 3)}
    if accuracy:
        out["mixed_accuracy_vs_fp64"] = accuracy
    if "i8x6" in res:
        k8 = res["i8x6"][1]["kernel"]
        out["int8_split_fit"] = {
            "s_per_fit": round(res["i8x6"][0], 4),
            "breakdown_ms": {k: (round(v, 3) if v is not None else None)
                             for k, v in res["i8x6"][1].items()},
            "kernel_int8_tops": round(21 * win_flop / (k8 * 1e-3) / 1e12, 3),
            "kernel_frac_int8_peak": round(21 * win_flop / (k8 * 1e-3) / 1e12
                                           / PEAK_INT8_MFMA_TOPS, 4),
            "kernel_fp64_equivalent_tflops": round(win_flop / (k8 * 1e-3) / 1e12, 3),
            "accuracy_vs_fp64": acc_i8}
    if args.ef_conformers > 0:
        out["energy_force_fit"] = ef_leg(args, L)
    if args.learned_pcg_leg:
        pcg = learned_pcg_leg(args, L)
        out["learned_pcg_fit"] = {k: pcg[k] for k in ("value", "unit", "config", "solves",
                                                       "training")}
    print(json.dumps(out), flush=True)
    return 0


def ctypes_name(L):
    import ctypes

    buf = ctypes.create_string_buffer(256)
    L.kf_device_name(buf, 256)
    return buf.value.decode(errors="replace")


if __name__ == "__main__":
    sys.exit(main())
