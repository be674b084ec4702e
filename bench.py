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

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--cpu-sample 3000]
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
SIGMA, LAMBDA = 2.0, 1e-8
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
        "sample": (f"oracle (C/OpenMP + scipy OpenBLAS) on the first {sample} of the 20,000 "
                   f"training molecules: descriptors {t1 - t0:.2f} s, FP64 symmetric RFP "
                   f"kernel {t2 - t1:.2f} s, dpftrf+dpftrs {t3 - t2:.3f} s; scaled to 20,000 "
                   f"by n (descriptors), sum_e R_e^2 (kernel) and n^3 (Cholesky): "
                   f"{s_desc:.1f} + {s_kern:.1f} + {s_chol:.1f} s"),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-train", type=int, default=20000)
    ap.add_argument("--n-test", type=int, default=2000)
    ap.add_argument("--cpu-sample", type=int, default=3000,
                    help="molecules in the CPU-baseline sample (0 skips it)")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both",
                    help="fp64: plain path; mixed: BF16x3 kernel + FP32 factor + PCG; both: "
                         "time both, report the mixed fit with its error vs the FP64 fit")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path would shard by molecule-pair tiles, see DESIGN.md)", file=sys.stderr)
        return 2

    from kernelforge_amd import capi, synthetic

    L = capi.lib()
    name = ctypes_name(L)
    log(f"device {name}")

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

    arf32 = DA(ntr * (ntr + 1) // 2, np.float32) if args.precision != "fp64" else None
    lwork = DA(ntr * (ntr + 1) // 2, np.float32) if args.precision != "fp64" else None
    alpha_m = DA(ntr) if args.precision != "fp64" else None
    pcg_info = {}

    def step_fp64():
        capi.fchl19(Ctr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, Xtr.ptr, 0, stream)
        capi.local_kernel_symm_rfp(capi.KF_PREC_F64, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                   SIGMA, arf.ptr, stream)
        capi.rfp_potrf(capi.KF_PREC_F64, arf.ptr, ntr, LAMBDA, stream)
        alpha.copy_from(ytr, stream)
        capi.rfp_potrs(capi.KF_PREC_F64, arf.ptr, ntr, alpha.ptr, 1, stream)

    def step_mixed():
        capi.fchl19(Ctr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, Xtr.ptr, 0, stream)
        capi.local_kernel_symm_rfp(capi.KF_PREC_BF16X3, Xtr.ptr, Qtr.ptr, Ntr.ptr, ntr, ma, rs,
                                   SIGMA, arf32.ptr, stream)
        it, rel, mu = capi.krr_solve_mixed(arf32.ptr, lwork.ptr, ntr, LAMBDA, ytr.ptr,
                                           alpha_m.ptr, tol=1e-10, max_iter=300, stream=stream)
        pcg_info.update(iters=it, rel_residual=rel, lambda_solved=mu)

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
    if args.precision in ("fp64", "both"):
        res["fp64"] = timed(step_fp64, "fp64")
    if args.precision in ("mixed", "both"):
        try:
            res["mixed"] = timed(step_mixed, "mixed")
        except RuntimeError as exc:  # reported, the plain path still stands
            log(f"mixed path failed: {exc}")
            if not res:
                raise
    head = "mixed" if "mixed" in res else "fp64"
    sec_per_fit, avg_ms = res[head]

    rows = label_rows(Q[:ntr], N[:ntr])
    rows_sq = sum(v * v for v in rows.values())
    kernel_flop = rs * rows_sq  # symmetric half of 2 * R_e^2 * rep, summed over labels
    if head == "mixed":  # three BF16 products per contraction term
        roof_flop, peak, kname = 3 * kernel_flop, PEAK_BF16_MFMA_TFLOPS, "local_bf16x3_kernel<SYMM_RFP>"
    else:
        roof_flop, peak, kname = kernel_flop, PEAK_F64_MFMA_TFLOPS, "local_scalar_f64_kernel<SYMM_RFP>"
    achieved = roof_flop / (avg_ms["kernel"] * 1e-3) / 1e12

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
    mae = None
    if "fp64" in res:
        a64 = alpha.numpy()
        pred64 = Kte_h @ a64 + cnt_te @ coef
        mae = float(np.mean(np.abs(pred64 - energies[ntr:])))
    if "mixed" in res:
        Kte32 = DA((nte, ntr), np.float32)
        capi.local_kernel(capi.KF_PREC_BF16X3, Xte.ptr, Xtr.ptr, Qte.ptr, Qtr.ptr, Nte.ptr,
                          Ntr.ptr, nte, ntr, ma, ma, rs, SIGMA, Kte32.ptr, stream)
        capi.synchronize(stream)
        am = alpha_m.numpy()
        predm = Kte32.numpy().astype(np.float64) @ am + cnt_te @ coef
        accuracy.update(pcg_iters=pcg_info.get("iters"),
                        pcg_rel_residual=pcg_info.get("rel_residual"),
                        lambda_solved=pcg_info.get("lambda_solved"),
                        test_mae_mixed_ha=float(np.mean(np.abs(predm - energies[ntr:]))))
        if "fp64" in res:
            accuracy.update(
                energy_rel_err_max=float(np.max(np.abs(predm - pred64) / np.abs(pred64))),
                energy_abs_err_max_ha=float(np.max(np.abs(predm - pred64))),
                alpha_rel_err_inf=float(np.max(np.abs(am - a64)) / np.max(np.abs(a64))),
                kernel_rel_err_test_block=float(np.max(np.abs(Kte32.numpy() - Kte_h)) /
                                                np.max(np.abs(Kte_h))),
                speedup_vs_fp64=res["fp64"][0] / res["mixed"][0])
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
        "dtype": "bf16x3/f32/f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"FCHL19 local-kernel KRR energy fit, {ntr} synthetic molecules "
                         f"(seed 1): descriptors + symmetric RFP kernel + solve; "
                         + ("mixed: BF16x3 MFMA kernel (FP32 RFP), FP32 RFP Cholesky, FP64-"
                            "residual PCG to 1e-10" if head == "mixed" else
                            "plain FP64 kernel and RFP Cholesky")),
            "n_train": ntr, "n_test": nte, "max_atoms": ma, "rep": rs, "sigma": SIGMA,
            "lambda": LAMBDA, "precision": head,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": None,
            "kernel": kname,
            "flop_per_launch": roof_flop,
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
                             "kernel_tflops": round(kernel_flop / (res["fp64"][1]["kernel"] * 1e-3)
                                                    / 1e12, 3),
                             "test_mae_ha": round(mae, 5)}
    if accuracy:
        out["mixed_accuracy_vs_fp64"] = accuracy
    print(json.dumps(out), flush=True)
    return 0


def ctypes_name(L):
    import ctypes

    buf = ctypes.create_string_buffer(256)
    L.kf_device_name(buf, 256)
    return buf.value.decode(errors="replace")


if __name__ == "__main__":
    sys.exit(main())
