"""Benchmark of the supernodal Cholesky numeric phase on one MI355X.

One step = one numeric factorization of the 1.57M-equation hex8 block (80^3 cells on the unit
cube, z = 0 clamped, E = 210 GPa, nu = 0.3) from matrix values already in HBM, followed by a
solve refined in FP64 to normwise backward error <= 1e-12 (berr = ||b - Kx||_inf /
(||K||_inf ||x||_inf + ||b||_inf)). The symbolic phase is done once, outside the timed region,
as the reference reuses one Symbolic for many factorizations.

Two legs are timed the same way: the plain method (FP64 factor) and the chosen precision form
(mixed factor + refinement). `value` is the chosen form's throughput: the factor's algorithmic
flops F = sum_s sum_{c<w_s} (m_s - c)^2 (the same for every precision) divided by the step's
time (factor + refined solve). The FP64 leg is reported beside it, with the speedup.

`roofline` is the dominant kernel of the chosen form's factorization, timed with HIP events
around its launches on the factor's stream during the timed steps; `cpu_baseline` is the
oracle (the C restatement of the reference's CPU factorization) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 80] [--form mixed|fp32|fp64]
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

METRIC = "factor+solve s & TFLOP/s, 1.6M-dof hex8: mixed-prec+IR vs FP64 at berr≤1e-12"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # TFLOP/s, dense MFMA, MI355X
HBM_PEAK = 8000.0  # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_leg(CH, sym, dvals, A, b_ptr, x_ptr, n, precision, steps, warmup, profile, bf16_terms=1):
    import torch

    f = CH.Factor(sym, precision=precision, device_values=dvals, profile=profile, bf16_terms=bf16_terms)
    rep = f.solve_refined_device(A, b_ptr, x_ptr)
    for _ in range(max(warmup - 1, 0)):
        f.refactor_device(dvals)
        rep = f.solve_refined_device(A, b_ptr, x_ptr)
    torch.cuda.synchronize()
    t_factor, t_solve, per_kind = [], [], {k: [0.0, 0.0] for k in profile}
    t0 = time.perf_counter()
    for _ in range(steps):
        f.refactor_device(dvals)
        t_factor.append(f.seconds())
        for k in profile:
            pr = f.profile(k)
            per_kind[k][0] += pr["seconds"]
            per_kind[k][1] += pr["launches"]
        rep = f.solve_refined_device(A, b_ptr, x_ptr)
        t_solve.append(rep["seconds"])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / max(steps, 1)
    kinds = {}
    for k in profile:
        pr = f.profile(k)
        kinds[k] = dict(seconds=per_kind[k][0] / steps, launches=per_kind[k][1] / steps, flops=pr["flops"],
                        entries=pr["entries"])
    out = dict(wall_s=wall, factor_s=float(np.mean(t_factor)), solve_s=float(np.mean(t_solve)),
               iterations=rep["iterations"], berr=rep["berr"], converged=rep["converged"], kinds=kinds)
    del f
    return out


def cpu_baseline(cells: int, threads: int):
    from fieldcraft_amd import problems as P
    from oracle import oracle as O

    p = P.cube_hex8(cells)
    up = p.upper()
    O.set_threads(threads)
    t0 = time.perf_counter()
    f = O.Factor(p.n, *up)
    t1 = time.perf_counter()
    x = f.solve(p.rhs)
    t2 = time.perf_counter()
    s = f.symbolic()
    w = np.diff(s["first"]).astype(np.float64)
    m = np.diff(s["rptr"]).astype(np.float64)
    flops = float(np.sum(w * m * m - m * w * (w - 1) + (w - 1) * w * (2 * w - 1) / 6))
    r = p.mul(x) - p.rhs
    berr = float(np.abs(r).max() / (p.norm_inf() * np.abs(x).max() + np.abs(p.rhs).max()))
    return dict(value=flops / (t2 - t0) / 1e12, unit="TFLOP/s", cores=threads, kind="port",
                sample=f"oracle FP64 factor+solve of the hex8 {cells}^3 unit cube ({p.n} equations, "
                       f"{flops:.3e} flop): factor {t1 - t0:.2f} s, solve {t2 - t1:.3f} s, berr {berr:.1e}, "
                       f"{threads} OpenMP threads on {os.cpu_count()} host cpus")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=80)
    ap.add_argument("--form", default="mixed", choices=["mixed", "fp32", "fp64"])
    ap.add_argument("--bf16-terms", type=int, default=3)
    ap.add_argument("--cpu-cells", type=int, default=28)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count())
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU: the factorization of one matrix is not split across GPUs here "
              "(see the design notes on sharding); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from fieldcraft_amd import cholesky as CH
    from fieldcraft_amd import problems as P

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    t0 = time.perf_counter()
    p = P.cube_hex8(args.cells)
    up = p.upper()
    t1 = time.perf_counter()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    t2 = time.perf_counter()
    info = sym.info()
    log(f"bench: hex8 {args.cells}^3 n={p.n} nnz={p.nnz} generated in {t1 - t0:.1f}s, analysed in {t2 - t1:.1f}s; "
        f"fronts {info['fronts']} levels {info['levels']} launches {info['launches']} F={info['flops']:.3e}")
    dev = torch.device("cuda:0")
    dvals = torch.from_numpy(up[2]).to(dev)
    b = torch.from_numpy(p.rhs).to(dev)
    x = torch.zeros_like(b)
    A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
    torch.cuda.synchronize()
    profile = ["syrk", "update", "extend_l", "extend_u", "trsm", "potrf", "reduce"]
    legs = {}
    if not args.no_fp64:
        log("bench: fp64 leg")
        legs["fp64"] = run_leg(CH, sym, dvals.data_ptr(), A, b.data_ptr(), x.data_ptr(), p.n, "fp64", args.steps,
                               args.warmup, profile)
        x64 = x.cpu().numpy().copy()
        log(f"bench: fp64 {legs['fp64']}")
    form = args.form
    if form != "fp64":
        log(f"bench: {form} leg")
        legs[form] = run_leg(CH, sym, dvals.data_ptr(), A, b.data_ptr(), x.data_ptr(), p.n, form, args.steps,
                             args.warmup, profile, bf16_terms=args.bf16_terms)
        xf = x.cpu().numpy()
        if "fp64" in legs:
            legs[form]["rel_err_vs_fp64"] = float(np.linalg.norm(xf - x64) / np.linalg.norm(x64))
        log(f"bench: {form} {legs[form]}")
    head = legs[form]
    F = info["flops"]
    t_step = head["factor_s"] + head["solve_s"]
    # the dominant kernel of the headline leg's factorization
    dom = max(head["kinds"], key=lambda k: head["kinds"][k]["seconds"])
    dk = head["kinds"][dom]
    if dom in ("extend_l", "extend_u"):
        es = 8 if form == "fp64" else 4
        per_launch_bytes = dk["entries"] * (3 * es + 4) / max(dk["launches"], 1)
        achieved = per_launch_bytes / (dk["seconds"] / max(dk["launches"], 1)) / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK, unit="GB/s", frac=achieved / HBM_PEAK,
                    traffic=None, kernel=f"k_extend_add ({dom})", launches=dk["launches"])
    else:
        bf16 = form == "mixed" and dom in ("syrk", "update")
        peak = PEAK["fp64"] if form == "fp64" else (PEAK["bf16"] if bf16 else PEAK["fp32"])
        achieved = dk["flops"] / dk["seconds"] / 1e12
        kname = (f"k_gemm_bf16<{args.bf16_terms}> ({dom})" if bf16 else
                 f"k_gemm<{'double' if form == 'fp64' else 'float'}> ({dom})" if dom in ("syrk", "update", "trsm") else
                 f"{dom}<{'double' if form == 'fp64' else 'float'}>")
        roof = dict(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak, traffic=None,
                    kernel=kname, launches=dk["launches"], mfma_products_per_flop=args.bf16_terms if bf16 else 1,
                    flops_per_launch=dk["flops"] / max(dk["launches"], 1),
                    avg_launch_s=dk["seconds"] / max(dk["launches"], 1))
    line = dict(metric=METRIC, value=F / t_step / 1e12, unit="TFLOP/s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=t_step * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype={"mixed": "bf16/fp32 factor, fp64 refinement", "fp32": "fp32 factor, fp64 refinement",
                       "fp64": "fp64"}[form],
                data="synthetic",
                config=dict(workload=f"hex8 {args.cells}^3 unit cube, z=0 clamped, E=210GPa nu=0.3, pressure + "
                                     f"seeded N(0,1) body force; {p.n} equations, nnz {p.nnz}",
                            equations=p.n, factor_entries=info["entries"], flops=F, form=form,
                            target_berr=1e-12),
                roofline=roof)
    if "fp64" in legs:
        f64 = legs["fp64"]
        t64 = f64["factor_s"] + f64["solve_s"]
        line["fp64"] = dict(factor_s=f64["factor_s"], solve_s=f64["solve_s"], tflops_factor=F / f64["factor_s"] / 1e12,
                            frac_fp64_peak=F / f64["factor_s"] / 1e12 / PEAK["fp64"], berr=f64["berr"],
                            iterations=f64["iterations"])
        line["speedup_vs_fp64"] = t64 / t_step
    line[form] = dict(factor_s=head["factor_s"], solve_s=head["solve_s"], iterations=head["iterations"],
                      berr=head["berr"], converged=head["converged"], rel_err_vs_fp64=head.get("rel_err_vs_fp64"))
    line["kernels"] = {leg: {k: dict(seconds=v["seconds"], launches=v["launches"],
                                     tflops=(v["flops"] / v["seconds"] / 1e12) if v["seconds"] and v["flops"] else None)
                             for k, v in legs[leg]["kinds"].items()} for leg in legs}
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(args.cpu_cells, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
