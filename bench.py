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
                      [--contrast] [--workload cube|batch|shift|assemble|modes|ldlt]

--contrast adds the harder variant of the same pattern (E in {210 GPa, 21 MPa} per element,
nu = 0.49) with the same symbolic analysis. --workload batch runs 512 independent hex8/tet10
models (2k-60k equations) as one block-diagonal factorization in FP32 with FP64 refinement of
64 load cases per model; --workload shift factors K - sigma M of a tet10 steel block
(sigma = (2 pi 50 Hz)^2) in the mixed form and times 40 refined block solves of 16 vectors;
--workload assemble times the step before the factor on the device: K_ff of the hex8 cube and
K - sigma M of the tet10 block assembled into the factor's input, the reference's values bit for
bit; --workload modes runs the modal solve on the device: K - sigma M of the tet10 block assembled,
factored in FP64 and its 20 lowest modes above sigma found by the shifted block Lanczos method;
--workload ldlt factors an indefinite K - sigma M (sigma between the tet10 block's 5th and 6th
modes) by LDL^T with pivoting and checks its inertia against the modes. Those five print their
own metric lines.
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


def run_leg(CH, sym, dvals, A, b_ptr, x_ptr, n, precision, steps, warmup, profile, bf16_terms=0, breakdown=()):
    import torch

    # the solves replay from hipGraphs (use_graph bit 1); the factorization's launches stay eager
    # so that every launch of the profiled kinds is timed with events
    f = CH.Factor(sym, precision=precision, device_values=dvals, profile=profile, bf16_terms=bf16_terms,
                  use_graph=2)
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
    if breakdown:
        # every kind's time from one more factorization, outside the timed steps: timing each
        # launch with events costs the step a few per cent, so the timed steps time only the
        # kinds the roofline needs
        torch.cuda.empty_cache()
        g = CH.Factor(sym, precision=precision, device_values=dvals, profile=breakdown, bf16_terms=bf16_terms)
        g.refactor_device(dvals)
        out["breakdown"] = {k: g.profile(k) for k in breakdown}
        del g
    return out


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload`, as recorded in a committed rocprofv3 --pmc
    summary under profiles/ (a line `traffic: <kernel> | <workload> | <bytes> bytes per launch`,
    FETCH_SIZE + WRITE_SIZE scaled by the plain-copy calibration); None when no summary covers it."""
    import glob
    import re

    pat = re.compile(r"^traffic: (.+?) \| (.+?) \| ([0-9.eE+]+) bytes per launch")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for ln in open(path):
            m = pat.match(ln.strip())
            if m and m.group(1) == kernel and m.group(2) == workload:
                return float(m.group(3))
    return None


def calibrate_copy(gib: int = 1):
    """A plain elementwise copy of a known size (torch.neg: reads and writes `gib` GiB once each),
    run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to scale those counters to bytes."""
    import torch

    n = gib * (1 << 30) // 4
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps(dict(calibration="torch.neg fp32", bytes_read=n * 4, bytes_written=n * 4, launches=3)))


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
    ap.add_argument("--contrast", action="store_true")
    ap.add_argument("--workload", default="cube", choices=["cube", "batch", "shift", "assemble", "modes", "ldlt",
                                                                 "zldlt"])
    ap.add_argument("--ldlt-cells", type=int, default=12)
    ap.add_argument("--modes", type=int, default=20)
    ap.add_argument("--models", type=int, default=512)
    ap.add_argument("--tet-cells", type=int, default=50)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default schedule (no autotuning)")
    ap.add_argument("--schedule", default="", help="run this schedule instead of tuning, e.g. "
                                                   "slab=2,split_target=2048,split_side_fp32=384")
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
    if args.calibrate_copy:
        return calibrate_copy()
    if args.workload == "batch":
        return batch_workload(args, CH, P)
    if args.workload == "shift":
        return shift_workload(args, CH, P)
    if args.workload == "assemble":
        return assemble_workload(args, P)
    if args.workload == "modes":
        return modes_workload(args, CH, P)
    if args.workload == "ldlt":
        return ldlt_workload(args, CH, P)
    if args.workload == "zldlt":
        return zldlt_workload(args, CH, P)
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
    breakdown = ["syrk", "update", "extend_l", "extend_u", "trsm", "potrf", "reduce", "small"]
    # timed with events in the timed steps: the kinds of the headline's dominant kernel only
    profile = ["syrk", "update"] if args.form == "mixed" else ["syrk", "update", "trsm"]
    legs = {}
    tuned = {}

    def tune(leg):
        # the autotuner (fc_tune) times candidate schedules on this matrix, outside the timed
        # steps, and installs the fastest for the leg's factorizations
        if args.schedule:
            sym.set_schedule(**{k: int(v) for k, v in (kv.split("=") for kv in args.schedule.split(","))})
        if args.no_tune or args.schedule:
            tuned[leg] = dict(chosen=sym.schedule(), tuned=False)
            return
        t0 = time.perf_counter()
        rep = sym.tune(dvals.data_ptr(), precision=leg, repeats=1)
        tuned[leg] = dict(chosen=rep["chosen"], tuned=True, seconds_start=rep["seconds_start"],
                          seconds_chosen=rep["seconds_chosen"], candidates=len(rep["tried"]),
                          tune_s=time.perf_counter() - t0)
        log(f"bench: {leg} schedule {tuned[leg]}")

    if not args.no_fp64:
        log("bench: fp64 leg")
        tune("fp64")
        legs["fp64"] = run_leg(CH, sym, dvals.data_ptr(), A, b.data_ptr(), x.data_ptr(), p.n, "fp64", args.steps,
                               args.warmup, [], breakdown=breakdown)
        x64 = x.cpu().numpy().copy()
        log(f"bench: fp64 {legs['fp64']}")
    form = args.form
    if form != "fp64":
        log(f"bench: {form} leg")
        tune(form)
        legs[form] = run_leg(CH, sym, dvals.data_ptr(), A, b.data_ptr(), x.data_ptr(), p.n, form, args.steps,
                             args.warmup, profile, bf16_terms=args.bf16_terms, breakdown=breakdown)
        xf = x.cpu().numpy()
        if "fp64" in legs:
            legs[form]["rel_err_vs_fp64"] = float(np.linalg.norm(xf - x64) / np.linalg.norm(x64))
        log(f"bench: {form} {legs[form]}")
    head = legs[form]
    F = info["flops"]
    # a step's time is the wall clock around the timed steps (synchronized on both sides), which
    # includes the host's launch work; the per-kernel figures below come from HIP events
    t_step = head["wall_s"]
    # the dominant kernel function of the headline leg's factorization (kinds grouped by the
    # kernel that runs them, so the figure matches a rocprofv3 per-kernel summary)
    el = "double" if form == "fp64" else "float"
    funcs = {}
    for kind, v in head["kinds"].items():
        if kind in ("update", "syrk") and form == "mixed":
            name, bound, peak = f"k_gemm_b16<{args.bf16_terms}>", "mfma", PEAK["bf16"]
        elif kind in ("update", "syrk", "trsm"):
            name, bound, peak = f"k_gemm<{el}>", "mfma", PEAK["fp64" if form == "fp64" else "fp32"]
        elif kind in ("extend_l", "extend_u"):
            name, bound, peak = f"k_extend_add<{el}>", "hbm", HBM_PEAK
        elif kind == "potrf":
            name, bound, peak = f"k_potrf_blk<{el}>", "latency", None
        elif kind == "small":
            name, bound, peak = f"k_small_fronts<{el}>", "latency", None
        else:
            name, bound, peak = f"k_reduce_parts<{el}>", "hbm", HBM_PEAK
        g = funcs.setdefault(name, dict(kinds=[], seconds=0.0, launches=0.0, flops=0.0, entries=0.0, bound=bound,
                                        peak=peak))
        g["kinds"].append(kind)
        for key in ("seconds", "launches", "flops", "entries"):
            g[key] += v[key]
    dom = max((k for k in funcs if funcs[k]["bound"] != "latency"), key=lambda k: funcs[k]["seconds"])
    g = funcs[dom]
    traffic = traffic_from_profiles(dom, f"hex8 {args.cells}^3 {form}")
    avg = g["seconds"] / max(g["launches"], 1)
    if g["bound"] == "hbm":
        es = 8 if form == "fp64" else 4
        per_launch = g["entries"] * (3 * es + 4) / max(g["launches"], 1)  # child entry read, target RMW, row map
        achieved = per_launch / avg / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK, unit="GB/s", frac=achieved / HBM_PEAK,
                    traffic=traffic, kernel=dom, kinds=g["kinds"], launches=g["launches"], bytes_per_launch=per_launch,
                    avg_launch_s=avg)
    else:
        achieved = g["flops"] / g["seconds"] / 1e12
        products = args.bf16_terms if dom.startswith("k_gemm_b16") else 1
        roof = dict(bound="mfma", achieved=achieved, peak=g["peak"], unit="TFLOP/s", frac=achieved / g["peak"],
                    traffic=traffic, kernel=dom, kinds=g["kinds"], launches=g["launches"],
                    flops_per_launch=g["flops"] / max(g["launches"], 1), avg_launch_s=avg,
                    mfma_products_per_flop=products, mfma_products_tflops=achieved * products,
                    mfma_products_frac=achieved * products / g["peak"])
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
    line["schedule"] = tuned  # per leg: the schedule its factorizations ran (the tuner's pick)
    # The factor's bound over its precision classes: Schur products (SYRK, UPDATE) at the MFMA
    # peak of their class (three BF16 products per flop in the mixed form), TRSM and diagonal
    # blocks at the FP32 (FP64) peak, the small fronts (FP64 arithmetic) at the FP64 peak.
    def factor_bound(leg, form_):
        k = legs[leg]["breakdown"]
        schur = k["syrk"]["flops"] + k["update"]["flops"]
        if form_ == "mixed":
            t_schur = args.bf16_terms * schur / (PEAK["bf16"] * 1e12)
        else:
            t_schur = schur / (PEAK["fp64" if form_ == "fp64" else "fp32"] * 1e12)
        t_trsm = k["trsm"]["flops"] / (PEAK["fp64" if form_ == "fp64" else "fp32"] * 1e12)
        t_diag = k["potrf"]["flops"] / (PEAK["fp64" if form_ == "fp64" else "fp32"] * 1e12)
        t_small = k["small"]["flops"] / (PEAK["fp64"] * 1e12)
        b = t_schur + t_trsm + t_diag + t_small
        return dict(bound_s=b, factor_s=legs[leg]["factor_s"], frac=b / legs[leg]["factor_s"],
                    schur_flops=schur, trsm_flops=k["trsm"]["flops"], diag_flops=k["potrf"]["flops"],
                    small_flops=k["small"]["flops"])
    line["factor_roofline"] = {leg: factor_bound(leg, leg) for leg in legs}
    if "fp64" in legs:
        f64 = legs["fp64"]
        t64 = f64["wall_s"]
        line["fp64"] = dict(factor_s=f64["factor_s"], solve_s=f64["solve_s"], wall_s=f64["wall_s"], tflops_factor=F / f64["factor_s"] / 1e12,
                            frac_fp64_peak=F / f64["factor_s"] / 1e12 / PEAK["fp64"], berr=f64["berr"],
                            iterations=f64["iterations"])
        line["speedup_vs_fp64"] = t64 / t_step
    line[form] = dict(factor_s=head["factor_s"], solve_s=head["solve_s"], wall_s=head["wall_s"],
                      iterations=head["iterations"],
                      berr=head["berr"], converged=head["converged"], rel_err_vs_fp64=head.get("rel_err_vs_fp64"))
    line["kernels"] = {leg: {k: dict(seconds=v["seconds"], launches=v["launches"],
                                     tflops=(v["flops"] / v["seconds"] / 1e12) if v["seconds"] and v["flops"] else None)
                             for k, v in legs[leg]["breakdown"].items()} for leg in legs}
    if args.contrast:
        pc = P.cube_hex8_contrast(args.cells)
        dv_c = torch.from_numpy(pc.upper()[2]).to(dev)
        Ac = CH.Matrix(pc.n, pc.ptr, pc.idx, pc.kval)
        bc = torch.from_numpy(pc.rhs).to(dev)
        xc = torch.zeros_like(bc)
        out = {}
        for leg in (["fp64"] if not args.no_fp64 else []) + [form]:
            r = run_leg(CH, sym, dv_c.data_ptr(), Ac, bc.data_ptr(), xc.data_ptr(), pc.n, leg, 1, 1, [],
                        bf16_terms=args.bf16_terms)
            out[leg] = dict(factor_s=r["factor_s"], solve_s=r["solve_s"], iterations=r["iterations"], berr=r["berr"],
                            converged=r["converged"])
            if leg == "fp64":
                xc64 = xc.cpu().numpy().copy()
            else:
                if not args.no_fp64:
                    out[leg]["rel_err_vs_fp64"] = float(np.linalg.norm(xc.cpu().numpy() - xc64) / np.linalg.norm(xc64))
            log(f"bench: contrast {leg} {out[leg]}")
        line["contrast"] = dict(workload=pc.meta.get("workload", "contrast"), **out)
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(args.cpu_cells, args.cpu_threads)
    print(json.dumps(line), flush=True)


def batch_workload(args, CH, P):
    """512 independent models as one block-diagonal matrix: FP32 factor, then 64 load cases per
    model refined in FP64 to berr <= 1e-12 per model (each model's own norms)."""
    import torch

    t0 = time.perf_counter()
    models = []
    for s in range(args.models):
        models.append(P.random_model(s))
        if s % 128 == 127:
            log(f"bench: {s + 1} models generated ({time.perf_counter() - t0:.0f} s)")
    bt = P.batch(models, nrhs=64)
    up = bt.upper()
    log(f"bench: block-diagonal matrix of {bt.n} equations assembled ({time.perf_counter() - t0:.0f} s)")
    sym = CH.Symbolic(CH.UpperColumns(bt.n, *up))
    info = sym.info()
    log(f"bench: batch of {len(models)} models, {bt.n} equations, generated and analysed in "
        f"{time.perf_counter() - t0:.1f}s; fronts {info['fronts']} levels {info['levels']} F={info['flops']:.3e}")
    dev = torch.device("cuda:0")
    dv = torch.from_numpy(up[2]).to(dev)
    A = CH.Matrix(bt.n, bt.ptr, bt.idx, bt.kval)
    A.set_segments(bt.segments)
    B = torch.from_numpy(np.ascontiguousarray(bt.rhs.T)).to(dev)
    X = torch.zeros_like(B)
    f = CH.Factor(sym, precision="fp32", device_values=dv.data_ptr())
    rep, berr = f.solve_refined_many_device(A, B.data_ptr(), X.data_ptr(), 64)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        f.refactor_device(dv.data_ptr())
        tf = f.seconds()
        rep, berr = f.solve_refined_many_device(A, B.data_ptr(), X.data_ptr(), 64)
        times.append((tf, rep["seconds"]))
    tf, ts = (float(np.mean([t[i] for t in times])) for i in (0, 1))
    F = info["flops"]
    line = dict(metric="batched factor + 64 refined load cases per model, 512 hex8/tet10 models (2k-60k dof)",
                value=F / (tf + ts) / 1e12, unit="TFLOP/s", n_gpus=1, steps=args.steps, warmup=1,
                ms_per_step=(tf + ts) * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="fp32 factor, fp64 refinement", data="synthetic",
                config=dict(workload=f"{len(models)} random hex8/tet10 models as one block-diagonal matrix",
                            equations=bt.n, flops=F, factor_entries=info["entries"], rhs_per_model=64),
                factor_s=tf, solve_s=ts, models_per_s=len(models) / (tf + ts), iterations=rep["iterations"],
                berr_max=float(np.max(berr)), converged=rep["converged"])
    print(json.dumps(line), flush=True)


def shift_workload(args, CH, P):
    """K - sigma M of a tet10 steel block (sigma = (2 pi 50)^2): FP64 factor for the reference
    solutions of the first block, then the mixed factor and 40 refined solves of 16 vectors."""
    import torch

    t0 = time.perf_counter()
    k = args.tet_cells
    p = P.block_tet10((k, k, k), (k / 100, k / 100, k / 100))
    log(f"bench: tet10 {k}^3 block of {p.n} equations generated ({time.perf_counter() - t0:.0f} s)")
    sigma = (2 * np.pi * 50.0) ** 2
    vals = p.kval - sigma * p.mval
    up = p.upper(vals)
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    info = sym.info()
    log(f"bench: tet10 {k}^3 shift-invert, {p.n} equations, set up in {time.perf_counter() - t0:.1f}s; "
        f"F={info['flops']:.3e} entries {info['entries']:.3e}")
    dev = torch.device("cuda:0")
    dv = torch.from_numpy(up[2]).to(dev)
    A = CH.Matrix(p.n, p.ptr, p.idx, vals)
    # the 40 blocks are drawn in order from one seeded generator, one at a time (15.7 GB together)
    rng = np.random.default_rng(20244)
    B0 = torch.from_numpy(rng.standard_normal((16, p.n))).to(dev)
    X = torch.zeros_like(B0)
    out = {}
    if not args.no_fp64:
        f64 = CH.Factor(sym, precision="fp64", device_values=dv.data_ptr())
        out["fp64_first_factor_s"] = f64.seconds()  # includes allocating the factor's storage
        f64.refactor_device(dv.data_ptr())
        out["fp64_factor_s"] = f64.seconds()
        X.copy_(B0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        f64.solve_device(X.data_ptr(), 16)
        torch.cuda.synchronize()
        out["fp64_solve_s"] = time.perf_counter() - t
        x64 = X.cpu().numpy().copy()
        log(f"bench: fp64 factor {out['fp64_factor_s']:.2f} s, 16-vector solve {out['fp64_solve_s'] * 1e3:.1f} ms")
        del f64
        torch.cuda.empty_cache()
    f = CH.Factor(sym, precision="mixed", device_values=dv.data_ptr(), bf16_terms=args.bf16_terms)
    out["mixed_first_factor_s"] = f.seconds()
    f.refactor_device(dv.data_ptr())
    out["mixed_factor_s"] = f.seconds()
    log(f"bench: mixed factor {out['mixed_factor_s']:.2f} s")
    lat, its, worst = [], [], 0.0
    for q in range(40):
        Bq = B0 if q == 0 else torch.from_numpy(rng.standard_normal((16, p.n))).to(dev)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rep, berr = f.solve_refined_many_device(A, Bq.data_ptr(), X.data_ptr(), 16)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
        its.append(rep["iterations"])
        worst = max(worst, float(np.max(berr)))
        if q == 0 and not args.no_fp64:
            xm = X.cpu().numpy()
            out["rel_err_vs_fp64"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
        del Bq
        if q % 10 == 9:
            log(f"bench: {q + 1} blocks, median {np.median(lat) * 1e3:.1f} ms, iterations {its[-1]}")
    F = info["flops"]
    line = dict(metric="shift-invert: mixed factor + 40 refined block solves (16 vectors), tet10 ~3M dof",
                value=float(np.median(lat)) * 1e3, unit="ms per 16-vector refined solve", n_gpus=1, steps=40,
                warmup=0, ms_per_step=float(np.median(lat)) * 1e3, higher_is_better=False, scaling="weak",
                vs_baseline=None, dtype="bf16/fp32 factor, fp64 refinement", data="synthetic",
                config=dict(workload=f"tet10 {k}^3 Kuhn block, steel, K - (2 pi 50)^2 M", equations=p.n, flops=F,
                            factor_entries=info["entries"]),
                factor_tflops=F / out["mixed_factor_s"] / 1e12, iterations_median=float(np.median(its)),
                berr_max=worst, **out)
    print(json.dumps(line), flush=True)


def assemble_workload(args, P):
    """The step before the factor, on the device (fc_assemble_device): the element matrices and
    their ordered sums into the upper columns of K_ff for the hex8 cube (the step) and of
    K - sigma M for the tet10 block, from the mesh and the plan already in HBM. The CPU baseline
    is the oracle's assembly of the same cube (the reference's loop, element matrices in
    parallel rounds, sums in element order) on all host threads."""
    import torch

    from fieldcraft_amd import assembly as AS

    sigma = (2 * np.pi * 50.0) ** 2
    k = args.tet_cells
    runs = {}
    for name, make, what in [("hex8", lambda: P.cube_hex8(args.cells), "K"),
                             ("tet10", lambda: P.block_tet10((k, k, k), (k / 100, k / 100, k / 100)), "shifted")]:
        t0 = time.perf_counter()
        p = make()
        uptr, uidx, _ = p.upper()
        t1 = time.perf_counter()
        a = AS.Assembly.of_problem(p, uptr, uidx, timing=True)
        plan_s = time.perf_counter() - t1
        log(f"bench: {name} {p.n} equations, {len(p.conn)} elements (mesh {t1 - t0:.1f} s, plan {plan_s:.2f} s)")
        out = torch.empty(a.nnz, dtype=torch.float64, device="cuda:0")
        for _ in range(args.warmup):
            a.assemble_device(out.data_ptr(), what, sigma)
        wall, el, ga = [], [], []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            a.assemble_device(out.data_ptr(), what, sigma)
            wall.append(time.perf_counter() - t)
            inf = a.info()
            el.append(inf["elements_s"])
            ga.append(inf["gather_s"])
        inf = a.info()
        npe, W = (8, 24) if p.kind == "hex8" else (10, 30)
        E, nnz, ck, cm = len(p.conn), a.nnz, inf["contributions_k"], inf["contributions_m"]
        # algorithmic bytes: the gather reads the ranges, positions and values of each
        # contribution and writes each entry; the element kernels write the packed stiffness
        # (and the 10 x 10 mass) and read the connectivity, the slot map and the coordinates
        gbytes = (nnz + 1) * 8 + ck * 12 + nnz * 8 + ((nnz + 1) * 8 + cm * 12 if cm else 0)
        ebytes = E * (W * (W + 1) // 2 * 8 + npe * 4 + 4 + npe * 24) + (E * (100 * 8 + npe * 28 + 4) if cm else 0)
        runs[name] = dict(p=p, a=a, what=what, wall=float(np.mean(wall)), elements_s=float(np.mean(el)),
                          gather_s=float(np.mean(ga)), gbytes=gbytes, ebytes=ebytes, E=E, nnz=nnz, ck=ck, cm=cm,
                          plan_s=plan_s)
        log(f"bench: {name} assemble {np.mean(wall) * 1e3:.2f} ms (elements {np.mean(el) * 1e3:.2f} ms, "
            f"gather {np.mean(ga) * 1e3:.2f} ms)")
    h = runs["hex8"]
    kern = {"k_hex8_stiffness": (h["elements_s"], h["ebytes"]), "k_assemble_gather": (h["gather_s"], h["gbytes"])}
    dom = max(kern, key=lambda n: kern[n][0])
    sec, byt = kern[dom]
    achieved = byt / sec / 1e9
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK, unit="GB/s", frac=achieved / HBM_PEAK,
                traffic=traffic_from_profiles(dom, f"hex8 {args.cells}^3 K"), kernel=dom, avg_launch_s=sec,
                bytes_per_launch=byt, kernels={n: dict(seconds=v[0], bytes=v[1], gbs=v[1] / v[0] / 1e9)
                                               for n, v in kern.items()})
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        threads = args.cpu_threads
        O.set_threads(threads)
        p = h["p"]
        uptr, uidx, _ = p.upper()
        t = time.perf_counter()
        O.assemble_upper(p.kind, p.xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx)
        cs = time.perf_counter() - t
        cpu = dict(value=h["E"] / cs, unit="elements/s", cores=threads, kind="port",
                   sample=f"oracle assembly of the whole hex8 {args.cells}^3 K_ff ({h['E']} elements): {cs:.2f} s "
                          f"on {threads} threads")
        log(f"bench: cpu baseline {cs:.2f} s")
    t10 = runs["tet10"]
    line = dict(metric=f"assembly of K_ff on the device (reference values bit for bit), hex8 {args.cells}^3: elements/s",
                value=h["E"] / h["wall"], unit="elements/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=h["wall"] * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64",
                data="synthetic",
                config=dict(workload=f"hex8 {args.cells}^3 unit cube, z=0 clamped, K_ff upper columns",
                            equations=h["p"].n, elements=h["E"], entries=h["nnz"], contributions=h["ck"]),
                roofline=roof, cpu_baseline=cpu, plan_s=h["plan_s"],
                tet10_shifted=dict(workload=f"tet10 {k}^3 Kuhn block, steel, K - (2 pi 50)^2 M", equations=t10["p"].n,
                                   elements=t10["E"], entries=t10["nnz"], ms=t10["wall"] * 1e3,
                                   elements_ms=t10["elements_s"] * 1e3, gather_ms=t10["gather_s"] * 1e3,
                                   gather_gbs=t10["gbytes"] / t10["gather_s"] / 1e9, plan_s=t10["plan_s"]))
    print(json.dumps(line), flush=True)


def modes_workload(args, CH, P):
    """The modal solve on the device (the reference's frequency step, analysis.rs:655-740, with
    the shift of the tet10 workload): K - sigma M assembled from the mesh (fc_assemble_device), its
    FP64 factor from those device values, and the lowest `--modes` modes above sigma by the shifted
    block Lanczos method (fc_lanczos_device, block 6, tolerance 1e-9). One step = assembly +
    refactorization + Lanczos; the first factorization (which also allocates the factor) is
    outside the timed steps. The CPU baseline is the oracle (numpy Lanczos over the C oracle's
    factor) on a smaller tet10 block."""
    import torch

    from fieldcraft_amd import assembly as AS

    k = args.tet_cells
    sigma = (2 * np.pi * 50.0) ** 2
    t0 = time.perf_counter()
    p = P.block_tet10((k, k, k), (k / 100, k / 100, k / 100))
    uptr, uidx, _ = p.upper()
    a = AS.Assembly.of_problem(p, uptr, uidx)
    sym = CH.Symbolic(CH.UpperColumns(p.n, uptr, uidx, np.zeros(len(uidx))))
    info = sym.info()
    M = CH.Matrix(p.n, p.ptr, p.idx, p.mval)
    log(f"bench: tet10 {k}^3, {p.n} equations, set up in {time.perf_counter() - t0:.1f} s")
    vals = torch.empty(len(uidx), dtype=torch.float64, device="cuda:0")
    a.assemble_device(vals.data_ptr(), "shifted", sigma)
    gemm_kinds = ("update", "syrk", "trsm")  # the kinds k_gemm<double> runs (HIP events per launch)
    f = CH.Factor(sym, precision="fp64", device_values=vals.data_ptr(), profile=gemm_kinds)
    first_factor_s = f.seconds()
    max_basis = 600
    runs = []
    for step in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        a.assemble_device(vals.data_ptr(), "shifted", sigma)
        ta = time.perf_counter()
        f.refactor_device(vals.data_ptr())
        torch.cuda.synchronize()
        tf = time.perf_counter()
        r = f.lanczos_device(M, sigma, args.modes, block=6, max_basis=max_basis, tolerance=1e-9)
        torch.cuda.synchronize()
        te = time.perf_counter()
        prof = [f.profile(kk) for kk in gemm_kinds]
        runs.append(dict(total=te - t, assemble=ta - t, factor=tf - ta, lanczos=te - tf, basis=r["basis"],
                         lanczos_device_s=r["seconds"], gemm_s=sum(x["seconds"] for x in prof),
                         gemm_flops=sum(x["flops"] for x in prof), gemm_launches=sum(x["launches"] for x in prof)))
        log(f"bench: step {step}: assemble {1e3 * (ta - t):.1f} ms, factor {tf - ta:.2f} s, "
            f"Lanczos {te - tf:.2f} s ({r['basis']} vectors, {len(r['values'])} modes)")
    timed = runs[args.warmup:]
    mean = {key: float(np.mean([x[key] for x in timed])) for key in timed[0]}
    # accuracy of the last run's first modes: ||K x - lambda M x||_inf / ||K x||_inf on the host
    import scipy.sparse as sps

    K = sps.csr_matrix((p.kval, p.idx, p.ptr), shape=(p.n, p.n))
    Ms = sps.csr_matrix((p.mval, p.idx, p.ptr), shape=(p.n, p.n))
    X = r["vectors"][:, :3].cpu().numpy()
    res = [float(np.abs(K @ X[:, c] - r["values"][c] * (Ms @ X[:, c])).max() / np.abs(K @ X[:, c]).max())
           for c in range(X.shape[1])]
    cpu = None
    if not args.no_cpu:
        from oracle import eigen as EG
        from oracle import oracle as O

        O.set_threads(args.cpu_threads)
        kc = 10
        q = P.block_tet10((kc, kc, kc), (kc / 100, kc / 100, kc / 100))
        qu = q.upper(q.kval - sigma * q.mval)
        t = time.perf_counter()
        of = O.Factor(q.n, *qu)
        Mq = sps.csr_matrix((q.mval, q.idx, q.ptr), shape=(q.n, q.n))
        op = EG.Operator(q.n, lambda x: Mq @ x, lambda Xb: np.stack([of.solve(Xb[:, c]) for c in range(Xb.shape[1])], 1))
        cr = EG.lanczos(op, sigma, args.modes, block=6, max_basis=max_basis, tolerance=1e-9)
        cs = time.perf_counter() - t
        cpu = dict(value=cs, unit="s", cores=args.cpu_threads, kind="port",
                   sample=f"oracle factor (C, {args.cpu_threads} threads) + numpy Lanczos of the tet10 {kc}^3 block "
                          f"({q.n} equations, {args.modes} modes, {cr['basis']} vectors): {cs:.2f} s")
        log(f"bench: cpu baseline {cs:.2f} s on {q.n} equations")
    F = info["flops"]
    # the step's dominant kernel: the FP64 factor's k_gemm<double> (panel updates, update matrices,
    # TRSM), its algorithmic flops over its event-timed seconds
    achieved = mean["gemm_flops"] / mean["gemm_s"] / 1e12
    roof = dict(bound="mfma", achieved=achieved, peak=PEAK["fp64"], unit="TFLOP/s", frac=achieved / PEAK["fp64"],
                traffic=None, kernel="k_gemm<double, false>", kinds=list(gemm_kinds), launches=mean["gemm_launches"],
                avg_launch_s=mean["gemm_s"] / max(mean["gemm_launches"], 1))
    # the Lanczos phase against HBM: per block of b = 6 vectors the FP64 solve reads L twice, the
    # two Gram-Schmidt passes read MQ and Q (n x m each) and write the block, and about 8 products
    # with M read its nnz values and indices
    b, m_final, nnzm = 6, int(timed[-1]["basis"]), int(p.ptr[-1])
    lbytes = sum(2 * info["entries"] * 8 + 2 * 2 * p.n * m * 8 + 8 * (nnzm * 12 + p.n * 8 * 2)
                 for m in range(b, m_final + 1, b))
    lanczos_gbs = lbytes / mean["lanczos_device_s"] / 1e9
    line = dict(metric=f"modal solve on the device, tet10 {k}^3 (~3M dof): K - sigma M assembled, FP64 factor, "
                       f"{args.modes} modes by shift-invert block Lanczos: s",
                value=mean["total"], unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=mean["total"] * 1e3, higher_is_better=False, scaling="weak", vs_baseline=None, dtype="f64",
                data="synthetic",
                config=dict(workload=f"tet10 {k}^3 Kuhn block, steel, z=0 clamped, sigma = (2 pi 50)^2, {args.modes} "
                                     f"modes, block 6, tolerance 1e-9", equations=p.n, factor_entries=info["entries"],
                            flops=F),
                assemble_ms=mean["assemble"] * 1e3, factor_s=mean["factor"], factor_tflops=F / mean["factor"] / 1e12,
                lanczos_s=mean["lanczos"], lanczos_device_s=mean["lanczos_device_s"], basis=int(timed[-1]["basis"]),
                first_factor_s=first_factor_s, frequencies_hz=(np.sqrt(r["values"]) / (2 * np.pi)).tolist(),
                residual_first3=res, roofline=roof,
                lanczos_roofline=dict(bound="hbm", achieved=lanczos_gbs, peak=HBM_PEAK, unit="GB/s",
                                      frac=lanczos_gbs / HBM_PEAK, bytes=lbytes),
                cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


def ldlt_workload(args, CH, P):
    """LDL^T with pivoting (LdltFactor, ldlt.rs) of an indefinite shifted stiffness: the tet10 block's
    first modes found on the device (assembly, FP64 Cholesky, Lanczos), sigma put between the 5th
    and 6th, K - sigma M factored by LDL^T on the device (its inertia must show the 5 modes below
    sigma) and solved. One step = the LDL^T factorization (host plan included) + one solve. The
    CPU baseline is the oracle's LDL^T of the same matrix."""
    import torch

    from fieldcraft_amd import assembly as AS

    k = args.ldlt_cells
    t0 = time.perf_counter()
    p = P.block_tet10((k, k, k), (k / 100, k / 100, k / 100))
    uptr, uidx, _ = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, uptr, uidx, np.zeros(len(uidx))))
    info = sym.info()
    a = AS.Assembly.of_problem(p, uptr, uidx)
    sigma0 = (2 * np.pi * 50.0) ** 2
    vals = a.assemble("shifted", sigma0)
    chol = CH.Factor(sym, device_values=vals.data_ptr())
    M = CH.Matrix(p.n, p.ptr, p.idx, p.mval)
    modes = chol.lanczos_device(M, sigma0, 12, block=6, max_basis=min(p.n, 600), vectors=False)["values"]
    # sigma between two well separated modes from the 5th on: that many eigenvalues lie below it
    below = next(i + 1 for i in range(4, len(modes) - 1) if modes[i + 1] > 1.05 * modes[i])
    sigma = 0.5 * (modes[below - 1] + modes[below])
    del chol
    up = p.upper(p.kval - sigma * p.mval)
    log(f"bench: tet10 {k}^3, {p.n} equations, sigma between {np.sqrt(modes[below - 1]) / 2 / np.pi:.1f} and "
        f"{np.sqrt(modes[below]) / 2 / np.pi:.1f} Hz (set up in {time.perf_counter() - t0:.1f} s)")
    b = np.cos(np.arange(p.n) * 0.3)
    runs = []
    for step in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        f = CH.LdltFactor(sym, up[2])
        tf = time.perf_counter()
        x = f.solve(b)
        te = time.perf_counter()
        runs.append(dict(total=te - t, factor=tf - t, solve=te - tf))
        log(f"bench: step {step}: factor {tf - t:.3f} s, solve {te - tf:.3f} s")
    timed = runs[args.warmup:]
    mean = {key: float(np.mean([r[key] for r in timed])) for key in timed[0]}
    inertia, two = f.inertia(), f.two_by_two()
    res = float(np.abs(p.mul(x, p.kval - sigma * p.mval) - b).max() / np.abs(b).max())
    F = info["flops"]
    achieved = F / mean["factor"] / 1e12
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        O.set_threads(args.cpu_threads)
        t = time.perf_counter()
        of = O.LdltFactor(p.n, *up)
        of.solve(b)
        cs = time.perf_counter() - t
        same = bool(np.array_equal(of.values().view(np.uint64), f.values().view(np.uint64)))
        cpu = dict(value=cs, unit="s", cores=args.cpu_threads, kind="port",
                   sample=f"oracle LDL^T factor + solve of the same matrix ({p.n} equations) on {args.cpu_threads} "
                          f"threads: {cs:.2f} s; factor bits equal to the device's: {same}")
        log(f"bench: cpu baseline {cs:.2f} s, same bits {same}")
    line = dict(metric=f"LDL^T with pivoting of an indefinite K - sigma M, tet10 {k}^3: factor + solve s",
                value=mean["total"], unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=mean["total"] * 1e3, higher_is_better=False, scaling="weak", vs_baseline=None, dtype="f64",
                data="synthetic",
                config=dict(workload=f"tet10 {k}^3 Kuhn block, steel, z=0 clamped, sigma between modes {below} and "
                                     f"{below + 1}",
                            equations=p.n, factor_entries=info["entries"], flops=F, sigma=float(sigma)),
                factor_s=mean["factor"], solve_s=mean["solve"], inertia=list(inertia), negative_expected=below,
                two_by_two=two, residual=res,
                roofline=dict(bound="mfma", achieved=achieved, peak=PEAK["fp64"], unit="TFLOP/s",
                              frac=achieved / PEAK["fp64"], traffic=None, kernel="all LDL^T factor kernels",
                              note="the Cholesky flop count of the same structure over the factor's wall time (host "
                                   "plan cached by the warmup); LDL^T is bound by chains of dependent steps (pivot "
                                   "searches, per-entry sums in a fixed order), not by flops or bytes"),
                cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


def zldlt_workload(args, CH, P):
    """Complex-symmetric LDL^T (LdltFactor<Complex64>, ldlt.rs over scalar.rs) of a harmonic
    response's dynamic stiffness K - w^2 M + i w C on the tet10 block, C = 1e-4 K + 10 M
    (Rayleigh damping), w = 2 pi 5 kHz (above the first modes of the 12^3 block: indefinite real
    part). One step = the factorization + one solve. The CPU baseline is the oracle's complex
    LDL^T of the same matrix."""
    import torch

    k = args.ldlt_cells
    t0 = time.perf_counter()
    p = P.block_tet10((k, k, k), (k / 100, k / 100, k / 100))
    uptr, uidx, _ = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, uptr, uidx, np.zeros(len(uidx))))
    info = sym.info()
    w = 2 * np.pi * 5000.0
    full = (p.kval - w * w * p.mval) + 1j * w * (1e-4 * p.kval + 10.0 * p.mval)
    _, _, uval_re = p.upper(full.real)
    _, _, uval_im = p.upper(full.imag)
    uval = uval_re + 1j * uval_im
    log(f"bench: tet10 {k}^3 harmonic, {p.n} equations (set up in {time.perf_counter() - t0:.1f} s)")
    b = np.cos(np.arange(p.n) * 0.3) + 1j * np.sin(np.arange(p.n) * 0.2)
    runs = []
    for step in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        f = CH.ZLdltFactor(sym, uval)
        tf = time.perf_counter()
        x = f.solve(b)
        te = time.perf_counter()
        runs.append(dict(total=te - t, factor=tf - t, solve=te - tf))
        log(f"bench: step {step}: factor {tf - t:.3f} s, solve {te - tf:.3f} s")
    timed = runs[args.warmup:]
    mean = {key: float(np.mean([r[key] for r in timed])) for key in timed[0]}
    rows = np.repeat(np.arange(p.n), np.diff(p.ptr))
    ax = np.zeros(p.n, np.complex128)
    np.add.at(ax, rows, full * x[p.idx])
    res = float(np.abs(ax - b).max() / np.abs(b).max())
    F = info["flops"]
    achieved = 4 * F / mean["factor"] / 1e12  # a complex multiply-add is 4 real ones
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O

        O.set_threads(args.cpu_threads)
        t = time.perf_counter()
        of = O.ZLdltFactor(p.n, uptr, uidx, uval)
        xo = of.solve(b)
        cs = time.perf_counter() - t
        same = bool(np.array_equal(of.values().view(np.uint64), f.values().view(np.uint64)) and
                    np.array_equal(xo.view(np.uint64), x.view(np.uint64)))
        cpu = dict(value=cs, unit="s", cores=args.cpu_threads, kind="port",
                   sample=f"oracle complex LDL^T factor + solve of the same matrix ({p.n} equations) on "
                          f"{args.cpu_threads} threads: {cs:.2f} s; factor and solution bits equal to the device's: "
                          f"{same}")
        log(f"bench: cpu baseline {cs:.2f} s, same bits {same}")
    line = dict(metric=f"complex-symmetric LDL^T of K - w^2 M + i w C, tet10 {k}^3: factor + solve s",
                value=mean["total"], unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=mean["total"] * 1e3, higher_is_better=False, scaling="weak", vs_baseline=None,
                dtype="c128", data="synthetic",
                config=dict(workload=f"tet10 {k}^3 Kuhn block, steel, z=0 clamped, w = 2 pi 5 kHz, "
                                     f"C = 1e-4 K + 10 M", equations=p.n, factor_entries=info["entries"],
                            flops_real_equivalent=4 * F),
                factor_s=mean["factor"], solve_s=mean["solve"], two_by_two=f.two_by_two(), residual=res,
                roofline=dict(bound="mfma", achieved=achieved, peak=PEAK["fp64"], unit="TFLOP/s",
                              frac=achieved / PEAK["fp64"], traffic=None, kernel="all LDL^T factor kernels",
                              note="four times the Cholesky flop count of the structure over the factor's wall "
                                   "time; bound by chains of dependent steps like the real LDL^T"),
                cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
