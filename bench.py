"""Benchmark: cell-RK-stage updates/s of the explicit finite-volume step on
one MI355X, mixed precision vs the FP64 path, with the HBM roofline of the
dominant kernel and a CPU baseline (OpenMP FP64 oracle on a bounded sample).

A step = SSP-RK3 (3 stages: pass G + pass F each, dt from the CFL condition
at stage 0) + the per-cell phase relaxation, on the 48M-tet shock-bubble
workload (200^3 jittered hexes x 6 tets, 64 vapour bubbles, WALE on).
Inputs are resident in HBM before the timed region.

Legs at 48M cells: mixed precision with the surrogate relaxation (the
headline), the same with the cell state in place of the network (cold
start, what the network adds), mixed with the reference's plain relaxation,
and FP64.  Errors against the FP64 CPU oracle come from the same legs run on
the oracle's sample (40^3 x 6 cells, the steps the oracle ran).  A
relaxation micro-benchmark times the three relaxations on far-from-
equilibrium states, where chemical cells are common.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 200]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(neq, s, fc, s_geo, s_grad):
    """Algorithmic bytes per cell-RK-stage, split by pass (compulsory
    traffic of the two-pass design, every array touched once per pass):
      pass G: U read, geometry, face states (4 NR values) written
      pass F: U read, U0 read (stages 2..s), U_new written, geometry,
              face states read.
    Geometry per pass: per face fc2el(8) + bc(4) + n.A (3 S) + x_f (3 S),
    per cell V + x_c (4 S) + 4 slot ids (16)."""
    geo = fc * (12 + 6 * s_geo) + (4 * s_geo + 16)
    fs = 4 * neq * s_grad
    g = 8 * neq + geo + fs
    f_nou0 = 8 * neq + 8 * neq + geo + fs
    f_u0 = f_nou0 + 8 * neq
    total = g + 8 * neq * (2 + (s - 1) / s) + geo + fs  # pass G + pass F averaged over the s stages
    return dict(g=g, f_nou0=f_nou0, f_u0=f_u0, total=total, geo=geo)


def relax_bytes(neq):
    return 8 * neq + 48


def kernel_bytes(name, B, neq, mixed=True):
    """Algorithmic bytes per cell of one launch of kernel `name` (None when
    the kernel has no per-cell byte model).  pass_f_relax is the last RK
    stage (U0 read) with the per-cell relaxation on the state in registers:
    it adds the FP32 primitive record of each relaxed cell (4 NEQ, mixed)."""
    return {
        "pass_g": B["g"],
        "pass_f": B["f_u0"],          # fused stage update (stages 2..s read U0)
        "pass_f_relax": B["f_u0"] + (4 * neq if mixed else 0),
        "pass_f_res": B["f_nou0"],    # residual written in place of U_new
        "pass_f_rhs": B["f_nou0"] + 8,  # + per-cell time-step bound
        "rk_update": 8 * neq * 4 + 8,  # U, U0, residual read, U_new written, V
        "relax": relax_bytes(neq),
        "relax_screen": 8 * 9 + 8 * 3,  # U[0..8] read, alpha, alpha_k rho_k e_k written
    }.get(name)


# Entries of the counter summary for the kernels a timed run launches: the
# mixed fused stage writes primitive records (pass_f_w); the FP64 entries
# carry the fp64_ prefix and its last stage runs the plain relaxation.
TRAFFIC_NAMES = {
    "mixed_mlp": {"pass_g": "pass_g", "pass_f_rhs": "pass_f_rhs", "pass_f": "pass_f_w",
                  "pass_f_relax": "pass_f_relax", "rk_update": "rk_update"},
    "fp64": {"pass_g": "fp64_pass_g", "pass_f_rhs": "fp64_pass_f_rhs", "pass_f": "fp64_pass_f",
             "pass_f_relax": "fp64_pass_f_relax_plain", "rk_update": "fp64_rk_update"},
}


def traffic_rate(kt, seconds, cells, leg):
    """Bytes per second a timed run moved between L2 and the memory fabric:
    the summary's counter bytes per launch of each kernel (FETCH_SIZE x 2 +
    WRITE_SIZE) times its launches in the timed steps, over the wall time.
    Gather over-fetch and Infinity Cache hits are included, so this bounds
    the HBM traffic from above; it is the figure to set against the copy
    rate.  None without a summary of this size; kernels without an entry
    (the small reductions and list kernels) are left out."""
    names = TRAFFIC_NAMES.get(leg)
    if names is None:
        return None
    total = 0.0
    for k, (n, _) in kt.items():
        if k not in names:
            continue
        b = measured_traffic(names[k], cells)
        if b is None:
            return None
        total += b * n
    return total / seconds


def measured_traffic(kernel, cells):
    """HBM bytes per launch of `kernel` (read + written) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated
    on a plain copy), or None when there is none for this kernel / size."""
    path = os.path.join(ROOT, "profiles", "shock_bubble_48M_traffic.json")
    if cells != 48_000_000 or not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else k["read_bytes_per_dispatch"] + k["write_bytes_per_dispatch"]


def run_gpu(case, precision, steps, warmup, relax_mode=0, device=0, train=None, fuse=1, stage_f=1, stage_g=1):
    """Time `steps` explicitLoop steps on the device (state resident in HBM).
    train: dict of surrogate-training options (done before the timed region)."""
    from fooskagpu_amd import GpuSolver, surrogate
    p = dict(case.params, precision=precision, relaxMode=0)
    g = GpuSolver(p, case.mesh, case.U, device=device)
    g.set_tuning("fuse_update", fuse)
    g.set_tuning("stage_f", stage_f)
    g.set_tuning("stage_g", stage_g)
    info = {}
    if train is not None:
        t0 = time.perf_counter()
        _, (l0, l1, n) = surrogate.train_on_solver(g, case.U, case.params, log=log, **train)
        info = {"train_s": time.perf_counter() - t0, "loss_first": l0, "loss_last": l1, "train_samples": n}
        log(f"surrogate trained on {n} states in {info['train_s']:.1f} s, loss {l0:.4f} -> {l1:.6f}")
    g.set_params(dict(p, relaxMode=relax_mode))
    g.sync()
    if warmup:
        g.run_steps(warmup)
    g.upload_state(case.U)
    g.sync()
    g.set_timing(True)
    if case.neq == 10:
        g.relax_totals(reset=True)
    t0 = time.perf_counter()
    dt, tphys = g.run_steps(steps)  # returns after a stream sync
    el = time.perf_counter() - t0
    kt = g.kernel_times()
    if case.neq == 10:
        rt = g.relax_totals()
        cs = max(1, case.mesh.n_cells * steps)
        info["relaxation"] = {"newton_iters_per_cell": rt["newton_iters"] / cs,
                              "tsat_iters_per_cell": rt["tsat_iters"] / cs,
                              "chem_fraction": rt["n_chem"] / cs, "warm_fraction": rt["n_warm"] / cs,
                              "fallback_fraction": rt["n_fallback"] / cs, "n_chem": rt["n_chem"],
                              "n_warm": rt["n_warm"]}
        log(f"relaxation: {info['relaxation']}")
    U = g.download_state()
    g.close()
    return el, kt, U, dt, info


def rel_errors(U, Uo):
    """Relative L1 of density, momentum, total energy and vapour fraction."""
    rho, rho_o = U[:, 1] + U[:, 2], Uo[:, 1] + Uo[:, 2]
    return {"density": float(np.abs(rho - rho_o).sum() / np.abs(rho_o).sum()),
            "momentum": float(np.abs(U[:, 3:6] - Uo[:, 3:6]).sum() / max(np.abs(Uo[:, 3:6]).sum(), 1e-300)),
            "energy": float(np.abs(U[:, 8] - Uo[:, 8]).sum() / np.abs(Uo[:, 8]).sum()),
            "vapour_fraction": float(np.abs(Uo[:, 0] - U[:, 0]).sum() / np.abs(1 - Uo[:, 0]).sum())}


def cpu_baseline(target_s=15.0):
    """FP64 oracle with OpenMP on a bounded sample of the same workload
    (shock-bubble at 40^3 hexes x 6 = 384k cells), steps until ~target_s.
    Returns the baseline and (case, steps, oracle state) for the error legs."""
    from fooskagpu_amd import cases
    from oracle import pyoracle
    from oracle.pyoracle import Oracle
    threads = pyoracle.threads(omp=True)
    c = cases.shock_bubble(n=40, nbub=64, order="morton")
    o = Oracle(c.params, c.mesh, c.U, omp=True)
    steps, el = 0, 0.0
    while el < target_s and steps < 50:
        t0 = time.perf_counter()
        o.run_steps(1)
        el += time.perf_counter() - t0
        steps += 1
    v = c.mesh.n_cells * 3 * steps / el
    base = dict(value=v, unit="cell-RK-stage updates/s", cores=threads, kind="port",
                nproc=len(os.sched_getaffinity(0)),
                sample=f"FP64 OpenMP oracle ({threads} threads), shock-bubble 40^3x6 = {c.mesh.n_cells} cells "
                       f"in Morton order, {steps} SSP-RK3 steps incl. relaxation, {el:.1f} s")
    return base, (c, steps, o.state())


def cpu_configs():
    """The OpenMP oracle (and the serial one on Sod) timed on the benchmark
    configurations: Sod in full (100 steps), the 64^3 TGV and 128^3 bubble
    meshes for a few steps each (their full runs take minutes on the CPU).
    Rates in cell-RK-stage updates/s, setup excluded, nothing extrapolated."""
    from fooskagpu_amd import cases
    from oracle import pyoracle
    from oracle.pyoracle import Oracle
    out = {"omp_threads": pyoracle.threads(omp=True), "nproc": len(os.sched_getaffinity(0))}
    for name, make, steps, omps in (("sod_32x8x8", lambda: cases.sod(), 100, (False, True)),
                                    ("tgv_64", lambda: cases.tgv(n=64), 3, (True,)),
                                    ("bubble_128", lambda: cases.bubble(n=128), 2, (True,))):
        c = make()
        for omp in omps:
            o = Oracle(c.params, c.mesh, c.U, omp=omp)
            t0 = time.perf_counter()
            o.run_steps(steps)
            el = time.perf_counter() - t0
            key = f"{name}_{'omp' if omp else 'serial'}"
            out[key] = {"cells": c.mesh.n_cells, "steps": steps, "seconds": el,
                        "value": c.mesh.n_cells * 3 * steps / el}
            log(f"cpu {key}: {out[key]['value']:.4g} cell-stage/s ({el:.1f} s)")
            del o
        del c
    return out


def relax_microbench(n=88):
    """The three relaxations on far-from-equilibrium states (phasic pressures
    and temperatures drawn independently, mixture densities 0.02 .. 1000
    kg/m^3; about a third of the cells are chemical): plain reference Newton
    solves, the surrogate pipeline from the cell state (cold) and with the
    MLP warm start.  One fk_relaxation each on n^3 x 6 cells."""
    from fooskagpu_amd import GpuSolver, RELAX_COLD, RELAX_MLP, RELAX_PLAIN, cases, surrogate
    c = cases.bubble(n=n)
    out = {"cells": c.mesh.n_cells}
    with GpuSolver(c.params, c.mesh, c.U) as g:
        sg = surrogate.Surrogate(g)
        x, y = surrogate.finite(*sg.label(surrogate.uniform_states(1_000_000, c.params, seed=6)))
        sg.set_data(x, y)
        sg.set_weights(surrogate.init_weights(0), surrogate.fit_norm(x, y))
        l0, l1 = sg.train(3000, batch=8192, lr=2e-3, seed=2)
        out["training"] = {"labels": len(x), "loss_first": l0, "loss_last": l1}
        U = surrogate.uniform_states(g.n_cells, c.params, seed=5)
        for name, mode in (("plain", RELAX_PLAIN), ("cold", RELAX_COLD), ("mlp", RELAX_MLP)):
            best = None
            for _ in range(2):
                g.upload_state(U)
                g.sync()
                t0 = time.perf_counter()
                st = g.relaxation(mode)  # returns after reading the counts (a sync)
                el = time.perf_counter() - t0
                best = el if best is None else min(best, el)
            out[name] = {"ms": best * 1e3, "newton_iters_per_cell": st["newton_iters"] / g.n_cells,
                         "chem_fraction": st["n_chem"] / g.n_cells, "fallbacks": st["n_fallback"],
                         "warm": st["n_warm"]}
            log(f"relaxation micro-benchmark {name}: {out[name]}")
    out["mlp_vs_cold_time"] = out["mlp"]["ms"] / out["cold"]["ms"]
    out["mlp_vs_cold_iters"] = out["mlp"]["newton_iters_per_cell"] / out["cold"]["newton_iters_per_cell"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=200, help="hexes per axis (200 -> 48M tets)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs (baseline, per-config timings)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-mlp", action="store_true", help="skip the surrogate runs")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the cold-start leg, the oracle error legs and the relaxation micro-benchmark")
    ap.add_argument("--order", default="morton", choices=["rcm", "morton", "none"], help="cell order of the mesh")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--stage-f", type=int, default=1, choices=[0, 1],
                    help="pass F stages the workgroup's face states in LDS (256 cells, FP64 128); same results")
    ap.add_argument("--fuse", type=int, default=1, choices=[0, 1],
                    help="RK update inside pass F (1) or residual + separate update kernel (0); same results")
    ap.add_argument("--stage-g", type=int, default=1, choices=[0, 1],
                    help="pass G writes its face records through LDS (256 cells, FP64 128); same results")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30,
                    help="size of the plain device copy measured as the HBM reference (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (halo exchange is out of scope); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from fooskagpu_amd import FP64, MIXED, RELAX_COLD, RELAX_MLP, RELAX_PLAIN, cases
    t0 = time.perf_counter()
    case = cases.shock_bubble(n=args.n, nbub=64, order=args.order)
    C = case.mesh.n_cells
    fc = case.mesh.n_faces / C
    log(f"mesh {C} cells, {case.mesh.n_faces} faces (F/C {fc:.4f}) built in {time.perf_counter() - t0:.1f} s")
    s = 3
    res = {}
    copy = None
    if args.copy_bytes:
        from fooskagpu_amd import GpuSolver
        with GpuSolver(dict(case.params)) as g0:
            ms = min(g0.stream_copy(args.copy_bytes) for _ in range(3))
        copy = {"bytes": args.copy_bytes, "ms": ms, "GBps": 2 * args.copy_bytes / ms / 1e6}
        log(f"stream_copy {args.copy_bytes} B: {ms:.3f} ms, {copy['GBps']:.0f} GB/s (read + write)")
    ntrain = min(4_000_000, C)
    train = dict(n_solver=ntrain, n_uniform=ntrain // 4, warm_steps=3, train_steps=args.train_steps)
    plan = [("mixed_mlp", MIXED, RELAX_MLP, train), ("mixed_cold", MIXED, RELAX_COLD, None),
            ("mixed", MIXED, RELAX_PLAIN, None), ("fp64", FP64, RELAX_PLAIN, None)]
    for name, prec, mode, tr in plan:
        if (name == "fp64" and args.no_fp64) or (name.startswith("mixed_") and args.no_mlp) or \
                (name == "mixed_cold" and args.no_extra):
            continue
        log(f"{name}: warmup {args.warmup}, timing {args.steps} steps")
        el, kt, U, dt, info = run_gpu(case, prec, args.steps, args.warmup, relax_mode=mode, train=tr, fuse=args.fuse,
                                      stage_f=args.stage_f, stage_g=args.stage_g)
        rate = C * s * args.steps / el
        log(f"{name}: {el:.3f} s, {rate:.4g} cell-stage/s, last dt {dt:.3e}, finite={bool(np.isfinite(U).all())}")
        for k, (n, ms) in sorted(kt.items(), key=lambda kv: -kv[1][1]):
            log(f"   {k:14s} {n:5d} launches {ms:9.2f} ms  avg {ms / max(n, 1):8.3f} ms")
        res[name] = dict(el=el, rate=rate, kt=kt, U=U, info=info)

    head = "mixed_mlp" if "mixed_mlp" in res else "mixed"
    mixed = res[head]
    sb = 4
    B = stage_bytes(10, s, fc, sb, sb)
    kt = mixed["kt"]
    # dominant kernel (largest total time) with a byte model
    dom = max((k for k in kt if kernel_bytes(k, B, 10)), key=lambda k: kt[k][1])
    n_f, ms_f = kt[dom]
    avg_f = ms_f / max(n_f, 1) / 1e3
    bpc = kernel_bytes(dom, B, 10)
    achieved = bpc * C / avg_f if avg_f > 0 else 0.0
    traffic = measured_traffic(dom, C)
    whole = (B["total"] * C * s * args.steps + relax_bytes(10) * C * args.steps) / mixed["el"]
    # the pass F family (residual stage, fused stage, last stage with the relaxation)
    fam = [k for k in ("pass_f_rhs", "pass_f", "pass_f_relax") if k in kt]
    fam_bytes = sum(kernel_bytes(k, B, 10) * C * kt[k][0] for k in fam)
    fam_s = sum(kt[k][1] for k in fam) / 1e3
    has_chem = head == "mixed_mlp" and mixed["info"].get("relaxation", {}).get("n_warm", 0) > 0
    out = {
        "metric": "cell-RK-stage updates/s, 1 MI355X: mixed+surrogate vs FP64; % of HBM peak",
        "value": mixed["rate"],
        "unit": "cell-RK-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["el"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"shock-bubble {args.n}^3 jittered hexes x 6 tets ({C} cells), 64 vapour bubbles, "
                               f"WALE, SSP-RK3, CFL 0.1, mixed precision, "
                               + ("surrogate relaxation (exact classification, analytic thermal Newton, MLP warm "
                                  "start + FP64 polish for chemical cells: "
                                  + ("present" if has_chem else "none in the timed steps") + ")"
                                  if head == "mixed_mlp" else "plain Newton relaxation"),
                   "cells": C, "faces": case.mesh.n_faces, "cell_order": args.order, "fuse_update": args.fuse,
                   "stage_f": args.stage_f, "stage_g": args.stage_g, "fuse_relax": 1},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": dom, "bytes_per_cell": bpc, "avg_launch_ms": avg_f * 1e3},
        "pass_f_family": {"kernels": fam, "achieved_GBps": fam_bytes / fam_s / 1e9 if fam_s else None,
                          "frac": fam_bytes / fam_s / HBM_PEAK if fam_s else None,
                          "bytes_per_cell": {k: kernel_bytes(k, B, 10) for k in fam}},
        "whole_step_hbm_frac": whole / HBM_PEAK,
        "whole_step_traffic": None,
        "stream_copy": copy,
    }
    if head == "mixed_mlp":
        out["surrogate_training"] = {k: v for k, v in mixed["info"].items() if k != "relaxation"}
    tr = traffic_rate(kt, mixed["el"], C, head)
    if tr is not None:
        out["whole_step_traffic"] = {"GBps": tr / 1e9, "frac_of_peak": tr / HBM_PEAK,
                                     "frac_of_copy": tr / (copy["GBps"] * 1e9) if copy and copy.get("GBps") else None,
                                     "source": "profiles/shock_bubble_48M_traffic.json bytes per launch x launches / "
                                               "wall time; includes over-fetch and Infinity Cache hits"}
    out["relaxation"] = {k: r["info"].get("relaxation") for k, r in res.items()}
    for other in ("mixed_cold", "mixed"):
        if other in res and other != head:
            out[other] = {"value": res[other]["rate"], "ms_per_step": res[other]["el"] / args.steps * 1e3}
    if "mixed_cold" in res and head == "mixed_mlp":
        out["speedup_network_vs_cold_start"] = mixed["rate"] / res["mixed_cold"]["rate"]
    if "fp64" in res:
        f64 = res["fp64"]
        Bd = stage_bytes(10, s, fc, 8, 8)
        out["fp64"] = {"value": f64["rate"], "ms_per_step": f64["el"] / args.steps * 1e3,
                       "whole_step_hbm_frac": (Bd["total"] * C * s * args.steps + relax_bytes(10) * C * args.steps)
                       / f64["el"] / HBM_PEAK}
        tr64 = traffic_rate(f64["kt"], f64["el"], C, "fp64")
        if tr64 is not None:
            out["fp64"]["whole_step_traffic_frac_of_peak"] = tr64 / HBM_PEAK
        out["speedup_mixed_vs_fp64"] = mixed["rate"] / f64["rate"]
        out["mixed_vs_fp64_gpu_rel_l1"] = rel_errors(mixed["U"], f64["U"])
    del res, case
    if not args.no_cpu:
        log("cpu baseline ...")
        base, (cs, osteps, Uo) = cpu_baseline()
        out["cpu_baseline"] = base
        log(f"cpu baseline {base['value']:.4g} cell-stage/s")
        if not args.no_extra:
            # the GPU legs on the oracle's sample, same steps: errors against the FP64 oracle
            errs = {}
            for name, prec, mode in (("mixed_mlp", MIXED, RELAX_MLP), ("mixed", MIXED, RELAX_PLAIN),
                                     ("fp64", FP64, RELAX_PLAIN)):
                if name == "mixed_mlp" and args.no_mlp:
                    continue
                tr = None
                if name == "mixed_mlp":
                    tr = dict(n_solver=cs.mesh.n_cells, n_uniform=cs.mesh.n_cells // 4, warm_steps=3,
                              train_steps=args.train_steps)
                _, _, Ug, _, _ = run_gpu(cs, prec, osteps, 0, relax_mode=mode, train=tr)
                errs[name] = rel_errors(Ug, Uo)
                log(f"{name} vs FP64 oracle on the sample: {errs[name]}")
            out["vs_fp64_oracle_rel_l1"] = {"sample": base["sample"], **errs}
            try:
                out["cpu_configs"] = cpu_configs()
            except MemoryError:
                log("cpu configs skipped: out of host memory")
    if not args.no_extra and not args.no_mlp:
        out["relaxation_microbench"] = relax_microbench()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
