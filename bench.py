#!/usr/bin/env python3
"""Benchmark: Davidson time to converge 5 SF-CPSd roots on one MI355X,
mixed-precision MFMA vs FP64 MFMA, on the seeded synthetic D2h Hamiltonian with
120 spatial orbitals [30,8,14,18,26,6,10,8], n_alpha=12, n_beta=10
(22 occupied / 218 virtual spin orbitals, |R1|+|R2| = 699 455).

One "step" is one complete Davidson solve from the spin-flip guesses with all
integrals already resident in HBM.  The FP64 run uses FP64 MFMA for every
contraction; the mixed run uses FP32 MFMA (FP32 operands, FP32 accumulation)
for every contraction while the Davidson vectors, orthogonalisation, residuals,
preconditioner and subspace eigensolve stay FP64, and finishes with one batched
FP64 sigma of the converged vectors (FP64 Rayleigh quotients).  `value` is the
mixed-precision seconds per solve; the FP64 time, the speedup and the per-root
|w_mixed - w_fp64| are reported beside it.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: this benchmark is one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "revised-cluster-perturbation-single-double_amd"), os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "golden")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

WORKLOADS = {
    "c2v-24": dict(nso=(11, 2, 4, 7), nalpha=6, nbeta=4, seed=1, nroot=3, group="C2v"),
    "d2h-120": dict(nso=(30, 8, 14, 18, 26, 6, 10, 8), nalpha=12, nbeta=10, seed=2, nroot=5, group="D2h"),
    # block Davidson: 16 roots of the D2h/120 Hamiltonian
    "d2h-120-16roots": dict(nso=(30, 8, 14, 18, 26, 6, 10, 8), nalpha=12, nbeta=10, seed=2, nroot=16, group="D2h"),
    # mixed precision at C2v/180 (|R1|+|R2| = 5 981 115)
    "c2v-180": dict(nso=(70, 18, 36, 56), nalpha=16, nbeta=14, seed=3, nroot=5, group="C2v"),
    # gradient of the lowest SF-CPSd excitation energy w.r.t. the orbital
    # energies and the integral blocks, checked by central differences
    "c2v-100-gradient": dict(nso=(40, 10, 20, 30), nalpha=10, nbeta=8, seed=5, nroot=1, group="C2v", kind="gradient"),
    # 32 independent C2v Hamiltonians (seeds 0..31): 30-80 spatial orbitals with a
    # random irrep split, n_alpha 4..10, n_beta = n_alpha - 2, 3 roots each
    "c2v-batch32": dict(nham=32, norb=(30, 80), nalpha=(4, 10), nroot=3, group="C2v", kind="batch"),
}


def batch_specs(cfg):
    """the seeded Hamiltonians of the batch workload"""
    import numpy as np

    import dpdhip
    out = []
    for s in range(cfg["nham"]):
        rng = np.random.default_rng(s)
        norb = int(rng.integers(cfg["norb"][0], cfg["norb"][1] + 1))
        w = rng.dirichlet([3.0] * 4)
        nso = tuple(int(x) + 1 for x in rng.multinomial(norb - 4, w))
        na = int(rng.integers(cfg["nalpha"][0], cfg["nalpha"][1] + 1))
        out.append(dpdhip.Spec(nso, na, na - 2, s))
    return out
# MI355X dense MFMA peaks per ALGORITHMIC flop: FP64, FP32, and BF16x3 (three
# BF16 MFMAs of the 2.5 PFLOP/s dense BF16 peak per algorithmic product)
PEAK_TFLOPS = {0: 78.6, 1: 157.3, 2: 2500.0 / 3}
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="d2h-120", choices=sorted(WORKLOADS))
    ap.add_argument("--conv", type=float, default=1e-7, help="FP64 max-abs residual tolerance")
    ap.add_argument("--conv-mixed", type=float, default=1e-5, help="mixed-precision residual tolerance")
    ap.add_argument("--max-dim", type=int, default=40)
    ap.add_argument("--max-cyc", type=int, default=200)
    ap.add_argument("--tracking", choices=["lowest", "overlap"], default="lowest")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", default=None, choices=["fp64", "fp32", "bf16x3"],
                    help="counter-run mode: a few batched sigma calls in this precision (tile shape --tile), "
                         "then a plain 512 MiB device copy for calibrating FETCH_SIZE/WRITE_SIZE; no JSON line")
    ap.add_argument("--tile", type=int, default=-1,
                    help="GEMM tile configuration for --probe and --probe-gathers (-1: autotuned)")
    ap.add_argument("--split", type=int, default=-1,
                    help="split-K for --probe and --probe-gathers: -1 automatic, 0 off, 1 on")
    ap.add_argument("--profile-solve", default=None, choices=["fp64", "fp32", "bf16x3"],
                    help="trace mode: a warm-up solve in this precision, then one solve between two marker "
                         "kernels (k_precond on one element) for rocprofv3 --kernel-trace; no JSON line")
    ap.add_argument("--summarize-trace", default=None, metavar="CSV",
                    help="summarise the kernel_trace.csv of a --profile-solve run (the kernels between the "
                         "markers, grouped by class) as markdown and exit")
    ap.add_argument("--probe-gathers", default=None, choices=["fp64", "fp32", "bf16x3"],
                    help="counter-run mode for the gathers and sorts: one profiled sigma call (5 vectors) in this "
                         "precision, then the DPD sorts at this workload's sizes, each measured call between two "
                         "marker kernels, then a plain 512 MiB device copy; writes the launches' algorithmic "
                         "bytes to --probe-out; no JSON line")
    ap.add_argument("--probe-out", default=None,
                    help="launch list (JSON) written by --probe and --probe-gathers (default for the latter: "
                         "gather_probe.json)")
    ap.add_argument("--summarize-kernels", nargs=3, default=None, metavar=("TRACE_CSV", "KERNEL", "GRID"),
                    help="markdown of a kernel_trace.csv: every kernel's dispatches and time, then the dispatches "
                         "of KERNEL with grid size GRID (one plan's launches) with their mean and median, and exit")
    ap.add_argument("--summarize-gemms", nargs=3, default=None, metavar=("TRACE_CSV", "PMC_CSV", "PROBE_JSON"),
                    help="join the kernel trace and an SQ counter run of a --probe program with its launch list: "
                         "per GEMM launch of the profiled sigma call, algorithmic and counted MFMA throughput and "
                         "MFMA busy share, as markdown, and exit")
    ap.add_argument("--summarize-gathers", nargs=4, default=None,
                    metavar=("TRACE_CSV", "FETCH_CSV", "WRITE_CSV", "PROBE_JSON"),
                    help="join the kernel trace and the FETCH_SIZE / WRITE_SIZE counter runs of a --probe-gathers "
                         "program with its launch list: achieved and calibrated-traffic GB/s per gather class "
                         "and per sort, as markdown, and exit")
    ap.add_argument("--fd-dirs", type=int, default=50, help="gradient workload: random integral directions")
    ap.add_argument("--fd-h", type=float, default=1e-4, help="gradient workload: central-difference step")
    ap.add_argument("--grad-conv", type=float, default=1e-9, help="gradient workload: FP64 residual tolerance")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this "
              "benchmark", file=sys.stderr)
        sys.exit(2)

    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.summarize_gathers:
        summarize_gathers(*args.summarize_gathers)
        return
    if args.summarize_gemms:
        summarize_gemms(*args.summarize_gemms)
        return
    if args.summarize_kernels:
        summarize_kernels(*args.summarize_kernels)
        return

    import numpy as np
    import torch

    import dpdhip

    cfg = WORKLOADS[args.workload]
    nroot = cfg["nroot"]
    if args.probe:
        probe(cfg, args)
        return
    if args.probe_gathers:
        probe_gathers(cfg, args)
        return
    if args.profile_solve:
        profile_solve(cfg, args)
        return
    if cfg.get("kind") == "gradient":
        gradient_bench(cfg, args)
        return
    if cfg.get("kind") == "batch":
        batch_bench(cfg, args)
        return
    lowest = args.tracking == "lowest"
    t0 = time.time()
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    spec = dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
    prob = dpdhip.Problem(ctx, spec, "sf-cpsd", 1)
    guess, _ = prob.guess(nroot)
    log(f"problem ready: dim={prob.dim} setup {time.time() - t0:.1f}s")

    def solve(prec, conv, refine):
        prob.set_precision(prec)
        R = guess.clone()
        w, R, info = prob.davidson(nroot, conv=conv, max_dim=args.max_dim, max_cyc=args.max_cyc, lowest=lowest,
                                   final_check=False, R=R)
        if refine:
            # FP64 Rayleigh quotients of the converged vectors: one batched FP64 sigma
            prob.set_precision("fp64")
            S = prob.sigma(R)
            Rm = R.view(nroot, -1)
            w = (Rm * S.view(nroot, -1)).sum(1) / (Rm * Rm).sum(1)
            w = w.cpu().numpy()
        return np.asarray(w), info

    def timed(prec, conv, refine, steps, warmup):
        for _ in range(warmup):
            solve(prec, conv, refine)
        torch.cuda.synchronize()
        ts = []
        out = None
        for _ in range(steps):
            torch.cuda.synchronize()
            a = time.perf_counter()
            out = solve(prec, conv, refine)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - a)
        return ts, out

    ts64, (w64, info64) = timed("fp64", args.conv, False, args.steps, args.warmup)
    log(f"fp64: {np.median(ts64):.3f}s/solve iterations={info64['iterations']} conv={info64['converged']} w={w64}")
    # the same code in FP64 at the mixed run's tolerance (equal-work comparison)
    ts64b, (w64b, info64b) = timed("fp64", args.conv_mixed, True, args.steps, args.warmup)
    log(f"fp64@conv_mixed: {np.median(ts64b):.3f}s/solve iterations={info64b['iterations']} w={w64b}")
    # mixed precision: FP32 MFMA and BF16x3 MFMA contractions; the timed region
    # also measures every GEMM launch with HIP events on the stream it runs on
    mixed = {}
    for prec in ("fp32", "bf16x3"):
        for _ in range(args.warmup):
            solve(prec, args.conv_mixed, True)
        ctx.profile(True, 200000)
        ts, (wm, infom) = timed(prec, args.conv_mixed, True, args.steps, 0)
        prof = ctx.profile_report()
        ctx.profile(False)
        dw = float(np.abs(wm - w64).max())
        mixed[prec] = dict(t=float(np.median(ts)), w=wm, info=infom, prof=prof, dw=dw)
        log(f"{prec}: {np.median(ts):.3f}s/solve iterations={infom['iterations']} conv={infom['converged']} "
            f"max|dw|={dw:.2e}")
    ok = {k: v for k, v in mixed.items() if v["info"]["converged"] and v["dw"] <= 1e-6}
    # a mixed run that did not converge or missed the 1e-6 Eh gate is no
    # measurement of the metric: value is null then, the runs stay in mixed_all
    gate_failed = not ok
    best = min(ok or mixed, key=lambda k: mixed[k]["t"])
    M = mixed[best]
    t_fp64 = float(np.median(ts64))
    t_mix = M["t"]
    infom, prof = M["info"], M["prof"]

    # dominant GEMM kernel of the measured mixed run; it is MFMA-bound when its
    # algorithmic intensity (flops per algorithmic byte) exceeds the ridge point
    # peak_flops / 8 TB/s, else HBM-bound
    gemm_prof = {k: v for k, v in prof.items() if v.get("kind", "gemm") == "gemm"}
    gather_prof = {k: v for k, v in prof.items() if v.get("kind") == "gather"}
    dom = max(gemm_prof.items(), key=lambda kv: kv[1]["ms"])
    tag, st = dom
    avg_ms = st["ms"] / st["count"]
    flops_per_launch = st["flops_total"] / st["count"]
    bytes_per_launch = st.get("bytes_total", 0) / st["count"]
    intensity = flops_per_launch / max(bytes_per_launch, 1.0)
    ridge = PEAK_TFLOPS[st["prec"]] * 1e12 / (PEAK_HBM_GBS * 1e9)
    if intensity >= ridge:
        bound, unit, peak = "mfma", "TFLOP/s", PEAK_TFLOPS[st["prec"]]
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    else:
        bound, unit, peak = "hbm", "GB/s", PEAK_HBM_GBS
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        ent = json.load(open(tpath)).get(
            f"{args.workload}:{tag.rsplit(':t', 1)[0]}:cfg{st.get('cfg')}:split{st.get('split', 1)}")
        traffic = ent["traffic_bytes"] if ent else None

    cpu = None
    if not args.no_cpu and args.workload != "c2v-180":
        # (at C2v/180 one oracle sigma alone exceeds the bounded sample)
        cpu = cpu_baseline(cfg, args, info64["sigma_calls"], lowest)

    tiles = {1: (64, 64), 2: (128, 64), 3: (64, 128), 4: (128, 128), 5: (32, 32), 6: (64, 192), 7: (128, 192)}
    bm, bn = tiles.get(st.get("cfg", 1), (64, 64))
    mode = {0: 0, 1: 1, 2: 2}[st["prec"]]
    if st.get("packed", 1):
        kname = "k_gemm_pkd<192>" if st.get("cfg") == 7 else f"k_gemm_pk<{mode}, {bm}, {bn}>"
    else:
        kname = f"k_gemm2<{bm}, {bn}>"
    line = {
        "metric": f"Davidson time to {nroot} roots at <=1e-6 Eh: mixed-prec MFMA vs FP64 GPU"
        if nroot != 5 else "Davidson time to 5 roots at <=1e-6 Eh: mixed-prec MFMA vs FP64 GPU",
        "value": None if gate_failed else round(t_mix, 6),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": f"{best}-mfma/fp64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"{args.workload}: SF-CPSd {cfg['group']} {sum(cfg['nso'])} spatial orbitals "
                               f"{list(cfg['nso'])} n_alpha={cfg['nalpha']} n_beta={cfg['nbeta']}, {nroot} roots, "
                               f"dim {prob.dim}",
                   "davidson": {"tracking": args.tracking, "max_dim": args.max_dim, "max_cyc": args.max_cyc,
                                "conv_fp64": args.conv, "conv_mixed": args.conv_mixed}},
        "mixed_precision": best,
        "accuracy_gate_failed": gate_failed,
        "fp64_s_per_solve": round(t_fp64, 6),
        "fp64_same_tol_s_per_solve": round(float(np.median(ts64b)), 6),
        # headline ratio: the same code in FP64 stopped at the mixed run's
        # tolerance (the eigenvalues agree to far below 1e-6 Eh either way)
        "speedup_mixed_vs_fp64_same_tol": None if gate_failed else round(float(np.median(ts64b)) / t_mix, 3),
        "speedup_mixed_vs_fp64_at_conv_fp64": None if gate_failed else round(t_fp64 / t_mix, 3),
        "fp64_iterations": info64["iterations"],
        "mixed_iterations": infom["iterations"],
        "converged": bool(info64["converged"] and infom["converged"]),
        "w_fp64": [round(float(x), 10) for x in w64],
        "abs_dw_per_root": {k: [float(abs(a - b)) for a, b in zip(v["w"], w64)] for k, v in mixed.items()},
        "max_abs_dw_mixed_vs_fp64": M["dw"],
        "mixed_all": {k: {"s_per_solve": round(v["t"], 6), "iterations": v["info"]["iterations"],
                          "max_abs_dw": v["dw"]} for k, v in mixed.items()},
        "roofline": {"bound": bound, "kernel": f"{kname} ({tag})", "tile_cfg": st.get("cfg"),
                     "split_k": bool(st.get("split", 0)),
                     "achieved": round(achieved, 3), "peak": peak, "unit": unit, "frac": round(achieved / peak, 4),
                     "traffic": traffic, "flops_per_launch": int(flops_per_launch),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch),
                     "intensity_flop_per_byte": round(intensity, 2), "ridge_flop_per_byte": round(ridge, 2),
                     "tflops": round(flops_per_launch / (avg_ms * 1e-3) / 1e12, 3),
                     "avg_launch_ms": round(avg_ms, 5), "launches": st["count"]},
        "kernels": {k: {"launches": v["count"], "ms_total": round(v["ms"], 3),
                        "tflops": round(v["flops_total"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] > 0 else None}
                    for k, v in gemm_prof.items()},
        # expansions (x1/x2) and epilogues (epi) of each tagged contraction:
        # algorithmic bytes (input once, output once, twice when accumulated)
        # over the event-timed duration
        "gathers": {k: {"launches": v["count"], "ms_total": round(v["ms"], 3),
                        "gbs": round(v["bytes_total"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                    for k, v in gather_prof.items()},
        "gathers_total": {"ms_per_solve": round(sum(v["ms"] for v in gather_prof.values()) / max(args.steps, 1), 3),
                          "gbs": round(sum(v["bytes_total"] for v in gather_prof.values()) /
                                       max(sum(v["ms"] for v in gather_prof.values()) * 1e-3, 1e-12) / 1e9, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def batch_bench(cfg, args):
    """32 Hamiltonians solved in lockstep: every contraction of a Davidson step
    is one grouped GEMM launch for all of them, the step is replayed from a
    hipGraph, FP32 MFMA with FP64 Rayleigh quotients of the converged vectors;
    against the same batch in FP64 and against one-at-a-time FP64 solves."""
    import numpy as np
    import torch

    import dpdhip

    t0 = time.time()
    ctx = dpdhip.Context(4, 0)
    specs = batch_specs(cfg)
    probs = [dpdhip.Problem(ctx, sp, "sf-cpsd", 1) for sp in specs]
    nroot = cfg["nroot"]
    guesses = [p.guess(nroot)[0] for p in probs]
    dims = [p.dim for p in probs]
    log(f"{len(probs)} problems ready: dims {min(dims)}..{max(dims)} (sum {sum(dims)}) setup {time.time() - t0:.1f}s")
    kw = dict(max_dim=args.max_dim, max_cyc=args.max_cyc, lowest=args.tracking == "lowest", final_check=False)

    def batch(prec, conv, grouped=True, graph=True, refine=False):
        for p in probs:
            p.set_precision(prec)
        Rs = [g.clone() for g in guesses]
        torch.cuda.synchronize()
        a = time.perf_counter()
        ws, Rs, infos = dpdhip.davidson_batch(probs, nroot, conv=conv, Rs=Rs, grouped=grouped, graph=graph, **kw)
        if refine:
            for p in probs:
                p.set_precision("fp64")
            S = dpdhip.sigma_group(probs, Rs)
            ws = []
            for R, Sv in zip(Rs, S):
                Rm, Sm = R.view(nroot, -1), Sv.view(nroot, -1)
                ws.append(((Rm * Sm).sum(1) / (Rm * Rm).sum(1)).cpu().numpy())
        torch.cuda.synchronize()
        return time.perf_counter() - a, [np.asarray(w) for w in ws], infos

    def single(prec, conv):
        for p in probs:
            p.set_precision(prec)
        torch.cuda.synchronize()
        a = time.perf_counter()
        ws, its = [], []
        for p, g in zip(probs, guesses):
            w, R, info = p.davidson(nroot, conv=conv, R=g.clone(), **kw)
            ws.append(np.asarray(w))
            its.append(info["iterations"])
        torch.cuda.synchronize()
        return time.perf_counter() - a, ws, its

    def timed(fn, *a, **k):
        for _ in range(args.warmup):
            fn(*a, **k)
        ts, out = [], None
        for _ in range(args.steps):
            out = fn(*a, **k)
            ts.append(out[0])
        return float(np.median(ts)), out

    t_seq64, (_, w_seq64, it_seq64) = timed(single, "fp64", args.conv)
    log(f"one at a time fp64: {t_seq64:.3f}s")
    t_b64, (_, w_b64, inf_b64) = timed(batch, "fp64", args.conv)
    log(f"batched fp64: {t_b64:.3f}s steps {inf_b64[0]['batch']}")
    t_b32g, (_, w_b32g, inf_b32g) = timed(batch, "fp32", args.conv_mixed, True, False, True)
    log(f"batched fp32 grouped, no graph: {t_b32g:.3f}s")
    t_b32s, (_, w_b32s, inf_b32s) = timed(batch, "fp32", args.conv_mixed, False, False, True)
    log(f"batched fp32 one launch per problem: {t_b32s:.3f}s")
    t_b32, (_, w_b32, inf_b32) = timed(batch, "fp32", args.conv_mixed, True, True, True)
    # the grouped GEMMs' own times from one more (untimed) run with the context
    # profiler on: the graph steps then run eagerly, one event pair per launch
    ctx.profile(True, 200000)
    batch("fp32", args.conv_mixed, True, False, True)
    prof = ctx.profile_report()
    ctx.profile(False)
    log(f"batched fp32 grouped + graph + fp64 refine: {t_b32:.3f}s steps {inf_b32[0]['batch']}")
    dw = max(float(np.max(np.abs(np.sort(a) - np.sort(b)))) for a, b in zip(w_b32, w_seq64))
    dw64 = max(float(np.max(np.abs(a - b))) for a, b in zip(w_b64, w_seq64))
    conv_all = all(i["converged"] for i in inf_b32) and all(i["converged"] for i in inf_b64)
    line = {
        "metric": "time to 3 roots of each of 32 Hamiltonians: batched grouped FP32 MFMA + FP64 refine vs FP64",
        "value": round(t_b32, 6),
        "unit": "s/batch",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_b32, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32-mfma/fp64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"c2v-batch32: {len(probs)} SF-CPSd C2v Hamiltonians, {cfg['norb'][0]}-{cfg['norb'][1]} "
                               f"spatial orbitals (random irrep splits, seeds 0..{len(probs) - 1}), n_alpha "
                               f"{cfg['nalpha'][0]}-{cfg['nalpha'][1]}, n_beta = n_alpha - 2, {nroot} roots each",
                   "dims": dims, "nso": [list(sp.nso) for sp in specs], "nalpha": [sp.nalpha for sp in specs],
                   "davidson": {"tracking": args.tracking, "max_dim": args.max_dim, "max_cyc": args.max_cyc,
                                "conv_fp64": args.conv, "conv_mixed": args.conv_mixed}},
        "one_at_a_time_fp64_s": round(t_seq64, 6),
        "batched_fp64_s": round(t_b64, 6),
        "batched_fp32_grouped_no_graph_s": round(t_b32g, 6),
        "batched_fp32_per_problem_launches_s": round(t_b32s, 6),
        "speedup_vs_one_at_a_time_fp64": round(t_seq64 / t_b32, 3),
        "speedup_vs_batched_fp64": round(t_b64 / t_b32, 3),
        "graph": inf_b32[0]["batch"],
        "converged": conv_all,
        "max_abs_dw_mixed_vs_fp64": dw,
        "max_abs_dw_batched_fp64_vs_one_at_a_time": dw64,
        "iterations_one_at_a_time_fp64": it_seq64,
        "iterations_batched_fp32": [i["iterations"] for i in inf_b32],
        "grouped_gemm": {k: {"launches": v["count"], "ms_total": round(v["ms"], 3), "tile_cfg": v.get("cfg"),
                             "tflops": round(v["flops_total"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] > 0 else None}
                         for k, v in prof.items()},
    }
    print(json.dumps(line), flush=True)


def gradient_bench(cfg, args):
    """Gradient of the lowest excitation energy (right and left Davidson solves,
    then the adjoint contractions) in FP64 and FP32/BF16x3, timed; the FP64
    gradient checked against central differences of the eigenvalue along seeded
    random unit directions over the spin-allowed elements of the integral blocks
    (and a few in the orbital energies), and the mixed-precision gradients
    against the FP64 one."""
    import numpy as np
    import torch

    import dpdhip
    from dpdhip.gradient import excitation_energy

    t0 = time.time()
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    names = prob.INPUTS
    base = {k: prob.get(k) for k in names}
    log(f"problem ready: dim={prob.dim} setup {time.time() - t0:.1f}s")
    solve = dict(nroot=cfg["nroot"], conv=args.grad_conv, max_dim=args.max_dim, max_cyc=args.max_cyc)

    def run(prec, conv, x0=None, warm=None):
        prob.set_precision(prec)
        x = {k: (x0[k] if x0 else base[k]).clone().requires_grad_(True) for k in names}
        info = {}
        extra = {"R0": warm["R"], "L0": warm["L"]} if warm else {}
        torch.cuda.synchronize()
        a = time.perf_counter()
        w = excitation_energy(prob, x, info=info, **dict(solve, conv=conv), **extra)
        torch.cuda.synchronize()
        b = time.perf_counter()
        w.backward()
        torch.cuda.synchronize()
        c = time.perf_counter()
        return float(w.detach()), {k: x[k].grad for k in names}, info, b - a, c - b

    res = {}
    for prec, conv in (("fp64", args.grad_conv), ("fp32", args.conv_mixed), ("bf16x3", args.conv_mixed)):
        for _ in range(args.warmup):
            run(prec, conv)
        tf, tb = [], []
        for _ in range(args.steps):
            w, g, info, a, b = run(prec, conv)
            tf.append(a)
            tb.append(b)
        res[prec] = dict(w=w, g=g, info=info, forward=float(np.median(tf)), backward=float(np.median(tb)))
        log(f"{prec}: w={w:.12f} forward {np.median(tf):.3f}s backward {np.median(tb):.3f}s "
            f"iterations right/left {info['right']['iterations']}/{info['left']['iterations']}")
    g64 = res["fp64"]["g"]
    gnorm = sum(float(v.norm() ** 2) for v in g64.values()) ** 0.5
    gn_eri = sum(float(g64[k].norm() ** 2) for k in names if k != "eorb") ** 0.5
    for prec in ("fp32", "bf16x3"):
        err = sum(float((res[prec]["g"][k] - g64[k]).norm() ** 2) for k in names) ** 0.5
        res[prec]["grad_rel_err_vs_fp64"] = err / gnorm
        res[prec]["dw_vs_fp64"] = abs(res[prec]["w"] - res["fp64"]["w"])

    # central differences along seeded random unit directions
    gen = torch.Generator(device="cuda:0").manual_seed(cfg["seed"])
    h = args.fd_h
    warm = res["fp64"]["info"]
    fd_rows = []
    t_fd = time.time()
    ndir_e = 5
    for i in range(args.fd_dirs + ndir_e):
        on_eorb = i >= args.fd_dirs
        keys = ["eorb"] if on_eorb else [k for k in names if k != "eorb"]
        # supported on the spin-allowed elements (nonzero in the inputs): a change
        # of a spin-forbidden element couples other spin sectors, whose lowest
        # state is another one
        d = {k: torch.randn(base[k].numel(), dtype=torch.float64, device="cuda:0", generator=gen)
             * (base[k] != 0) for k in keys}
        nrm = sum(float(v.norm() ** 2) for v in d.values()) ** 0.5
        d = {k: v / nrm for k, v in d.items()}
        vals = []
        for sgn in (1.0, -1.0):
            xs = {k: (base[k] + sgn * h * d[k]) if k in d else base[k] for k in names}
            vals.append(run("fp64", args.grad_conv, xs, warm)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        an = sum(float(g64[k] @ d[k]) for k in keys)
        scale = gnorm if on_eorb else gn_eri
        fd_rows.append(dict(space="eorb" if on_eorb else "eri", fd=fd, analytic=an, abs_err=abs(fd - an),
                            err_over_gnorm=abs(fd - an) / scale))
        if (i + 1) % 10 == 0:
            log(f"directions {i + 1}/{args.fd_dirs + ndir_e} done ({time.time() - t_fd:.1f}s)")
    eri = [r for r in fd_rows if r["space"] == "eri"]
    eo = [r for r in fd_rows if r["space"] == "eorb"]
    line = {
        "metric": "gradient of the lowest SF-CPSd excitation energy (right+left Davidson + adjoint contractions)",
        "value": round(res["fp64"]["forward"] + res["fp64"]["backward"], 6),
        "unit": "s/gradient",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (res["fp64"]["forward"] + res["fp64"]["backward"]), 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"c2v-100-gradient: SF-CPSd {cfg['group']} {sum(cfg['nso'])} spatial orbitals {list(cfg['nso'])} "
                               f"n_alpha={cfg['nalpha']} n_beta={cfg['nbeta']}, lowest root, dim {prob.dim}",
                   "inputs": {k: int(base[k].numel()) for k in names}, "conv_fp64": args.grad_conv,
                   "conv_mixed": args.conv_mixed},
        "w_fp64": res["fp64"]["w"],
        "precisions": {p: {"forward_s": round(r["forward"], 6), "backward_s": round(r["backward"], 6),
                           "w": r["w"], "iterations_right": r["info"]["right"]["iterations"],
                           "iterations_left": r["info"]["left"]["iterations"],
                           **({"grad_rel_err_vs_fp64": r["grad_rel_err_vs_fp64"], "abs_dw_vs_fp64": r["dw_vs_fp64"]}
                              if p != "fp64" else {})} for p, r in res.items()},
        "grad_norm": {"all": gnorm, "eri": gn_eri, "eorb": float(g64["eorb"].norm())},
        "finite_difference": {"h": h, "eri_directions": len(eri), "eorb_directions": len(eo),
                              "max_abs_err_eri": max(r["abs_err"] for r in eri),
                              "max_err_over_gnorm_eri": max(r["err_over_gnorm"] for r in eri),
                              "median_abs_analytic_eri": float(np.median([abs(r["analytic"]) for r in eri])),
                              "max_abs_err_eorb": max(r["abs_err"] for r in eo),
                              "max_err_over_gnorm_eorb": max(r["err_over_gnorm"] for r in eo),
                              "seconds": round(time.time() - t_fd, 1)},
    }
    print(json.dumps(line), flush=True)


def probe(cfg, args):
    """The program profiled by the rocprofv3 --pmc runs: batched sigma calls of
    the benchmark problem (5 vectors) with a fixed GEMM tile shape, then a plain
    copy of 2^26 doubles whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch

    import dpdhip
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    ctx.set_gemm_config(args.tile, 1)
    ctx.set_split_k(args.split)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    prob.set_precision(args.probe)
    R = torch.randn(cfg["nroot"], prob.dim, device="cuda:0", dtype=torch.float64)
    for _ in range(2):
        prob.sigma(R)
    # the third call between two marker kernels, with the context profiler on
    one = torch.ones(1, dtype=torch.float64, device="cuda:0")
    out1 = torch.empty_like(one)
    torch.cuda.synchronize()
    ctx.profile(True)
    dpdhip.precondition(ctx, 0.0, one, one, out1)
    prob.sigma(R)
    dpdhip.precondition(ctx, 0.0, one, one, out1)
    rep = ctx.profile_report()
    ctx.profile(False)
    torch.cuda.synchronize()
    a = torch.empty(1 << 26, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    if args.probe_out:
        records = [[k, v["kind"], v["count"], v["bytes_total"], v["flops_total"], v["prec"]] for k, v in rep.items()]
        with open(args.probe_out, "w") as f:
            json.dump({"precision": args.probe, "workload": args.workload, "sigma_records": records,
                       "copy_bytes": 8 << 26}, f, indent=1)
    log("probe done")


def probe_gathers(cfg, args):
    """The program profiled for the gather and sort bandwidths: three warm
    sigma calls (codes built, tiles tuned, integral panels packed), one sigma
    call with the context profiler on (its per-launch algorithmic bytes), then
    each DPD sort at the workload's spin-orbital populations, called once to
    warm up and once more between two marker kernels (k_precond on one
    element), then a plain copy of 2^26 doubles that calibrates FETCH_SIZE and
    WRITE_SIZE.  The launch list goes to --probe-out as JSON."""
    import torch

    import dpdhip
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    # a fixed tile shape (--tile) keeps the launches of the trace and counter runs the same
    ctx.set_gemm_config(args.tile, 1)
    ctx.set_split_k(args.split)
    spec = dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"])
    prob = dpdhip.Problem(ctx, spec, "sf-cpsd", 1)
    prob.set_precision(args.probe_gathers)
    R = torch.randn(cfg["nroot"], prob.dim, device="cuda:0", dtype=torch.float64)
    for _ in range(3):
        prob.sigma(R)
    one = torch.ones(1, dtype=torch.float64, device="cuda:0")
    out1 = torch.empty_like(one)

    def marker():
        dpdhip.precondition(ctx, 0.0, one, one, out1)
    torch.cuda.synchronize()
    ctx.profile(True)
    marker()
    prob.sigma(R)
    marker()
    rep = ctx.profile_report()
    ctx.profile(False)
    # one sigma call: the report's keys are in first-launch order
    records = [[k, v["kind"], v["count"], v["bytes_total"]] for k, v in rep.items()]
    lay = spec.layout()
    V, Oc = lay["popv"], lay["popo"]
    sym = 1
    sorts = []

    def sort_case(name, n_in, n_out, call):
        x = torch.randn(n_in, dtype=torch.float64, device="cuda:0")
        y = torch.empty(n_out + 1, dtype=torch.float64, device="cuda:0")
        call(x, y)
        torch.cuda.synchronize()
        marker()
        call(x, y)
        marker()
        torch.cuda.synchronize()
        sorts.append([name, 8 * (n_in + n_out)])
        del x, y
    vovo = ctx.idsymsoc(sym, V, Oc, 1, V, Oc, 1)
    for t in ("1243", "2143", "3142", "2413", "4321"):
        sort_case(f"sstgenso '{t}' (V,O,V,O) -> permuted, {vovo} elements", vovo, vovo,
                  lambda x, y, t=t: dpdhip.sstgenso(ctx, x, y, V, Oc, V, Oc, sym, t))
    vvvo = ctx.idsymsoc(sym, V, V, 1, V, Oc, 1)
    for t in ("1243", "2134", "3412", "4321"):
        sort_case(f"sstgenso '{t}' (V,V,V,O) -> permuted, {vvvo} elements", vvvo, vvvo,
                  lambda x, y, t=t: dpdhip.sstgenso(ctx, x, y, V, V, V, Oc, sym, t))
    vvoo_p = ctx.idsymsoc(sym, V, V, 0, Oc, Oc, 0)
    vvoo_e = ctx.idsymsoc(sym, V, V, 1, Oc, Oc, 0)
    sort_case(f"iexpso ab packed -> full (V,V,O,O), {vvoo_p} -> {vvoo_e} elements", vvoo_p, vvoo_e,
              lambda x, y: dpdhip.iexpso(ctx, x, y, V, V, 0, 1, Oc, Oc, 0, 0, sym))
    sort_case(f"icompso ab full -> packed (V,V,O,O), {vvoo_e} -> {vvoo_p} elements", vvoo_e, vvoo_p,
              lambda x, y: dpdhip.icompso(ctx, x, y, V, V, 1, 0, Oc, Oc, 0, 1, sym))
    sort_case(f"itranspso (V,O|V,O) pair transpose, {vovo} elements", vovo, vovo,
              lambda x, y: dpdhip.itranspso(ctx, x, y, V, Oc, 1, V, Oc, 1, sym))
    a = torch.empty(1 << 26, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    with open(args.probe_out or "gather_probe.json", "w") as f:
        json.dump({"precision": args.probe_gathers, "workload": args.workload, "sigma_records": records,
                   "sorts": sorts, "copy_bytes": 8 << 26}, f, indent=1)
    log(f"probe done: {len(records)} sigma records, {len(sorts)} sorts -> {args.probe_out}")


def _trace_rows(path):
    """(start, end, name) of every dispatch of a rocprofv3 kernel_trace.csv, in start order"""
    import csv
    rows = list(csv.DictReader(open(path)))
    return sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in rows),
                  key=lambda x: x[0])


def _pmc_rows(path, counter):
    """(dispatch id, name, value, start, end) per dispatch of a rocprofv3 counter_collection.csv
    (rows of one dispatch and counter summed), in dispatch order"""
    import csv
    acc = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        e = acc.setdefault(d, [r["Kernel_Name"], 0.0, int(r["Start_Timestamp"]), int(r["End_Timestamp"])])
        e[1] += float(r["Counter_Value"])
    return [(d, e[0], e[1], e[2], e[3]) for d, e in sorted(acc.items())]


def _base(name):
    return name.split("(")[0].replace("void ", "").strip()


def _is_marker(name):
    return name.split("(")[0].strip().endswith("k_precond")


def _sections(names):
    """index ranges strictly between marker pairs (0,1), (2,3), ... of a dispatch-name list"""
    marks = [i for i, n in enumerate(names) if _is_marker(n)]
    return [(marks[k] + 1, marks[k + 1]) for k in range(0, len(marks) - 1, 2)]


def summarize_gathers(trace_csv, fetch_csv, write_csv, probe_json):
    """Markdown: per gather class of the profiled sigma call and per sort, the
    duration (kernel trace), the algorithmic bytes (the probe's launch list),
    the HBM traffic (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled
    by what it reads for the probe's plain copy) and both as GB/s."""
    meta = json.load(open(probe_json))
    tr = _trace_rows(trace_csv)
    fe = _pmc_rows(fetch_csv, "FETCH_SIZE")
    wr = _pmc_rows(write_csv, "WRITE_SIZE")
    kib = meta["copy_bytes"] / 1024.0
    # calibration: the last dispatch of each counter run is the plain copy
    f_scale = kib / fe[-1][2]
    w_scale = kib / wr[-1][2]
    st, sf, sw = _sections([r[2] for r in tr]), _sections([r[1] for r in fe]), _sections([r[1] for r in wr])
    nsort = len(meta["sorts"])
    if not (len(st) == len(sf) == len(sw) == 1 + nsort):
        raise SystemExit(f"marker sections differ: trace {len(st)}, fetch {len(sf)}, write {len(sw)}, "
                         f"expected {1 + nsort}")

    notes = []

    def dispatches(k):
        """(kernel, us, fetched bytes, written bytes) of section k; the durations come from the
        kernel trace, or from the FETCH_SIZE run when the trace launched other kernels there"""
        (a, b), (fa, fb), (wa, wb) = st[k], sf[k], sw[k]
        nf = [_base(fe[fa + i][1]) for i in range(fb - fa)]
        if nf != [_base(wr[wa + i][1]) for i in range(wb - wa)]:
            raise SystemExit(f"section {k}: the two counter runs launched different kernels")
        if nf == [_base(tr[a + i][2]) for i in range(b - a)]:
            dur = [(tr[a + i][1] - tr[a + i][0]) * 1e-3 for i in range(b - a)]
        else:
            notes.append(k)
            dur = [(fe[fa + i][4] - fe[fa + i][3]) * 1e-3 for i in range(fb - fa)]
        return [(nf[i], dur[i], fe[fa + i][2] * f_scale * 1024.0, wr[wa + i][2] * w_scale * 1024.0)
                for i in range(fb - fa)]
    print(f"Precision {meta['precision']}, workload {meta['workload']}.  Calibration copy of "
          f"{meta['copy_bytes'] / 2**20:.0f} MiB: FETCH_SIZE reads {fe[-1][2]:.0f} KiB (scale {f_scale:.3f}), "
          f"WRITE_SIZE {wr[-1][2]:.0f} KiB (scale {w_scale:.3f}).\n")
    # the sigma call: gather records <-> k_gather dispatches, pack records <-> k_pack dispatches, in order
    sig = dispatches(0)
    recs = []
    for key, kind, cnt, by in meta["sigma_records"]:
        if kind != "gather":
            continue
        for _ in range(cnt):
            recs.append((key, by / cnt))
    g_disp = [d for d in sig if "k_gather" in d[0] or "k_epi" in d[0]]
    p_disp = [d for d in sig if "k_pack" in d[0]]
    g_rec = [r for r in recs if not r[0].endswith(":pack")]
    p_rec = [r for r in recs if r[0].endswith(":pack")]
    if len(g_disp) != len(g_rec) or len(p_disp) != len(p_rec):
        raise SystemExit(f"sigma call: {len(g_disp)} gather / {len(p_disp)} pack dispatches against "
                         f"{len(g_rec)} / {len(p_rec)} profiler records")
    rows = [(r[0], r[1], d) for r, d in zip(g_rec + p_rec, g_disp + p_disp)]

    def cls(key):
        suf = key.rsplit(":", 1)[-1]
        return {"x1": "expansions", "x2": "expansions", "epi": "antisymmetrising epilogues",
                "pack": "operand packing"}.get(suf, suf)
    agg = {}
    for key, by, d in rows:
        e = agg.setdefault(cls(key), [0, 0.0, 0.0, 0.0])
        e[0] += 1
        e[1] += d[1]
        e[2] += by
        e[3] += d[2] + d[3]

    def line(name, n, us, alg, traffic):
        gbs = alg / (us * 1e-6) / 1e9 if us > 0 else 0.0
        tgbs = traffic / (us * 1e-6) / 1e9 if us > 0 else 0.0
        return (f"| {name} | {n} | {us:.1f} | {alg / 1e6:.1f} | {gbs:.0f} | {gbs / PEAK_HBM_GBS:.2f} | "
                f"{traffic / 1e6:.1f} | {tgbs:.0f} | {tgbs / PEAK_HBM_GBS:.2f} |")
    head = ("| {0} | launches | us | algorithmic MB | GB/s | of 8 TB/s | traffic MB | traffic GB/s | "
            "of 8 TB/s |\n|---|---:|---:|---:|---:|---:|---:|---:|---:|")
    print("One sigma call, 5 vectors, per gather class:\n")
    print(head.format("class"))
    for k, (n, us, alg, traffic) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(line(k, n, us, alg, traffic))
    print("\nThe same launches one by one (largest first):\n")
    print(head.format("launch (kernel)"))
    for key, by, d in sorted(rows, key=lambda r: -r[2][1]):
        print(line(f"{key} (`{d[0].replace('void ', '')}`)", 1, d[1], by, d[2] + d[3]))
    print("\nDPD sorts at the workload's spin-orbital populations (the second of two identical calls):\n")
    print(head.format("sort (kernel)"))
    for k, (name, by) in enumerate(meta["sorts"]):
        ds = dispatches(1 + k)
        if len(ds) != 1:
            raise SystemExit(f"sort {name}: {len(ds)} dispatches")
        d = ds[0]
        print(line(f"{name} (`{d[0].replace('void ', '')}`)", 1, d[1], by, d[2] + d[3]))
    if notes:
        print(f"\nSections {notes} launched other kernels in the kernel-trace run (the autotuner times tile "
              "shapes); their durations are the FETCH_SIZE run's own timestamps.")


def summarize_kernels(trace_csv, kernel, grid, top=25):
    """Markdown of a whole kernel trace: the `top` kernels by total time, then
    the dispatches of one kernel name with one grid size (the launches of one
    plan) with their count, mean and median duration."""
    import csv
    import statistics
    rows = list(csv.DictReader(open(trace_csv)))
    per = {}
    sel = []
    total = 0.0
    for r in rows:
        name = _base(r["Kernel_Name"])
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        total += us
        e = per.setdefault(name, [0, 0.0])
        e[0] += 1
        e[1] += us
        g = int(r["Grid_Size"]) if "Grid_Size" in r else int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        if name == kernel and g == int(grid):
            sel.append(us)
    print(f"Total kernel time: {total / 1e3:.1f} ms over {len(rows)} dispatches.\n")
    print("| kernel | calls | total ms | avg us | % |\n|---|---:|---:|---:|---:|")
    for name, (cnt, us) in sorted(per.items(), key=lambda kv: -kv[1][1])[:top]:
        print(f"| `{name}` | {cnt} | {us / 1e3:.2f} | {us / cnt:.2f} | {100 * us / total:.1f} |")
    if sel:
        print(f"\n`{kernel}` with grid {grid}: {len(sel)} dispatches, mean {statistics.mean(sel):.1f} us, "
              f"median {statistics.median(sel):.1f} us.")
    else:
        print(f"\nno dispatch of `{kernel}` with grid {grid}")


def summarize_gemms(trace_csv, pmc_csv, probe_json, clock_ghz=2.4, simds=1024):
    """Markdown: per GEMM launch of a --probe program's profiled sigma call,
    the duration (kernel trace), the algorithmic TFLOP/s against the precision's
    dense MFMA peak, and the SQ counters of the counter run with two derived
    figures: the MFMA busy share per SIMD (SQ_VALU_MFMA_BUSY_CYCLES over
    simds x duration x clock) and the counted MFMA TFLOP/s
    (SQ_INSTS_VALU_MFMA_MOPS_* x 512 flops over the duration)."""
    import csv
    meta = json.load(open(probe_json))
    tr = _trace_rows(trace_csv)
    per = {}
    for r in csv.DictReader(open(pmc_csv)):
        d = int(r["Dispatch_Id"])
        e = per.setdefault(d, {"name": r["Kernel_Name"], "c": {}, "t": (int(r["Start_Timestamp"]),
                                                                          int(r["End_Timestamp"]))})
        e["c"][r["Counter_Name"]] = e["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    pm = [(d, per[d]["name"], per[d]["c"], per[d]["t"]) for d in sorted(per)]
    st, sp = _sections([r[2] for r in tr]), _sections([r[1] for r in pm])
    if not st or not sp:
        raise SystemExit("no marker pair in the trace or the counter run")
    (a, b), (pa, pb) = st[0], sp[0]
    names = [_base(pm[pa + i][1]) for i in range(pb - pa)]
    # durations from the kernel trace when it launched the same kernels, else the counter run's own
    same = names == [_base(tr[a + i][2]) for i in range(b - a)]
    dur = [((tr[a + i][1] - tr[a + i][0]) if same else (pm[pa + i][3][1] - pm[pa + i][3][0])) * 1e-3
           for i in range(pb - pa)]

    def is_gemm(n):
        return "k_gemm" in n and "reduce" not in n
    disp = [(names[i], dur[i], pm[pa + i][2]) for i in range(pb - pa) if is_gemm(names[i])]
    recs = []
    for key, kind, cnt, _by, fl, prec in meta["sigma_records"]:
        if kind == "gemm":
            recs += [(key, fl / cnt, prec)] * cnt
    if len(disp) != len(recs):
        raise SystemExit(f"{len(disp)} GEMM dispatches against {len(recs)} profiler records")
    cnames = sorted({k for _, _, c in disp for k in c})
    mops = [k for k in cnames if k.startswith("SQ_INSTS_VALU_MFMA_MOPS")]
    print(f"Precision {meta['precision']}, workload {meta['workload']}; one sigma call (5 vectors).  Peaks per "
          f"algorithmic flop: FP64 78.6, FP32 157.3, BF16x3 833 TFLOP/s (a third of 2.5 PFLOP/s).  MFMA busy = "
          f"SQ_VALU_MFMA_BUSY_CYCLES / ({simds} SIMDs x duration x {clock_ghz} GHz).  Durations from the "
          + ("kernel trace." if same else "counter run's timestamps (the kernel-trace run tuned other tiles).")
          + "\n")
    hdr = ["launch", "kernel", "us", "GFLOP", "TFLOP/s", "of peak"]
    if "SQ_VALU_MFMA_BUSY_CYCLES" in cnames:
        hdr.append("MFMA busy")
    if mops:
        hdr.append("counted TFLOP/s")
    hdr += cnames
    print("| " + " | ".join(hdr) + " |\n|" + "---|" * 2 + "---:|" * (len(hdr) - 2))
    tot_us = tot_fl = 0.0
    for (kname, us, c), (key, fl, prec) in zip(disp, recs):
        tot_us += us
        tot_fl += fl
        tf = fl / (us * 1e-6) / 1e12
        row = [key, "`" + kname + "`", f"{us:.1f}", f"{fl / 1e9:.2f}", f"{tf:.1f}", f"{tf / PEAK_TFLOPS[prec]:.2f}"]
        if "SQ_VALU_MFMA_BUSY_CYCLES" in cnames:
            row.append(f"{c.get('SQ_VALU_MFMA_BUSY_CYCLES', 0.0) / (simds * us * 1e-6 * clock_ghz * 1e9):.2f}")
        if mops:
            row.append(f"{sum(c.get(k, 0.0) for k in mops) * 512 / (us * 1e-6) / 1e12:.1f}")
        row += [f"{c.get(k, 0.0):.0f}" for k in cnames]
        print("| " + " | ".join(row) + " |")
    if tot_us > 0:
        print(f"\nAll GEMM launches of the call: {tot_us:.1f} us, {tot_fl / 1e9:.2f} GFLOP, "
              f"{tot_fl / (tot_us * 1e-6) / 1e12:.1f} TFLOP/s.")


def profile_solve(cfg, args):
    """The program traced for the per-iteration split: one warm-up solve (GEMM
    tiles tuned, integral panels packed), then exactly one solve of the same
    precision between two marker kernels; summarize_trace() reads the trace."""
    import torch

    import dpdhip
    ctx = dpdhip.Context(len(cfg["nso"]), 0)
    prob = dpdhip.Problem(ctx, dpdhip.Spec(cfg["nso"], cfg["nalpha"], cfg["nbeta"], cfg["seed"]), "sf-cpsd", 1)
    nroot = cfg["nroot"]
    guess, _ = prob.guess(nroot)
    conv = args.conv if args.profile_solve == "fp64" else args.conv_mixed
    prob.set_precision(args.profile_solve)
    kw = dict(conv=conv, max_dim=args.max_dim, max_cyc=args.max_cyc, lowest=args.tracking == "lowest",
              final_check=False)
    prob.davidson(nroot, R=guess.clone(), **kw)
    one = torch.ones(1, dtype=torch.float64, device="cuda:0")
    out = torch.empty_like(one)
    R = guess.clone()
    torch.cuda.synchronize()
    dpdhip.precondition(ctx, 0.0, one, one, out)  # marker: start
    a = time.perf_counter()
    w, R, info = prob.davidson(nroot, R=R, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter() - a
    dpdhip.precondition(ctx, 0.0, one, one, out)  # marker: end
    torch.cuda.synchronize()
    log(f"profiled solve: {args.profile_solve} conv {conv:g} {info['iterations']} iterations "
        f"{info['sigma_calls']} sigma calls {t:.4f}s")


# kernel classes of the per-iteration split (substring of the kernel name)
KERNEL_CLASSES = [
    ("GEMM (MFMA)", ("k_gemm_pk", "k_gemm2", "k_gemm<", "k_splitk_reduce")),
    ("operand packing", ("k_pack",)),
    ("gathers (sorts, expansions, epilogues)", ("k_gather", "k_epi_fused", "k_sort")),
    ("Davidson vector algebra", ("k_multidot", "k_reduce_rows", "k_lincomb", "k_dots_", "k_cgs_need",
                                 "k_tri_apply", "k_mgs", "k_pair_dots", "k_renorm_pair",
                                 "k_resid", "k_amax", "k_precond_batched", "k_scale_by_norm", "k_block_dots",
                                 "k_davidson", "k_ritz_resid", "k_ritz_final", "k_cgs_update",
                                 "k_compact_rows")),
    ("elementwise sigma terms, copies", ("k_hadamard", "k_div", "k_orbital_energy", "k_convert", "k_copy",
                                         "k_axpy", "k_scal", "k_fill", "k_singles", "copyBuffer")),
]


def summarize_trace(path):
    """Markdown summary of the kernels between the two k_precond markers of a
    --profile-solve trace: time per class and per kernel, and the idle time of
    the GPU inside the window (window length minus kernel time)."""
    import csv
    rows = list(csv.DictReader(open(path)))
    name_key = "Kernel_Name"
    recs = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r[name_key]) for r in rows),
                  key=lambda x: x[0])

    def is_marker(n):
        return n.split("(")[0].strip().endswith("k_precond")
    marks = [i for i, r in enumerate(recs) if is_marker(r[2])]
    if len(marks) < 2:
        raise SystemExit("no marker pair in the trace")
    i0, i1 = marks[-2], marks[-1]
    win = recs[i0 + 1:i1]
    t0, t1 = recs[i0][1], recs[i1][0]
    span = (t1 - t0) * 1e-3  # us
    per_kernel = {}
    for a, b, n in win:
        base = n.split("(")[0].strip()
        e = per_kernel.setdefault(base, [0, 0.0])
        e[0] += 1
        e[1] += (b - a) * 1e-3
    cls = {c: 0.0 for c, _ in KERNEL_CLASSES}
    cls["other"] = 0.0
    for base, (cnt, us) in per_kernel.items():
        for c, keys in KERNEL_CLASSES:
            if any(k in base for k in keys):
                cls[c] += us
                break
        else:
            cls["other"] += us
    busy = sum(cls.values())
    print(f"window: {span / 1e3:.3f} ms between the markers, {len(win)} kernels, busy {busy / 1e3:.3f} ms, "
          f"idle {(span - busy) / 1e3:.3f} ms\n")
    print("| class | ms | share |\n|---|---:|---:|")
    for c, us in sorted(cls.items(), key=lambda kv: -kv[1]):
        if us > 0:
            print(f"| {c} | {us / 1e3:.3f} | {100 * us / span:.1f}% |")
    print(f"| GPU idle inside the window | {(span - busy) / 1e3:.3f} | {100 * (span - busy) / span:.1f}% |\n")
    print("| kernel | launches | ms total | us avg |\n|---|---:|---:|---:|")
    for base, (cnt, us) in sorted(per_kernel.items(), key=lambda kv: -kv[1][1])[:30]:
        print(f"| `{base}` | {cnt} | {us / 1e3:.3f} | {us / cnt:.1f} |")


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, args, fp64_sigma_calls, lowest):
    """The FP64 oracle (C restatement of the reference engines + OpenBLAS +
    numpy restatement of mod_davidson.f90) on the host cores, on a bounded
    sample: the same Davidson from the same guesses runs for ~cpu-seconds of
    wall time; after its first iteration (warm-up) the sample is cut into three
    windows of whole iterations, the median cost per sigma call of the windows
    is taken, and the full solve is priced at the FP64 GPU solve's sigma-call
    count (the same iteration semantics: the GPU reproduces the oracle's
    iteration count)."""
    import numpy as np

    import oracle as O
    import oracle_sigma as OS
    from make_oracle_references import oracle_problem

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    # the box gives one GPU a share of its host cores, announced through
    # OMP_NUM_THREADS; without it every core the process may run on is used
    cores = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    cores = max(1, min(cores, avail))
    O.set_threads(cores)
    op = oracle_problem(cfg, "sf-cpsd", 1)
    g, _ = op.guess(cfg["nroot"])
    calls = {"n": 0}

    def sig(x):
        calls["n"] += 1
        return op.sigma(x)

    t_start = time.perf_counter()
    marks = []  # (time, sigma calls) at the end of every iteration
    budget = args.cpu_seconds

    class Stop(Exception):
        pass

    def logcb(it, w, res):
        marks.append((time.perf_counter(), calls["n"]))
        if time.perf_counter() - t_start > budget and len(marks) >= 4:
            raise Stop()

    try:
        OS.davidson(sig, op.diag(), g, cfg["nroot"], conv=args.conv, max_dim=args.max_dim, max_cyc=args.max_cyc,
                    lowest=lowest, final_check=False, log=logcb)
    except Stop:
        pass
    wall = time.perf_counter() - t_start
    # three windows of whole iterations after the first (the warm-up)
    per_call = []
    last = len(marks) - 1
    if last >= 3:
        idx = [round(q * last / 3) for q in range(4)]
        for a, b in zip(idx[:-1], idx[1:]):
            (t0, n0), (t1, n1) = marks[a], marks[b]
            if n1 > n0:
                per_call.append((t1 - t0) / (n1 - n0))
    med = float(np.median(per_call)) if per_call else wall / max(calls["n"], 1)
    return {"value": round(med * fp64_sigma_calls, 3), "unit": "s/solve", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(), "host_cores_visible": avail,
            "sample": f"{len(marks)} Davidson iterations ({calls['n']} FP64 sigma calls, {wall:.1f} s wall) of the "
                      f"same SF-CPSd solve on the oracle (C engines + {O.blas_name()} BLAS, {cores} threads); "
                      f"median cost per sigma call over 3 windows after the first iteration "
                      f"({', '.join(f'{x:.3f}' for x in per_call)} s) times the {fp64_sigma_calls} sigma calls of "
                      f"the FP64 solve at tolerance {args.conv:g}"}


if __name__ == "__main__":
    main()
