"""Benchmark: ms per ILU-preconditioned Krylov solve on one MI355X.

Default workload (the headline): 3-D linear elasticity on a 96^3 hex grid
split into Kuhn tetrahedra (912,673 nodes, 2,738,019 DOF, 3x3 blocks),
per-element Young's modulus lognormal(0, 1.0), fixed at z = 0, body force
(0, 0, -1); synthetic mesh and coefficients from a fixed seed.  One step =
ILU(0) numeric factorisation + preconditioned CG to 1e-10 relative residual,
inputs resident in HBM.  `value` is the mixed-precision path (FP32 vectors
and matrix, BF16 ILU factors, FP64 dot accumulators, FP64 true residual and
reliable updates, convergence on the FP64 true residual); the plain FP64
path (reference semantics) is timed beside it.

Other workloads: --config cfg1 (2-D Poisson 128^2, ILU(0)-PCG to 1e-8),
cfg2 (elasticity 48^3, FP64 ILU(0)-PCG to 1e-10), cfg4 (Oseen P1-P1 PSPG
lid-driven cavity 256^2, ILU(1)-BiCGSTAB to 1e-9), cfg5 (512 independent
Helmholtz systems, batched ILU(0)-PCG to 1e-8 in one persistent launch),
zhelm (complex twin: complex Helmholtz with Sommerfeld lines 1024^2,
ILU(0)-PCOCG to 1e-8), mseg (elasticity 48^3 split into two segments,
ILU(0) across segments, PCG to 1e-10), mseg_fluid (Oseen 256^2 split into
velocity and pressure segments, ILU(1) with fill between segments,
BiCGSTAB to 1e-9).

Prints one JSON line:  python bench.py [--steps K] [--warmup W] [--config cfg3]

Profiling aids (no JSON line):
  python bench.py --kernel-loop WHICH --kernel-precision P [--nrep N] [--copy-mib M]
      the default workload's system, then one of the iteration's kernels
      (WHICH 0 SpMV, 1 lower sweep, 2 upper sweep) launched N times on its
      real operands, and a copy kernel of M MiB: the program to run under
      rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE.
  python bench.py --summarize-pmc FETCH_CSV WRITE_CSV --kernel NAME --out FILE
      HBM bytes per launch of NAME from the two counter passes, each counter
      scaled by what it reads for the copy kernel of known size.
  python bench.py --summarize-trace KERNEL_TRACE_CSV --kernel NAME --out FILE
      per-kernel statistics of a rocprofv3 --kernel-trace run, with NAME's
      average over its launches that did work (at least half the median).
"""
import argparse
import csv
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "ms per ILU-PCG solve to 1e-10 rel-res (mixed vs FP64); HBM GB/s % of peak"

CONFIGS = {
    "cfg1": dict(kind="poisson2d", n=128, sigma=0.5, tol=1e-8, lev=0, ordering="multicolour", solver="pcg",
                 mixed="mixed_f32_bf16", headline="fp64",
                 workload="2-D Poisson P1, 128x128 cells (16,641 nodes), lognormal(0,0.5) diffusivity, "
                          "ILU(0)-PCG to 1e-8"),
    "cfg2": dict(kind="solid3d", n=48, sigma=0.0, tol=1e-10, lev=0, ordering="multicolour", solver="pcg",
                 mixed="mixed_f32_bf16", headline="fp64",
                 workload="3-D linear elasticity P1 tets, 48^3 grid (117,649 nodes, 352,947 DOF), E=1 nu=0.3, "
                          "ILU(0)-PCG to 1e-10"),
    "cfg3": dict(kind="solid3d", n=96, sigma=1.0, tol=1e-10, lev=0, ordering="multicolour", solver="pcg",
                 mixed="mixed_f32_bf16",
                 workload="3-D linear elasticity P1 tets, 96^3 grid (912,673 nodes, 2,738,019 DOF), "
                          "E lognormal(0,1.0), nu=0.3, ILU(0)-PCG to 1e-10"),
    "cfg5": dict(kind="batch", count=512, tol=1e-8, maxit=20000,
                 workload="512 independent 2-D P1 positive Helmholtz systems (rectangles, 500-50,000 nodes "
                          "log-uniform, aspect 1-4, lognormal(0,0.5) diffusivity, c uniform [0,10]), "
                          "multicolour ILU(0)-PCG to 1e-8 each, one persistent launch"),
    "zhelm": dict(kind="zhelmholtz", n=1024, wave_length=1.0, tol=1e-8, lev=0, ordering="multicolour",
                  maxit=20000,
                  workload="complex twin: Helmholtz P1 triangles + Sommerfeld radiation lines on the unit square, "
                           "1024x1024 cells (1,050,625 nodes), wave_length 1.0, incoming wave u = 1 on x = 0, "
                           "ILU(0)-PCOCG to 1e-8"),
    "mseg_fluid": dict(kind="mseg", problem="oseen2d", n=256, groups=[[0, 1], [2]], lev=1, solver="bicg", tol=1e-5,
                       maxit=100, ordering="tile_multicolour",
                       workload="multi-segment system: Oseen P1-P1 PSPG 256x256 cells (66,049 nodes) split into a "
                                "velocity segment and a pressure segment, ILU(1) across segments (fill between "
                                "segments; nodes numbered by coloured 4x4 tiles), BiCGSTAB to 1e-5 or 100 "
                                "iterations (the fluid solver's settings)"),
    "mseg": dict(kind="mseg", n=48, sigma=0.0, groups=[[0, 1], [2]], tol=1e-10, maxit=5000,
                 workload="multi-segment system: 3-D linear elasticity 48^3 (117,649 nodes) split into an (ux, uy) "
                          "segment and a uz segment, ILU(0) across segments (nodes multicolour-numbered), "
                          "PCG to 1e-10"),
    # FAC32 BiCGSTAB keeps FP64 vectors: its recurrence needs no intermediate
    # reliable updates (each would restart BiCGSTAB's recurrences), so the FP64
    # true residual is checked only when the recurrence has converged
    "cfg4": dict(kind="oseen2d", n=256, tol=1e-9, lev=1, ordering="block_multicolour", block_nodes=16,
                 solver="bicg", mixed="fac32", others=("fac16",), mixed_opts={"fac32": {"reliable_delta": 1e-150}},
                 workload="Oseen P1-P1 PSPG lid-driven cavity, 256x256 cells (66,049 nodes, 198,147 DOF), "
                          "nu=0.01, Fourier advection, ILU(1)-BiCGSTAB to 1e-9"),
}
PREC_NAMES = {"fp64": 0, "mixed": 1, "mixed_bf16": 2, "mixed_f32": 3, "mixed_f32_bf16": 4, "fac32": 5, "fac16": 6}
OTHERS = ("mixed", "mixed_bf16", "mixed_f32", "mixed_f32_bf16", "fac32", "fac16")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spmv_f32_elasticity96_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(cfg, seed):
    from delfem_amd import meshgen
    if cfg["kind"] == "poisson2d":
        return meshgen.poisson2d(cfg["n"], seed=seed, sigma=cfg["sigma"])
    if cfg["kind"] == "oseen2d":
        return meshgen.oseen2d(cfg["n"], seed=seed)
    return meshgen.solid3d(cfg["n"], seed=seed, sigma=cfg["sigma"])


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("{"):
                    d = json.loads(line)
                    if d.get("kernel") == kernel:
                        return d
    except OSError:
        pass
    return None


def kernel_loop(args):
    """One kernel of the default workload's iteration, isolated, for counter runs."""
    import torch
    from delfem_amd import linsys, _lib
    linsys.init(0)
    cfg = CONFIGS["cfg3"]
    t0 = time.time()
    s = linsys.FEMSystem(make_problem(cfg, args.seed))
    s.assemble()
    M = linsys.CPreconditioner_ILU(s.A, cfg["lev"], ordering=cfg["ordering"])
    M.SetValue(s.A)
    torch.cuda.synchronize()
    log(f"[loop] {cfg['workload']}: ready in {time.time() - t0:.1f}s")
    a, b = C.c_double(), C.c_double()
    _lib.call("dfm_bench_kernel", s.A.h, M.h, args.kernel_loop, args.kernel_precision, args.nrep, C.byref(a),
              C.byref(b))
    log(f"[loop] kernel which={args.kernel_loop} precision={args.kernel_precision}: {a.value:.5f} ms/launch, "
        f"{b.value / 1e6:.2f} MB algorithmic, {b.value / a.value / 1e6:.1f} GB/s")
    nbytes = args.copy_mib << 20
    _lib.call("dfm_bench_copy", C.c_size_t(nbytes), args.nrep, C.byref(a))
    log(f"[loop] copy {args.copy_mib} MiB: {a.value:.5f} ms/launch, {2 * nbytes / a.value / 1e6:.1f} GB/s "
        f"(read + write)")


def _mean_counter(path, kernel_substr, counter):
    vals = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if kernel_substr not in row.get("Kernel_Name", "") or row.get("Counter_Name") != counter:
                continue
            did = row.get("Dispatch_Id")
            vals[did] = vals.get(did, 0.0) + float(row["Counter_Value"])
    if not vals:
        raise SystemExit(f"{path}: no {counter} rows for a kernel matching {kernel_substr!r}")
    return sum(vals.values()) / len(vals), len(vals)


def _out_file(path):
    """The summary file, or standard output when no --out is given."""
    import contextlib
    return open(path, "w") if path else contextlib.nullcontext(sys.stdout)


def summarize_pmc(args):
    """HBM bytes per launch from a FETCH_SIZE pass and a WRITE_SIZE pass."""
    fetch_csv, write_csv = args.summarize_pmc
    nbytes = float(args.copy_mib << 20)
    fk, nf = _mean_counter(fetch_csv, args.kernel, "FETCH_SIZE")
    fc, _ = _mean_counter(fetch_csv, "k_copy4", "FETCH_SIZE")
    wk, nw = _mean_counter(write_csv, args.kernel, "WRITE_SIZE")
    wc, _ = _mean_counter(write_csv, "k_copy4", "WRITE_SIZE")
    f_fetch, f_write = nbytes / fc, nbytes / wc  # bytes per counter unit, from the copy kernel
    out = dict(kernel=args.kernel, bytes_per_launch=round(fk * f_fetch + wk * f_write),
               fetch_bytes=round(fk * f_fetch), write_bytes=round(wk * f_write),
               fetch_counter=fk, write_counter=wk, copy_fetch_counter=fc, copy_write_counter=wc,
               fetch_scale=f_fetch, write_scale=f_write, dispatches=[nf, nw], copy_bytes=nbytes)
    with _out_file(args.out) as f:
        f.write(f"# HBM traffic per launch of {args.kernel}\n")
        f.write(f"# {args.note}\n" if args.note else "")
        f.write("# FETCH_SIZE and WRITE_SIZE (rocprofv3 --pmc, one pass each) averaged over the kernel's\n")
        f.write("# dispatches, each scaled by copy_bytes / (the same counter for k_copy4 moving copy_bytes).\n")
        f.write(json.dumps(out) + "\n")
    print(json.dumps(out))


def summarize_trace(args):
    """Per-kernel statistics of a kernel trace; the named kernel also over its working launches."""
    import re
    per = {}
    rows = []
    with open(args.summarize_trace, newline="") as f:
        for row in csv.DictReader(f):
            name = row["Kernel_Name"].replace("(anonymous namespace)::", "")
            rows.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"]), re.sub(r"\(.*", "", name).strip()))
    rows.sort()
    for t0, t1, name in rows:
        per.setdefault(name, []).append((t1 - t0) / 1e3)
    total = sum(sum(v) for v in per.values())
    lines = []
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{100 * sum(v) / total:6.2f}% {len(v):8d} {sum(v) / len(v):10.2f} {min(v):10.2f} / "
                     f"{max(v):9.2f}  {name}")
    key = None
    for name in per:
        if args.kernel and name.endswith(args.kernel):
            key = name
    with _out_file(args.out) as f:
        f.write(f"# {args.note}\n" if args.note else "")
        f.write("# columns: share of GPU time, calls, average duration (us), min / max duration (us), kernel\n")
        for ln in lines[:40]:
            f.write(ln + "\n")
        def working(durs):
            half = 0.5 * sorted(durs)[len(durs) // 2]
            return [d for d in durs if d >= half]

        if key:
            v = per[key]
            w = working(v)
            f.write(f"# {key}: {len(v)} launches, average {sum(v) / len(v):.2f} us; the {len(w)} that did work "
                    f"(>= half the median) average {sum(w) / len(w):.2f} us\n")
            if args.tail and args.tail <= len(v):
                t = v[-args.tail:]
                tw = working(t)
                f.write(f"# the last {len(t)} launches (bench.py's in-solve timing solve, launched directly with "
                        f"start / stop events): average {sum(t) / len(t):.2f} us; the {len(tw)} that did work "
                        f"average {sum(tw) / len(tw):.2f} us\n")
    if args.out:
        print(open(args.out).read())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg3", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=40, help="Krylov iterations in the CPU sample")
    ap.add_argument("--skip-natural", action="store_true",
                    help="multi-segment configs: leave out the natural-numbering comparison solve")
    ap.add_argument("--kernel-loop", type=int, default=None, metavar="WHICH",
                    help="profiling aid: run one iteration kernel of the default workload in a loop")
    ap.add_argument("--kernel-precision", type=int, default=4)
    ap.add_argument("--nrep", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--summarize-trace", metavar="KERNEL_TRACE_CSV")
    ap.add_argument("--tail", type=int, default=0, help="--summarize-trace: also average the kernel's last N launches")
    ap.add_argument("--kernel", default="")
    ap.add_argument("--out", default="")
    ap.add_argument("--note", default="")
    args = ap.parse_args()
    if args.summarize_pmc:
        return summarize_pmc(args)
    if args.summarize_trace:
        return summarize_trace(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1", "n_gpus": args.gpus}),
              flush=True)
        sys.exit(2)

    if args.kernel_loop is not None:
        return kernel_loop(args)
    import torch
    from delfem_amd import linsys, _lib

    cfg = CONFIGS[args.config]
    linsys.init(0)
    if cfg["kind"] == "batch":
        return bench_batch(cfg, args, torch, linsys, _lib)
    if cfg["kind"] == "zhelmholtz":
        return bench_complex(cfg, args, torch, linsys, _lib)
    if cfg["kind"] == "mseg":
        return bench_mseg(cfg, args, torch, linsys, _lib)
    stream = torch.cuda.ExternalStream(int(_lib.lib().dfm_stream()))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    log(f"[bench] {args.config}: {cfg['workload']}")
    t0 = time.time()
    prob = make_problem(cfg, args.seed)
    s = linsys.FEMSystem(prob)
    torch.cuda.synchronize()
    ev[0].record(stream)
    s.assemble()
    ev[1].record(stream)
    ev[1].synchronize()
    asm_ms = ev[0].elapsed_time(ev[1])
    rhs = s.res.clone()
    log(f"[bench] mesh + pattern + assembly {time.time() - t0:.1f}s (device assembly {asm_ms:.2f} ms)")
    t0 = time.time()
    M = linsys.CPreconditioner_ILU(s.A, cfg["lev"], ordering=cfg["ordering"], block_nodes=cfg.get("block_nodes", 0))
    info = M.info()
    log(f"[bench] ILU({cfg['lev']}) symbolic + layouts {time.time() - t0:.1f}s: {info}")
    solve = linsys.Solve_PCG if cfg["solver"] == "pcg" else linsys.Solve_PBiCGSTAB

    def step(precision, tol=cfg["tol"], maxit=50000):
        ok_f = M.SetValue(s.A)
        r = rhs.clone()
        x = torch.zeros_like(r)
        extra_opts = cfg.get("mixed_opts", {}).get(precision, {})
        _, conv, it, extra = solve(tol, maxit, s.A, M, r, x, precision=precision, want_true_relres=1, **extra_opts)
        return extra["converged"] and ok_f, conv, it, extra, x

    def timed(precision, steps, warmup, maxit=50000):
        for _ in range(warmup):
            step(precision, maxit=maxit)
        torch.cuda.synchronize()
        ev[0].record(stream)
        outs = [step(precision, maxit=maxit) for _ in range(steps)]
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_time(ev[1]) / steps, outs[-1]

    mixed = cfg["mixed"]
    log(f"[bench] timed {mixed} steps")
    mixed_ms, (ok_m, conv_m, it_m, ex_m, x_m) = timed(mixed, args.steps, args.warmup)
    log(f"[bench] {mixed}: {mixed_ms:.2f} ms/solve, inner iters {ex_m['inner_iters']}, "
        f"outer {ex_m['outer_iters']}, true relres {ex_m['true_relres']:.3e}, ok {ok_m}")
    log("[bench] timed fp64 steps")
    fp64_ms, (ok_d, conv_d, it_d, ex_d, x_d) = timed("fp64", max(1, args.steps), min(1, args.warmup))
    iters_d = it_d + 1 if cfg["solver"] == "pcg" else it_d
    log(f"[bench] fp64: {fp64_ms:.2f} ms/solve, iterations {iters_d}, conv {conv_d:.3e}, "
        f"true relres {ex_d['true_relres']:.3e}, ok {ok_d}")
    # other mixed modes, one solve each (reported, not the headline)
    others = {}
    for prec in cfg.get("others", OTHERS):
        if prec == mixed:
            continue
        ms, (ok_o, _, it_o, ex_o, _) = timed(prec, 1, 0, maxit=4 * iters_d)
        others[prec] = dict(ms=round(ms, 3), inner_iters=ex_o["inner_iters"], outer_iters=ex_o["outer_iters"],
                            relres=ex_o["true_relres"], ok=bool(ok_o))
        log(f"[bench] {prec}: {ms:.2f} ms/solve, {ex_o['inner_iters']} iterations, ok {ok_o}")
    # FP64 solution converged far below the target: the error reference
    _, _, _, ex_ref, x_ref = step("fp64", tol=1e-13 if cfg["solver"] == "pcg" else 1e-12)
    nref = torch.linalg.norm(x_ref).item()
    err_m = torch.linalg.norm(x_m - x_ref).item() / nref
    err_d = torch.linalg.norm(x_d - x_ref).item() / nref

    # the iteration's kernels on their real operands, timed with HIP events
    # on the library's stream (dfm_bench_kernel)
    kern = {}
    for which, name in ((0, "spmv"), (1, "ilu_lower"), (2, "ilu_upper")):
        for prec in (mixed, "fp64"):
            a, b = C.c_double(), C.c_double()
            _lib.call("dfm_bench_kernel", s.A.h, M.h, which, PREC_NAMES[prec], 50, C.byref(a), C.byref(b))
            gbs = b.value / (a.value * 1e-3) / 1e9
            kern[f"{name}_{prec}"] = dict(ms=round(a.value, 5), bytes=b.value, gbs=round(gbs, 1),
                                          frac=round(gbs / HBM_PEAK_GBS, 4))
    for k, v in kern.items():
        log(f"[bench] kernel {k}: {v['ms']:.4f} ms, {v['bytes'] / 1e6:.1f} MB, {v['gbs']:.0f} GB/s "
            f"({100 * v['frac']:.1f}% of peak)")
    ceiling = read_ceiling(_lib)
    log(f"[bench] streaming read of 2 GiB (library read kernel): {ceiling:.0f} GB/s")
    # the configuration's own precision: FP64 where it asks for plain FP64
    # (cfg1, cfg2), the mixed mode otherwise
    head = cfg.get("headline", mixed)
    head_ms = fp64_ms if head == "fp64" else mixed_ms
    dom = kern[f"spmv_{head}"]
    # the same kernel inside a solve: one more solve of the headline mode whose
    # iteration SpMVs carry kernel start / stop events (launched directly, not
    # from graphs); the launches after convergence inside the last block of
    # iterations return at once and are left out
    _lib.call("dfm_set_kernel_timing", 1)
    try:
        step(head)
        t_all, t_work, n_all, n_work = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        _lib.call("dfm_get_kernel_timing", C.byref(t_all), C.byref(t_work), C.byref(n_all), C.byref(n_work))
    finally:
        _lib.call("dfm_set_kernel_timing", 0)
    in_solve = dict(avg_ms=round(t_work.value, 5), launches=n_work.value, all_avg_ms=round(t_all.value, 5),
                    all_launches=n_all.value)
    in_gbs = dom["bytes"] / (t_work.value * 1e-3) / 1e9 if t_work.value > 0 else 0.0
    log(f"[bench] in-solve SpMV ({head}): {t_work.value:.4f} ms over {n_work.value} working launches "
        f"({n_all.value} in all, {t_all.value:.4f} ms), {in_gbs:.0f} GB/s ({100 * in_gbs / HBM_PEAK_GBS:.1f}% of peak)")
    kname = {"mixed_f32": "k_spmv<3, float, float, true, true>", "mixed_f32_bf16": "k_spmv<3, float, float, true, true>",
             "mixed": "k_spmv<3, float, double, true, true>", "mixed_bf16": "k_spmv<3, float, double, true, true>"
             }.get(head, "k_spmv<3, double, double, true, true>")
    tr = read_traffic(kname) if args.config == "cfg3" else None

    cpu = None
    if not args.no_cpu:
        # the reference runs its own natural ordering where that converges;
        # otherwise the benchmark's ordering, given to the reference's ordered ILU
        iters_nat = natural_iterations(linsys, s, cfg, rhs, solve, torch) if cfg["ordering"] != "natural" else iters_d
        if iters_nat is None:
            cpu = cpu_baseline(prob, s, cfg, args, iters_d, M.pattern()["order"])
        else:
            cpu = cpu_baseline(prob, s, cfg, args, iters_nat, None)

    out = {
        "metric": METRIC,
        "value": round(head_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head in ("mixed_f32", "mixed_f32_bf16") else "f64",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "config": args.config, "seed": args.seed,
                   "preconditioner": f"ILU({cfg['lev']}), {cfg['ordering']} ordering",
                   "ncolour": info["ncolour"], "levels": [info["nlevel_lower"], info["nlevel_upper"]],
                   "tol": cfg["tol"], "precision": head, "solve_options": cfg.get("mixed_opts", {}).get(head, {})},
        "mixed_ms": round(mixed_ms, 3),
        "mixed_precision": mixed,
        "fp64_ms": round(fp64_ms, 3),
        "speedup_mixed_vs_fp64": round(fp64_ms / mixed_ms, 3),
        "iters_fp64": iters_d,
        "inner_iters_mixed": ex_m["inner_iters"],
        "outer_iters_mixed": ex_m["outer_iters"],
        "relres_mixed": ex_m["true_relres"],
        "relres_fp64": ex_d["true_relres"],
        "converged": bool(ok_m and ok_d),
        "err_mixed_vs_fp64_tight": err_m,
        "err_fp64_vs_fp64_tight": err_d,
        "other_precisions": others,
        "assembly_ms": round(asm_ms, 3),
        "kernels": kern,
        "roofline": {"bound": "hbm", "achieved": round(in_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(in_gbs / HBM_PEAK_GBS, 4), "traffic": tr["bytes_per_launch"] if tr else None,
                     "kernel": kname, "algorithmic_bytes": dom["bytes"], "avg_launch_ms": in_solve["avg_ms"],
                     "in_solve": in_solve, "isolated_loop": {"avg_ms": dom["ms"], "frac": dom["frac"]},
                     "measured_read_ceiling_gbs": round(ceiling, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_batch(cfg, args, torch, linsys, _lib):
    """Batched workload: every step = ILU(0) numeric + PCG for all systems in
    one persistent launch (graph replay), FP32 with FP64 reliable updates;
    the FP64 batch (reference semantics per system) beside it."""
    import numpy as np
    from delfem_amd import meshgen
    stream = torch.cuda.ExternalStream(int(_lib.lib().dfm_stream()))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    log(f"[bench] {args.config}: {cfg['workload']}")
    t0 = time.time()
    parts = meshgen.helmholtz_batch(cfg["count"], seed=args.seed)
    prob = meshgen.concat(parts)
    fem = linsys.FEMSystem(prob)
    fem.assemble()
    B = linsys.BatchSystem(fem, prob.sys_ptr)
    B.SetValue()
    rhs = fem.res.clone()
    x = torch.zeros_like(rhs)
    nn = np.diff(prob.sys_ptr.astype(np.int64))
    log(f"[bench] {len(parts)} systems, {prob.nnode} nodes (min {nn.min()}, max {nn.max()}), "
        f"setup {time.time() - t0:.1f}s, {B.info()}")
    tune = B.autotune(rhs, x, cfg["tol"], cfg["maxit"], "mixed_f32")
    shape = min(tune, key=tune.get)
    log("[bench] autotune (block, workgroups per CU, team rows, team size) -> ms: " +
        ", ".join(f"{k}: {v:.1f}" for k, v in sorted(tune.items())) + f"; chosen {shape}")

    def timed(prec, steps, warmup):
        for _ in range(warmup):
            B.SetValue()
            B.solve(rhs, x, cfg["tol"], cfg["maxit"], prec)
        torch.cuda.synchronize()
        ev[0].record(stream)
        for _ in range(steps):
            B.SetValue()
            ok, out = B.solve(rhs, x, cfg["tol"], cfg["maxit"], prec)
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_time(ev[1]) / steps, ok, out

    mixed_ms, ok_m, out_m = timed("mixed_f32", args.steps, args.warmup)
    x_m = x.clone()
    fp64_ms, ok_d, out_d = timed("fp64", max(1, args.steps), min(1, args.warmup))
    x_d = x.clone()
    B.solve(rhs, x, 1e-12, cfg["maxit"], "fp64")
    errs = []
    xm, xd, xt = x_m.cpu().numpy(), x_d.cpu().numpy(), x.cpu().numpy()
    for s in range(len(parts)):
        r0, r1 = int(prob.sys_ptr[s]), int(prob.sys_ptr[s + 1])
        nt = np.linalg.norm(xt[r0:r1])
        errs.append((np.linalg.norm(xm[r0:r1] - xt[r0:r1]) / nt, np.linalg.norm(xd[r0:r1] - xt[r0:r1]) / nt))
    errs = np.array(errs)
    sm, dm = out_m["solve_ms"], out_m["done_ms"]
    log(f"[bench] mixed: {mixed_ms:.2f} ms per batch, {len(parts) / mixed_ms * 1e3:.0f} solves/s, per-system solve "
        f"p50 {np.percentile(sm, 50):.2f} p99 {np.percentile(sm, 99):.2f} ms; fp64 {fp64_ms:.2f} ms")
    # algorithmic bytes of the mixed step: per iteration the FP32 matrix and
    # factor slots (column + value), diagonals, and ten vector passes
    w_nnz = np.diff(fem.A.pattern()[0].astype(np.int64))
    nnz_s = np.add.reduceat(w_nnz, prob.sys_ptr[:-1].astype(np.int64))
    per_it = 2 * nnz_s * 8 + nn * (4 + 4 + 10 * 4)
    setup = nnz_s * (8 + 8 + 4 + 4) + nn * (8 + 8 + 4 + 4)
    total_bytes = float(np.sum(per_it * out_m["iters"].astype(np.int64) + setup))
    gbs = total_bytes / (mixed_ms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline_batch(prob, fem, cfg, out_d, args)
    out = {
        "metric": METRIC,
        "value": round(mixed_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "config": args.config, "seed": args.seed, "systems": len(parts),
                   "nodes": int(prob.nnode), "tol": cfg["tol"], "launch": list(shape),
                   "autotune_ms": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in tune.items()}},
        "solves_per_s": round(len(parts) / mixed_ms * 1e3, 1),
        "latency_ms": {"solve_p50": round(float(np.percentile(sm, 50)), 3),
                       "solve_p99": round(float(np.percentile(sm, 99)), 3),
                       "solve_max": round(float(sm.max()), 3),
                       "done_p50": round(float(np.percentile(dm, 50)), 3),
                       "done_p99": round(float(np.percentile(dm, 99)), 3)},
        "fp64_ms": round(fp64_ms, 3),
        "fp64_solves_per_s": round(len(parts) / fp64_ms * 1e3, 1),
        "converged": bool(ok_m and ok_d),
        "iters_mixed": {"p50": float(np.percentile(out_m["iters"], 50)), "max": int(out_m["iters"].max())},
        "iters_fp64": {"p50": float(np.percentile(out_d["iters"], 50)), "max": int(out_d["iters"].max())},
        "relres_mixed_max": float(out_m["relres"].max()),
        "relres_fp64_max": float(out_d["relres"].max()),
        "err_mixed_vs_fp64_tight_max": float(errs[:, 0].max()),
        "err_fp64_vs_fp64_tight_max": float(errs[:, 1].max()),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_batch_solve<6, true> (whole step)", "algorithmic_bytes": total_bytes,
                     "avg_launch_ms": round(mixed_ms, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_complex(cfg, args, torch, linsys, _lib):
    """Complex twin: every step = complex ILU(0) numeric factorisation +
    Solve_PCOCG (FP64 complex, the reference's arithmetic), inputs resident in
    HBM.  The multicolour ordering keeps the sweeps at a few level sets; the
    reference's natural-order factor needs about 2.5x the iterations here
    (2,006 against 797 at 256x256) and one launch per level, so the CPU
    sample is scaled to the multicolour count, which favours the CPU."""
    import numpy as np
    from delfem_amd import meshgen
    stream = torch.cuda.ExternalStream(int(_lib.lib().dfm_stream()))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    log(f"[bench] {args.config}: {cfg['workload']}")
    t0 = time.time()
    prob = meshgen.zhelmholtz2d(cfg["n"], cfg["n"], 1.0, 1.0, wave_length=cfg["wave_length"])
    S = linsys.ZHelmholtzSystem(prob)
    torch.cuda.synchronize()
    ev[0].record(stream)
    S.assemble()
    ev[1].record(stream)
    ev[1].synchronize()
    asm_ms = ev[0].elapsed_time(ev[1])
    rhs = S.res.clone()
    log(f"[bench] mesh + pattern + assembly {time.time() - t0:.1f}s (device assembly {asm_ms:.2f} ms)")
    t0 = time.time()
    M = linsys.CZMatDiaFrac_BlkCrs(cfg["lev"], S.A, ordering=cfg["ordering"])
    info = M.info()
    log(f"[bench] ILU({cfg['lev']}) symbolic + layouts {time.time() - t0:.1f}s: {info}")

    def step(Mx, maxit=cfg["maxit"]):
        Mx.SetValue(S.A)
        r = rhs.clone()
        u = torch.zeros_like(r)
        ok, conv, it = linsys.Solve_PCOCG(cfg["tol"], maxit, S.A, Mx, r, u)
        return conv, it, u

    for _ in range(args.warmup):
        step(M)
    torch.cuda.synchronize()
    ev[0].record(stream)
    for _ in range(args.steps):
        conv, it, u = step(M)
    ev[1].record(stream)
    ev[1].synchronize()
    ms = ev[0].elapsed_time(ev[1]) / args.steps
    converged = conv < cfg["tol"]
    # true relative residual ||b - A u|| / ||b|| (u lacks the last step, as the reference's)
    y = rhs.clone()
    S.A.MatVec(-1.0, u, 1.0, y)
    relres = float(np.sqrt(linsys.znorm2(y) / linsys.znorm2(rhs)))
    log(f"[bench] {cfg['ordering']} ILU({cfg['lev']})-PCOCG: {ms:.2f} ms/solve, {it} iterations, "
        f"conv {conv:.3e}, true relres {relres:.3e}")
    # algorithmic bytes of one iteration: matrix and factor slots (16-byte value
    # + 4-byte column), diagonals, and the vector passes of the nine kernels
    nb = prob.nnode
    nnz = S.A.NCrs() + M.NCrs()
    per_it = 20 * nnz + 348 * nb
    gbs = per_it * it / (ms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline_complex(prob, S, cfg, args, it)
    out = {
        "metric": "ms per complex ILU(0)-PCOCG solve (FP64 complex); HBM GB/s % of peak",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "config": args.config, "tol": cfg["tol"],
                   "preconditioner": f"ILU({cfg['lev']}), {cfg['ordering']} ordering",
                   "ncolour": info["ncolour"], "levels": [info["nlevel_lower"], info["nlevel_upper"]]},
        "iterations": it,
        "converged": bool(converged),
        "conv_ratio": conv,
        "true_relres": relres,
        "ms_per_iteration": round(ms / max(it, 1), 4),
        "assembly_ms": round(asm_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole PCOCG iteration (SpMV, sweeps, vector kernels)",
                     "algorithmic_bytes_per_iteration": per_it},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_mseg(cfg, args, torch, linsys, _lib):
    """Multi-segment workload: every step = cross-segment ILU(k) SetValue +
    Solve_PCG or Solve_PBiCGSTAB on the segments (device-driven loop: blocks
    of iterations replayed from graphs, the stop flag polled between them)."""
    from delfem_amd import meshgen
    log(f"[bench] {args.config}: {cfg['workload']}")
    t0 = time.time()
    if cfg.get("problem") == "oseen2d":
        prob = meshgen.oseen2d(cfg["n"], seed=args.seed)
    else:
        prob = meshgen.solid3d(cfg["n"], seed=args.seed, sigma=cfg["sigma"])
    lev, solver = cfg.get("lev", 0), cfg.get("solver", "pcg")
    names = f"ILU({lev}) across segments + {'BiCGSTAB' if solver == 'bicg' else 'PCG'}"
    runs = {}
    dev_ordering = cfg.get("ordering", "multicolour")
    for ordering in (dev_ordering,) if args.skip_natural else (dev_ordering, "natural"):
        S = linsys.MultiSegmentSystem(prob, cfg["groups"], ordering=ordering)
        S.assemble()
        rhs = S.res.clone()
        M = S.preconditioner(lev)
        log(f"[bench] {ordering}: mesh + split assembly + symbolic {time.time() - t0:.1f}s")

        def step():
            M.SetValue()
            r = rhs.clone()
            x = torch.zeros_like(r)
            return M.solve(solver, cfg["tol"], cfg["maxit"], r, x)

        steps = args.steps if ordering == dev_ordering else 1
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ok, conv, it = step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) * 1e3 / steps
        # Solve_PCG reports the 0-based index of its last iteration, Solve_PBiCGSTAB a count from 1
        runs[ordering] = (ms, ok, conv, it + 1 if solver == "pcg" else it)
        log(f"[bench] {ordering} {names}: {ms:.2f} ms/solve, {runs[ordering][3]} iterations, conv {conv:.3e}, "
            f"ok {ok}")
        t0 = time.time()
    ms, ok, conv, iters = runs[dev_ordering]
    it = iters - 1
    cpu = None
    if not args.no_cpu:
        # the reference's natural numbering, scaled to its own iteration count
        cpu = cpu_baseline_mseg(prob, cfg, args, runs.get("natural", runs[dev_ordering])[3])
    out = {
        "metric": f"ms per multi-segment {names} solve to {cfg['tol']:g} (device-driven loop)",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "config": args.config, "tol": cfg["tol"],
                   "segments": [len(g) for g in cfg["groups"]],
                   "ordering": f"nodes renumbered ({dev_ordering}) for the value; natural beside it"},
        "natural_ms": round(runs["natural"][0], 3) if "natural" in runs else None,
        "natural_iterations": runs["natural"][3] if "natural" in runs else None,
        "iterations": it + 1,
        "converged": bool(ok),
        "conv_ratio": conv,
        "ms_per_iteration": round(ms / (it + 1), 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_mseg(prob, cfg, args, iters):
    """The reference's multi-segment CLinearSystem (compiled from its sources)
    on one core: CPreconditioner_ILU(k) across segments (symbolic + numeric)
    and a sample of Solve_PCG / Solve_PBiCGSTAB iterations, scaled to the
    device's count."""
    from oracle import oracle
    try:
        ref = oracle.Ref()
    except RuntimeError:
        return None
    orc = oracle.Orc()
    sp = oracle.split_problem(orc, prob, cfg["groups"])
    R = oracle.RefMS(ref, sp, prob.conn)
    rhs = R.vector(-1)
    lev, bicg = cfg.get("lev", 0), cfg.get("solver", "pcg") == "bicg"
    t1 = time.time()
    R.ilu(lev)
    t_ilu = time.time() - t1
    t1 = time.time()
    R.solve(bicg, rhs, cfg["tol"], args.cpu_iters)
    t_it = (time.time() - t1) / args.cpu_iters
    log(f"[bench] cpu reference: ILU {t_ilu:.2f}s, {1e3 * t_it:.1f} ms/iteration, scaled to {iters} iterations")
    return {
        "value": round(1e3 * (t_ilu + t_it * iters), 1),
        "unit": "ms per solve (ILU + iterations x sampled ms/iteration)",
        "cores": 1,
        "kind": "reference",
        "sample": (f"the same split system: CPreconditioner_ILU({lev}) across the segments (symbolic + numeric), "
                   f"then {args.cpu_iters} {'Solve_PBiCGSTAB' if bicg else 'Solve_PCG'} iterations on one core, "
                   f"scaled to {iters} iterations (the device's count with the reference's natural numbering)"),
        "ilu_s": round(t_ilu, 3),
        "ms_per_iteration": round(1e3 * t_it, 2),
        "iterations": iters,
    }


def cpu_baseline_complex(prob, S, cfg, args, iters):
    """The reference's complex classes (compiled from its sources) on one host
    core: assembly by Mearge, CZMatDiaFrac_BlkCrs(lev) (symbolic + numeric)
    and a sample of Solve_PCOCG iterations, scaled to the device's iteration
    count."""
    from oracle import oracle
    orc = oracle.Orc()
    try:
        ref = oracle.Ref()
    except RuntimeError:
        return None
    rowptr, col = S.A.pattern()
    et, _ = orc.zemat_helmholtz2d(prob.coord, prob.conn, prob.wave_length, prob.value)
    el, _ = orc.zemat_sommerfeld(prob.coord, prob.lconn, prob.wave_length, prob.value)
    rs = oracle.RefZSystem(ref, prob.nnode, rowptr, col)
    rs.merge(prob.conn, et)
    rs.merge(prob.lconn, el)
    rs.set_bc(prob.flags)
    t1 = time.time()
    rs.factor(cfg["lev"])
    t_ilu = time.time() - t1
    rhs = S.res.cpu().numpy()
    t1 = time.time()
    rs.pcocg(rhs, cfg["tol"], args.cpu_iters)
    t_it = (time.time() - t1) / args.cpu_iters
    log(f"[bench] cpu reference: ILU({cfg['lev']}) {t_ilu:.2f}s, {1e3 * t_it:.1f} ms/iteration, "
        f"scaled to {iters} iterations")
    return {
        "value": round(1e3 * (t_ilu + t_it * iters), 1),
        "unit": "ms per solve (ILU + iterations x sampled ms/iteration)",
        "cores": 1,
        "kind": "reference",
        "sample": (f"the same assembled system: CZMatDiaFrac_BlkCrs({cfg['lev']}) symbolic + numeric, then "
                   f"{args.cpu_iters} Solve_PCOCG iterations timed on one core, scaled to {iters} iterations "
                   f"(the device's multicolour iteration count; the reference's natural order needs more)"),
        "ilu_s": round(t_ilu, 3),
        "ms_per_iteration": round(1e3 * t_it, 2),
        "iterations": iters,
    }


def cpu_baseline_batch(prob, fem, cfg, out_d, args):
    """The reference (compiled from its sources) solving every system of the
    batch: per system its natural-order CPreconditioner_ILU(0) (symbolic part
    set up beforehand, as the device batch is) then SetValue + Solve_PCG to the
    same tolerance.  All systems on a thread pool over the host cores (largest
    first; the reference calls release the GIL), timed as a whole; and a
    one-core figure from every 4th system, scaled by rows x iterations."""
    import concurrent.futures as cf
    import numpy as np
    from oracle import oracle
    try:
        ref = oracle.Ref()
    except RuntimeError:
        return None
    rowptr, col = fem.A.pattern()
    val, dia = fem.A.values()
    rhs = fem.res.cpu().numpy()
    nn = np.diff(prob.sys_ptr.astype(np.int64))
    order = np.argsort(-nn, kind="stable")
    t0 = time.time()
    systems = {}
    for s in order:
        r0, r1 = int(prob.sys_ptr[s]), int(prob.sys_ptr[s + 1])
        q0, q1 = int(rowptr[r0]), int(rowptr[r1])
        rs = oracle.RefSystem(ref, r1 - r0, 1, (rowptr[r0:r1 + 1] - q0).astype(np.uint32),
                              (col[q0:q1] - r0).astype(np.uint32))
        rs.set_matrix(val[q0:q1], dia[r0:r1])
        rs.set_vector(-1, rhs[r0:r1])
        systems[int(s)] = (rs, oracle.RefILU(rs, 0))
    t_setup = time.time() - t0

    def run(s):
        rs, pr = systems[s]
        t1 = time.time()
        ok = pr.set_value()
        ok2, conv, it = pr.solve("pcg", cfg["tol"], cfg["maxit"])
        return time.time() - t1, ok and ok2

    sample = [int(s) for s in order[::4]]
    t_one = sum(run(s)[0] for s in sample)
    work = nn * (out_d["iters"].astype(np.int64) + 1)
    scale = float(work.sum()) / float(work[sample].sum())
    ncores = max(1, min(16, os.cpu_count() or 1))
    for s in order:
        systems[int(s)][0].set_vector(-1, rhs[int(prob.sys_ptr[s]):int(prob.sys_ptr[s + 1])])
    t1 = time.time()
    with cf.ThreadPoolExecutor(max_workers=ncores) as ex:
        res = list(ex.map(run, [int(s) for s in order]))
    t_all = time.time() - t1
    nok = sum(1 for _, ok in res if ok)
    log(f"[bench] cpu reference: setup {t_setup:.1f}s; {len(nn)} systems on {ncores} threads {t_all:.2f}s "
        f"({nok} converged); one core, estimated from {len(sample)} systems: {t_one * scale:.1f}s")
    return {
        "value": round(1e3 * t_all, 1),
        "unit": "ms per batch of all systems",
        "cores": ncores,
        "kind": "reference",
        "sample": (f"all {len(nn)} systems, each the reference's natural-order ILU(0) SetValue + Solve_PCG to "
                   f"{cfg['tol']}, on a pool of {ncores} host threads, largest first (symbolic ILU set up "
                   f"beforehand, {t_setup:.1f}s, not timed)"),
        "solves_per_s": round(len(nn) / t_all, 2),
        "one_core_ms_estimated": round(1e3 * t_one * scale, 1),
        "one_core_sample": (f"{len(sample)} systems (every 4th by size) timed on one core, {t_one:.1f}s, scaled by "
                            f"sum(rows x device FP64 iterations) over all systems / over the sample ({scale:.2f})"),
    }


def read_ceiling(lib, gib=2, reps=10):
    """Streaming-read rate of this device: the library's 16-byte-load read
    kernel over 2 GiB, timed with HIP events on the library stream.  Context
    for the roofline fraction (what a read-dominated kernel can reach here),
    not its peak."""
    ms = C.c_double(0.0)
    lib.call("dfm_bench_read", C.c_size_t(gib << 30), reps, C.byref(ms))
    return (gib << 30) / (ms.value * 1e-3) / 1e9


def natural_iterations(linsys, s, cfg, rhs, solve, torch):
    """Iterations of the reference's own (natural-order) preconditioner on the
    same system, FP64 on the device: the count the CPU sample is scaled to."""
    if cfg["kind"] == "oseen2d":
        return None   # natural-order ILU(1)-BiCGSTAB does not reach 1e-9 here; the CPU runs the same ordering
    Mn = linsys.CPreconditioner_ILU(s.A, cfg["lev"])
    Mn.SetValue(s.A)
    r = rhs.clone()
    x = torch.zeros_like(r)
    t0 = time.time()
    ok, conv, it, _ = solve(cfg["tol"], 50000, s.A, Mn, r, x)
    log(f"[bench] natural-order FP64: {it + 1} iterations ({time.time() - t0:.1f}s)")
    return it + 1 if cfg["solver"] == "pcg" else it


def cpu_baseline(prob, s, cfg, args, iters, order):
    """The DelFEM reference's solver (compiled from its sources) on the host,
    one core, sampled: ILU numeric factorisation + a fixed number of Krylov
    iterations, scaled to the iteration count of a full solve."""
    from oracle import oracle
    rowptr, col = s.A.pattern()
    val, dia = s.A.values()
    rhs = s.res.cpu().numpy()
    kind = "reference"
    try:
        ref = oracle.Ref()
    except RuntimeError:
        ref = None
        kind = "port"
    t0 = time.time()
    if ref is not None:
        rs = oracle.RefSystem(ref, prob.nnode, prob.b, rowptr, col)
        rs.set_matrix(val, dia)
        rs.set_vector(-1, rhs)
        pr = oracle.RefILU(rs, cfg["lev"], order)
        t1 = time.time()
        pr.set_value()
        t_ilu = time.time() - t1
        t1 = time.time()
        pr.solve(cfg["solver"], cfg["tol"], args.cpu_iters)
        t_it = (time.time() - t1) / args.cpu_iters
    else:
        orc = oracle.Orc()
        F = orc.iluk_symbolic(rowptr, col, cfg["lev"], order)
        t1 = time.time()
        fv, fd, _ = orc.ilu_values(prob.nnode, prob.b, rowptr, col, val, dia, F, order)
        t_ilu = time.time() - t1
        t1 = time.time()
        orc.solve(cfg["solver"], prob.nnode, prob.b, rowptr, col, val, dia, F, fv, fd, rhs, cfg["tol"],
                  args.cpu_iters, order)
        t_it = (time.time() - t1) / args.cpu_iters
    log(f"[bench] cpu {kind}: ILU numeric {t_ilu:.2f}s, {1e3 * t_it:.1f} ms/iteration, sampled "
        f"{time.time() - t0:.1f}s, scaled to {iters} iterations")
    what = "natural-order" if order is None else f"{cfg['ordering']}-ordered"
    return {
        "value": round(1e3 * (t_ilu + t_it * iters), 1),
        "unit": "ms per solve (ILU numeric + iterations x sampled ms/iteration)",
        "cores": 1,
        "kind": kind,
        "sample": (f"the same assembled system: ILU({cfg['lev']}) numeric ({what}) + {args.cpu_iters} "
                   f"{'PCG' if cfg['solver'] == 'pcg' else 'BiCGSTAB'} iterations timed on one core, scaled to "
                   f"{iters} iterations (the device FP64 iteration count of the {what} preconditioner)"),
        "ilu_numeric_s": round(t_ilu, 3),
        "ms_per_iteration": round(1e3 * t_it, 2),
        "iterations": iters,
    }


if __name__ == "__main__":
    main()
