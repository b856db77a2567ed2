"""Kernel probe for profiling runs, and the summaries made from them.

Run under rocprofv3 to measure one kernel of the elasticity solve in
isolation on its real operands, next to a plain copy kernel of known size:

    python -m delfem_amd.kernel_probe run [--n 96] [--which 0] [--precision 4] [--nrep 50]

which: 0 SpMV, 1 lower ILU sweep, 2 upper ILU sweep (the PCG instantiations);
precision: 0 FP64 .. 4 FP32 vectors / FP32 matrix / BF16 factors.  The copy
kernel (k_copy4) reads and writes --copy-mib MiB per launch.

    python -m delfem_amd.kernel_probe summarize FETCH_CSV WRITE_CSV --kernel NAME --copy-mib M --out FILE

reads the counter_collection CSVs of a `--pmc FETCH_SIZE` run and a
`--pmc WRITE_SIZE` run, scales each counter of the kernel by the factor that
turns the copy kernel's reading into its known byte count, and writes the HBM
bytes per launch as one JSON line into a text summary.
"""
import argparse
import csv
import ctypes as C
import json
import os
import sys
import time


def run(args):
    import torch
    from delfem_amd import linsys, meshgen, _lib
    linsys.init(0)
    t0 = time.time()
    prob = meshgen.solid3d(args.n, seed=12345, sigma=1.0)
    s = linsys.FEMSystem(prob)
    s.assemble()
    M = linsys.CPreconditioner_ILU(s.A, 0, ordering="multicolour")
    M.SetValue(s.A)
    torch.cuda.synchronize()
    print(f"[probe] elasticity {args.n}^3 ready in {time.time() - t0:.1f}s", flush=True)
    a, b = C.c_double(), C.c_double()
    _lib.call("dfm_bench_kernel", s.A.h, M.h, args.which, args.precision, args.nrep, C.byref(a), C.byref(b))
    print(f"[probe] kernel which={args.which} precision={args.precision}: {a.value:.5f} ms/launch, "
          f"{b.value / 1e6:.2f} MB algorithmic, {b.value / a.value / 1e6:.1f} GB/s", flush=True)
    nbytes = args.copy_mib << 20
    _lib.call("dfm_bench_copy", C.c_size_t(nbytes), args.nrep, C.byref(a))
    print(f"[probe] copy {args.copy_mib} MiB: {a.value:.5f} ms/launch, {2 * nbytes / a.value / 1e6:.1f} GB/s "
          f"(read + write)", flush=True)


def _counter_rows(path):
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            yield row


def _mean_counter(path, kernel_substr, counter):
    vals = {}
    for row in _counter_rows(path):
        name = row.get("Kernel_Name", "")
        if kernel_substr not in name or row.get("Counter_Name") != counter:
            continue
        did = row.get("Dispatch_Id")
        vals[did] = vals.get(did, 0.0) + float(row["Counter_Value"])
    if not vals:
        raise SystemExit(f"{path}: no {counter} rows for a kernel matching {kernel_substr!r}")
    return sum(vals.values()) / len(vals), len(vals)


def summarize(args):
    nbytes = float(args.copy_mib << 20)
    fk, nf = _mean_counter(args.fetch_csv, args.kernel, "FETCH_SIZE")
    fc, _ = _mean_counter(args.fetch_csv, "k_copy4", "FETCH_SIZE")
    wk, nw = _mean_counter(args.write_csv, args.kernel, "WRITE_SIZE")
    wc, _ = _mean_counter(args.write_csv, "k_copy4", "WRITE_SIZE")
    f_fetch = nbytes / fc            # bytes per counter unit, from the copy kernel
    f_write = nbytes / wc
    out = dict(kernel=args.kernel, bytes_per_launch=round(fk * f_fetch + wk * f_write),
               fetch_bytes=round(fk * f_fetch), write_bytes=round(wk * f_write),
               fetch_counter=fk, write_counter=wk, copy_fetch_counter=fc, copy_write_counter=wc,
               fetch_scale=f_fetch, write_scale=f_write, dispatches=[nf, nw], copy_bytes=nbytes)
    with open(args.out, "w") as f:
        f.write(f"# HBM traffic per launch of {args.kernel}\n")
        f.write(f"# {args.note}\n" if args.note else "")
        f.write("# FETCH_SIZE and WRITE_SIZE (rocprofv3 --pmc, one pass each) averaged over the kernel's\n")
        f.write("# dispatches, each scaled by copy_bytes / (the same counter for k_copy4 moving copy_bytes).\n")
        f.write(json.dumps(out) + "\n")
    print(json.dumps(out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--n", type=int, default=96)
    r.add_argument("--which", type=int, default=0)
    r.add_argument("--precision", type=int, default=4)
    r.add_argument("--nrep", type=int, default=50)
    r.add_argument("--copy-mib", type=int, default=1024)
    sm = sub.add_parser("summarize")
    sm.add_argument("fetch_csv")
    sm.add_argument("write_csv")
    sm.add_argument("--kernel", required=True)
    sm.add_argument("--copy-mib", type=int, default=1024)
    sm.add_argument("--out", required=True)
    sm.add_argument("--note", default="")
    args = ap.parse_args(argv)
    if args.cmd == "run":
        run(args)
    else:
        summarize(args)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
