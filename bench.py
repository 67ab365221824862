"""Benchmark: UW shallow-cumulus columns/s on one MI355X.

Headline workload: the FP64 forward of the column physics on a 32 768-column chunk
(a 256x128 grid) of 35 levels and 3 constituents, inputs resident in HBM.  One
step = one call of the forward path over the whole chunk.  Timing: W untimed
warm-up steps, then K steps between two synchronisations, with HIP events
recorded on the stream the kernels run on.

The JSON line also carries:
  roofline      HBM view of the forward kernel: algorithmic bytes per launch (the
                FP64 inputs the physics reads plus the full output set, per column
                8*[9*(pver+1) + 24*pver + 2*ncnst*pver + 9] bytes) over the
                kernel's average duration, against 8 TB/s;
  cpu_baseline  the CPU oracle (FP64 C restatement of the reference with the host
                libm, OpenMP over columns) on a bounded sample of the same chunk;
  error_vs_oracle  on that sample: whether the GPU outputs are bit-identical to the
                portable-math oracle, and the per-column relative error against the
                host-libm oracle (the tendency error the metric is stated at).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hpc-and-climate-model_amd"))

METRIC = "UWshcu columns/s (FP64 vs mixed; fwd & fwd+adjoint) at stated tendency error"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncol", type=int, default=32768)
    ap.add_argument("--pver", type=int, default=35)
    ap.add_argument("--ncnst", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--inversion-frac", type=float, default=0.25)
    ap.add_argument("--cpu-sample", type=int, default=8192, help="columns timed on the CPU oracle")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the column physics shards by columns; see the design notes)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from uwshcu_amd import _capi, device
    from uwshcu_amd.generator import algorithmic_bytes_per_column, make_chunk

    torch.cuda.set_device(0)
    prec = _capi.FP64 if args.precision == "fp64" else _capi.MIXED
    ch = make_chunk(args.ncol, args.pver, args.ncnst, seed=args.seed, inversion_frac=args.inversion_frac)
    ctx = device.Context(0)
    inp = device.chunk_to_device(ch, "cuda:0")
    out = ctx.alloc_outputs(args.ncol, args.pver, args.ncnst, diag=True)
    cush = torch.empty_like(inp["cush"])
    stream = torch.cuda.current_stream()

    def step():
        cush.copy_(inp["cush"])
        ctx.forward_into(inp, out, cush, dt=ch.dt, precision=prec, stream=stream)

    for _ in range(args.warmup):
        step()
    ctx.check(stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for s in range(args.steps):
        step()
        ev[s + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rc, bad = ctx.status(stream)
    if rc:
        raise _capi.UwshcuError(rc, bad)
    per_step_ms = [ev[s].elapsed_time(ev[s + 1]) for s in range(args.steps)]
    ms = float(np.mean(per_step_ms))
    cols_per_s = args.ncol / (ms * 1e-3)
    diag = out["diag"].cpu().numpy()
    conv_frac = float((diag[7] > 0).mean())
    bpc = algorithmic_bytes_per_column(args.pver, args.ncnst)
    achieved = bpc * args.ncol / (ms * 1e-3) / 1e9
    res = {
        "metric": METRIC,
        "value": cols_per_s,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == _capi.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.precision} forward of compute_uwshcu_inv, {args.ncol} columns x {args.pver} levels x "
                        f"{args.ncnst} constituents, dt={ch.dt:g} s",
            "ncol": args.ncol, "pver": args.pver, "ncnst": args.ncnst,
            "convecting_fraction": conv_frac, "inversion_frac": args.inversion_frac, "seed": args.seed,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes_per_column": bpc, "note": "VALU-bound kernel; HBM view of algorithmic I/O",
        },
        "step_ms_min": float(np.min(per_step_ms)),
        "step_ms_median": float(np.median(per_step_ms)),
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        res.update(cpu_leg(args, ch, out, cush))
    print(json.dumps(res), flush=True)


def cpu_leg(args, ch, out, cush):
    """CPU baseline (the oracle on a bounded sample) and the GPU-vs-oracle error on it."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    import parity
    from uwshcu_amd.generator import Chunk

    n = min(args.cpu_sample, ch.ncol)
    sub = Chunk(n, ch.pver, ch.ncnst, ch.dt, {k: np.ascontiguousarray(v[..., :n]) for k, v in ch.fields.items()})
    nthr = args.cpu_threads
    small = Chunk(min(256, n), ch.pver, ch.ncnst, ch.dt,
                  {k: np.ascontiguousarray(v[..., :min(256, n)]) for k, v in ch.fields.items()})
    oracle.run(small, nthreads=nthr)                      # warm the library and the thread pool
    t0 = time.perf_counter()
    ref, rdiag, rc, _ = oracle.run(sub, nthreads=nthr)    # host libm: the faithful reading of the reference
    dt_cpu = time.perf_counter() - t0
    pm, pmdiag, _, _ = oracle.run(sub, nthreads=nthr, portable=True)
    got = {k: v.cpu().numpy()[..., :n] for k, v in out.items() if k != "diag"}
    got["cush"] = cush.cpu().numpy()[:n]
    diag = out["diag"].cpu().numpy()[:, :n].T
    errs = parity.compare(got, ref)
    worst_col = np.zeros(n)
    for k in ref:
        worst_col = np.maximum(worst_col, parity.column_rel_err(got[k], ref[k]))
    identical = all(np.array_equal(got[k], pm[k]) for k in pm) and np.array_equal(diag, pmdiag)
    return {
        "cpu_baseline": {"value": n / dt_cpu, "unit": "columns/s", "cores": nthr, "kind": "port",
                         "sample": f"first {n} columns of the same chunk, FP64 C oracle (host libm), "
                                   f"OpenMP over columns, {nthr} threads"},
        "error_vs_oracle": {
            "bit_identical_to_portable_math_oracle": bool(identical),
            "host_libm_oracle_max_col_rel_err": float(worst_col.max()),
            "host_libm_oracle_cols_above_1e-10": int((worst_col > 1e-10).sum()),
            "host_libm_oracle_worst_field": max(errs, key=errs.get),
            "integer_diag_mismatch_columns": int((diag != rdiag).any(axis=1).sum()),
            "sample_columns": n,
        },
    }


if __name__ == "__main__":
    main()
