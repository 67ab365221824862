"""Benchmark: UW shallow-cumulus columns/s on one MI355X.

Headline workload: the FP64 forward of the column physics on a 32 768-column chunk
(a 256x128 grid) of 35 levels and 3 constituents, inputs resident in HBM.  One
step = one call of the forward path over the whole chunk.  Timing: W untimed
warm-up steps, then K steps between two synchronisations, with HIP events
recorded on the stream the kernels run on.

Other modes (same JSON contract; metric and config name the mode):
  --precision mixed   the mixed-precision forward (FP32 plume arithmetic, FP64
                      saturation/CIN/closure/fix-ups, FP64 recompute of columns that
                      fail the conservation or branch check); reports its error against
                      the FP64 path on the same chunk and the recomputed fraction.
                      The 1 036 800-column workload: --ncol 1036800.
  --adjoint           forward + adjoint (gradients of the calibration loss
                      sum (sten/cp - tau_T)^2 + (qvten - tau_q)^2 with respect to t0,
                      qv0, u0, v0, tke, pblh, rpen, rle, rkm); one step = FP64 forward
                      plus the adjoint.  The 65 536-column workload: --ncol 65536.

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
    ap.add_argument("--scheme", choices=["uw", "zw"], default="uw",
                    help="uw: compute_uwshcu_inv; zw: the modified scheme compute_zwshcu_inv (FP64, default options)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--inversion-frac", type=float, default=0.25)
    ap.add_argument("--cpu-sample", type=int, default=8192, help="columns timed on the CPU oracle")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint", action="store_true", help="time forward + adjoint")
    ap.add_argument("--tile", type=int, default=1,
                    help="generate ncol/tile distinct columns and repeat them tile times on the GPU "
                         "(for the multi-million-column workloads; columns are independent)")
    ap.add_argument("--workspace-gib", type=float, default=None,
                    help="context workspace limit (default: the library's 24 GiB; the adjoint uses 160 GiB)")
    ap.add_argument("--chunks", type=int, default=1,
                    help="split the workload into this many physics chunks and time one uwshcu_forward_chunks_device "
                         "call over them (value), beside one uwshcu_forward_device call per chunk")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 256 MiB device buffer three times: the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    return ap.parse_args()


def copy_calibration():
    """Three device-to-device copies of 256 MiB (32 Mi doubles), nothing else."""
    import torch
    n = 32 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_copy": n * 8, "copies": 3}), flush=True)


def timed(step, steps, warmup, stream):
    """Run warmup + steps calls of step(); per-step milliseconds from HIP events on stream."""
    import numpy as np
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for s in range(steps):
        step()
        ev[s + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return np.array([ev[s].elapsed_time(ev[s + 1]) for s in range(steps)]), wall


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(the column physics shards by columns; see the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import torch

    from uwshcu_amd import _capi, device
    from uwshcu_amd.generator import add_zw_inputs, algorithmic_bytes_per_column, make_chunk

    torch.cuda.set_device(0)
    prec = _capi.FP64 if args.precision == "fp64" else _capi.MIXED
    zw = args.scheme == "zw"
    if zw and (prec != _capi.FP64 or args.adjoint or args.chunks > 1 or args.tile > 1):
        sys.exit("bench.py: --scheme zw times the FP64 forward alone")
    if args.ncol % args.tile:
        sys.exit("bench.py: --ncol must be a multiple of --tile")
    ptop = 100.0 if args.pver > 60 else 200.0
    ch = make_chunk(args.ncol // args.tile, args.pver, args.ncnst, seed=args.seed,
                    inversion_frac=args.inversion_frac, ptop=ptop)
    if zw:
        add_zw_inputs(ch, args.seed)
        zw_par, zw_opts = _capi.zw_default_params()
    ws_gib = args.workspace_gib or (160.0 if args.adjoint else None)
    ctx = device.Context(0, workspace_limit=int(ws_gib * (1 << 30)) if ws_gib else None)
    inp = device.chunk_to_device(ch, "cuda:0")
    if args.tile > 1:
        inp = {k: v.repeat(*([1] * (v.dim() - 1)), args.tile).contiguous() for k, v in inp.items()}
    out = ctx.alloc_outputs(args.ncol, args.pver, args.ncnst, diag=True)
    cush = torch.empty_like(inp["cush"])
    stream = torch.cuda.current_stream()

    def fwd(precision):
        cush.copy_(inp["cush"])
        if zw:
            ctx.zw_forward_into(inp, out, cush, zw_opts, dt=ch.dt, params=zw_par, stream=stream)
        else:
            ctx.forward_into(inp, out, cush, dt=ch.dt, precision=precision, stream=stream)

    extra = {}
    if args.adjoint:
        # the calibration loss' cotangents; the adjoint rewrites out/cush like the forward
        rng = np.random.default_rng(args.seed + 1)
        dev = torch.device("cuda:0")
        tau_t = torch.from_numpy(rng.normal(0, 1e-5, (args.pver, args.ncol))).to(dev)
        tau_q = torch.from_numpy(rng.normal(0, 1e-8, (args.pver, args.ncol))).to(dev)
        grads = {k: torch.empty((args.pver, args.ncol), dtype=torch.float64, device=dev)
                 for k in ("t0", "qv0", "u0", "v0")}
        grads["tke"] = torch.empty((args.pver + 1, args.ncol), dtype=torch.float64, device=dev)
        grads["pblh"] = torch.empty(args.ncol, dtype=torch.float64, device=dev)
        grads["par_sum"] = torch.empty(3, dtype=torch.float64, device=dev)
        cpair = 1.00464e3

        def step():
            fwd(_capi.FP64)
            cot = {"sten": 2.0 * (out["sten"] / cpair - tau_t) / cpair, "qvten": 2.0 * (out["qvten"] - tau_q)}
            # the adjoint rewrites the forward's outputs (with the same values)
            cush.copy_(inp["cush"])
            ctx.adjoint_into(inp, out, cush, cot, grads, dt=ch.dt, stream=stream)
        per_step_ms, wall = timed(lambda: fwd(_capi.FP64), args.steps, args.warmup, stream)
        fwd_ms = float(np.mean(per_step_ms))
        per_step_ms, wall = timed(step, args.steps, args.warmup, stream)
        ctx.check(stream)
        extra["adjoint"] = {"fp64_forward_ms": fwd_ms, "forward_plus_adjoint_ms": float(np.mean(per_step_ms)),
                            "ratio_to_fp64_forward": fwd_ms / float(np.mean(per_step_ms)),
                            "param_gradient": grads["par_sum"].cpu().tolist()}
    elif args.chunks > 1:
        # the chunk arrays of a timestep: separate allocations, as CAM holds them
        bounds = np.linspace(0, args.ncol, args.chunks + 1).astype(int)
        calls = []
        for c in range(args.chunks):
            a, b = int(bounds[c]), int(bounds[c + 1])
            sub = {k: v[..., a:b].contiguous() for k, v in inp.items()}
            calls.append((sub, ctx.alloc_outputs(b - a, args.pver, args.ncnst), sub["cush"].clone(), None))

        table = ctx.chunk_table(calls)
        cush_all = torch.cat([sub["cush"] for sub, _, _, _ in calls])

        def batched():
            # cush is in/out: restore every chunk's input value with one multi-tensor copy
            torch._foreach_copy_([cu for _, _, cu, _ in calls], list(cush_all.split([cu.numel() for _, _, cu, _ in calls])))
            ctx.forward_chunks(table, dt=ch.dt, precision=prec, stream=stream)

        def one_by_one():
            for sub, o, cu, _ in calls:
                cu.copy_(sub["cush"])
                ctx.forward_into(sub, o, cu, dt=ch.dt, precision=prec, stream=stream)
        loop_ms, _ = timed(one_by_one, args.steps, args.warmup, stream)
        per_step_ms, wall = timed(batched, args.steps, args.warmup, stream)
        ctx.check(stream)
        fwd(prec)                 # the diagnostics below come from the whole chunk
        extra["chunks"] = {"count": args.chunks, "columns_per_chunk": args.ncol // args.chunks,
                           "one_call_per_chunk_ms": float(np.mean(loop_ms)),
                           "batched_ms": float(np.mean(per_step_ms)),
                           "speedup": float(np.mean(loop_ms)) / float(np.mean(per_step_ms))}
    else:
        if zw:
            # the UW scheme on the same chunk with the same rpen and rkm, for comparison
            def uw_fwd():
                cush.copy_(inp["cush"])
                ctx.forward_into(inp, out, cush, dt=ch.dt, params=zw_par, stream=stream)
            uw_ms, _ = timed(uw_fwd, args.steps, args.warmup, stream)
            extra["uw_forward_ms_same_chunk_and_params"] = float(np.mean(uw_ms))
        per_step_ms, wall = timed(lambda: fwd(prec), args.steps, args.warmup, stream)
    rc, bad = ctx.status(stream)
    if rc:
        raise _capi.UwshcuError(rc, bad)
    ms = float(np.mean(per_step_ms))
    cols_per_s = args.ncol / (ms * 1e-3)
    diag = out["diag"].cpu().numpy()
    conv_frac = float((diag[7] > 0).mean())
    if prec == _capi.MIXED and not args.adjoint:
        nfix = ctx.mixed_recomputed(stream)
        mix = {k: v.cpu().numpy() for k, v in out.items() if k != "diag"}
        mix["cush"] = cush.cpu().numpy()
        f64_ms, _ = timed(lambda: fwd(_capi.FP64), args.steps, 1, stream)
        extra["fp64_forward_ms_same_chunk"] = float(np.mean(f64_ms))
        extra["speedup_vs_fp64"] = float(np.mean(f64_ms)) / ms
        f64 = {k: v.cpu().numpy() for k, v in out.items() if k != "diag"}
        f64["cush"] = cush.cpu().numpy()
        raw_ms, _ = timed(lambda: fwd(_capi.MIXED_UNCHECKED), args.steps, 1, stream)
        raw = {k: v.cpu().numpy() for k, v in out.items() if k != "diag"}
        raw["cush"] = cush.cpu().numpy()
        extra["fp32_pass_alone_ms"] = float(np.mean(raw_ms))
        extra["fp32_pass_alone"] = mixed_error(raw, f64, 0, args.ncol)
        fwd(_capi.MIXED)          # leave the mixed result in place
        extra["mixed"] = mixed_error(mix, f64, nfix, args.ncol)
    bpc = algorithmic_bytes_per_column(args.pver, args.ncnst)
    achieved = bpc * args.ncol / (ms * 1e-3) / 1e9
    mode = "fp64 forward" if prec == _capi.FP64 else "mixed forward"
    if args.adjoint:
        mode = "fp64 forward + adjoint"
    if args.chunks > 1:
        mode += f" over {args.chunks} chunks in one call"
    scheme = "compute_zwshcu_inv (default options, rpen 15, rkm 10)" if zw else "compute_uwshcu_inv"
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
        "dtype": "f64" if (prec == _capi.FP64 or args.adjoint) else "f32+f64",
        "data": "synthetic" if args.tile == 1 else f"synthetic, {args.ncol // args.tile} distinct columns tiled {args.tile}x",
        "config": {
            "workload": f"{mode} of {scheme}, {args.ncol} columns x {args.pver} levels x "
                        f"{args.ncnst} constituents, dt={ch.dt:g} s",
            "ncol": args.ncol, "pver": args.pver, "ncnst": args.ncnst,
            "convecting_fraction": conv_frac, "inversion_frac": args.inversion_frac, "seed": args.seed,
            "workspace_gib": ws_gib if ws_gib else 24.0,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(mode, args),
            "algorithmic_bytes_per_launch": bpc * args.ncol, "algorithmic_bytes_per_column": bpc,
            "timed": ("the forward pipeline (two level-parallel set-up kernels + the column kernel, which is 97 % "
                      "of it in rocprofv3), HIP events on its stream") if args.chunks == 1 else
                     "one batched call: table upload, gather, the forward pipeline and scatter, HIP events on its stream",
            "note": "latency-bound FP64 VALU kernel; HBM view of the algorithmic I/O",
        },
        "step_ms_min": float(np.min(per_step_ms)),
        "step_ms_median": float(np.median(per_step_ms)),
        "wall_s_timed": wall,
    }
    res.update(extra)
    if not args.no_cpu:
        res.update(cpu_leg(args, ch, out, cush))
    print(json.dumps(res), flush=True)


def measured_traffic(mode, args):
    """HBM bytes per launch of this workload from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement (profiles/traffic.json, copy-calibrated), or None."""
    key = f"{mode.split()[0]} forward {args.ncol}x{args.pver}x{args.ncnst}"
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            ent = json.load(f).get(key)
    except OSError:
        return None
    ok = ent and "adjoint" not in mode and "chunks" not in mode and args.scheme == "uw"
    return ent["traffic_bytes"] if ok else None


def mixed_error(mix, f64, nfix, ncol):
    """Error of the mixed-precision outputs against the FP64 path on the same chunk."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import parity
    worst = np.zeros(ncol)
    for k in f64:
        worst = np.maximum(worst, parity.column_rel_err(mix[k], f64[k]))
    rel_l2 = {}
    for k in ("qvten", "qlten", "qiten", "sten", "uten", "vten", "umf", "precip"):
        den = float((f64[k] ** 2).sum())
        rel_l2[k] = float(np.sqrt(((mix[k] - f64[k]) ** 2).sum() / den)) if den > 0 else 0.0
    return {"recomputed_fp64_columns": int(nfix), "recomputed_fraction": nfix / ncol,
            "tendency_rel_l2_vs_fp64": rel_l2,
            "column_rel_err_vs_fp64": {"median": float(np.median(worst)), "p90": float(np.quantile(worst, 0.9)),
                                       "p99": float(np.quantile(worst, 0.99)), "max": float(worst.max())},
            "reference_harness_rms_vs_fp64": parity.reference_rms_errors(mix, f64)}


def host_cpu():
    """Model name and logical CPU count of the host the CPU baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


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
    if args.scheme == "zw":
        zc = oracle.default_consts(rpen=15.0, rkm=10.0)

        def orun(chunk, nthreads, portable=False):
            return oracle.run_zw(chunk, oracle.zw_defaults(), zc, portable=portable, nthreads=nthreads)
    else:
        def orun(chunk, nthreads, portable=False):
            return oracle.run(chunk, nthreads=nthreads, portable=portable)
    orun(small, nthr)                                     # warm the library and the thread pool
    t0 = time.perf_counter()
    ref, rdiag, rc, _ = orun(sub, nthr)                   # host libm: the faithful reading of the reference
    dt_cpu = time.perf_counter() - t0
    n1 = min(2048, n)                                     # one thread, on a smaller prefix
    one = Chunk(n1, ch.pver, ch.ncnst, ch.dt, {k: np.ascontiguousarray(v[..., :n1]) for k, v in ch.fields.items()})
    t1 = time.perf_counter()
    orun(one, 1)
    dt_one = time.perf_counter() - t1
    pm, pmdiag, _, _ = orun(sub, nthr, portable=True)
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
                                   f"{nthr} threads over columns",
                         "single_thread_value": n1 / dt_one, "single_thread_sample": f"first {n1} columns, 1 thread",
                         "host_cpu": host_cpu()},
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
