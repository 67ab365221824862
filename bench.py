"""Benchmark: UW shallow-cumulus columns/s on one MI355X.

Headline workload: the forward of the column physics on a 1 036 800-column chunk (a
1440x720 grid) of 35 levels and 3 constituents with half the columns convecting,
inputs resident in HBM, timed in FP64 (UWSHCU_FP64, the reference's arithmetic bit for
bit) and, in the same run, in the mixed-precision mode (UWSHCU_MIXED: the fast FP64
pass whose saturation adjustment is solved to convergence, the conservation/branch
check and the FP64 recompute of the columns it lists).  `value` is the FP64 forward's
columns/s: it is exact and, on this chip, the faster of the two.  The line also
carries the mixed mode's columns/s, its speed relative to FP64 and its error per
output against the FP64 path (all columns) and against the oracle (sample).
One step = one call of the forward over the whole chunk.  Timing: W untimed warm-up
steps, then K steps between two synchronisations, with HIP events recorded on the
stream the kernels run on.

Other modes (same JSON contract; metric and config name the mode):
  --no-mixed          the FP64 forward alone (the 32 768-column chunk: --ncol 32768
                      --inversion-frac 0.25 --no-mixed).
  --precision mixed   time the mixed mode as the line's mode (and FP64 beside it
                      unless --no-mixed).
  --adjoint           forward + adjoint (gradients of the calibration loss
                      sum (sten/cp - tau_T)^2 + (qvten - tau_q)^2 with respect to t0,
                      qv0, u0, v0, tke, pblh, rpen, rle, rkm); one step = FP64 forward
                      plus the adjoint.  The 65 536-column workload: --ncol 65536.

The JSON line also carries:
  roofline      the binding axis: FP64 vector ALU.  `achieved` = the reference
                algorithm's FP64 operations (add, mul, div, sqrt; the transcendental
                functions as the operations of their portable evaluation), counted per
                column by the instrumented oracle on a random sample of the chunk, times
                ncol, over the timed step; peak 78.6 TFLOP/s (MI355X FP64 vector, which
                counts a fused multiply-add as two).  `hbm` beside it: the algorithmic
                bytes per column (the FP64 inputs the physics reads plus the full output
                set, 8*[9*(pver+1) + 24*pver + 2*ncnst*pver + 9]) over the same time,
                against 8 TB/s, and the measured traffic (rocprofv3, profiles/);
  cpu_baseline  the CPU oracle (FP64 C restatement of the reference with the host
                libm, OpenMP over columns) on a bounded sample of the same chunk;
  error_vs_oracle  on that sample: the FP64 path's bit-identity to the portable-math
                oracle and the mixed mode's error against it;
  transfers     host-to-device copy of the inputs and device-to-host copy of the
                outputs, timed separately (they are not part of `value`).
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
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncol", type=int, default=1036800)
    ap.add_argument("--pver", type=int, default=35)
    ap.add_argument("--ncnst", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="the mode `value` times; the other is timed beside it unless --no-mixed")
    ap.add_argument("--no-mixed", action="store_true", help="time the --precision mode alone")
    ap.add_argument("--scheme", choices=["uw", "zw"], default="uw",
                    help="uw: compute_uwshcu_inv; zw: the modified scheme compute_zwshcu_inv (FP64, default options)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--inversion-frac", type=float, default=0.48,
                    help="fraction of columns capped by an inversion (0.48: half the columns convect)")
    ap.add_argument("--cpu-sample", type=int, default=262144, help="columns timed on the CPU oracle")
    ap.add_argument("--op-sample", type=int, default=4096,
                    help="columns whose FP64 operations the instrumented oracle counts")
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
    mix_out = None
    compare = not args.adjoint and args.chunks == 1 and not zw and not args.no_mixed
    if compare:
        # the other precision mode on the same chunk, then the mixed mode's error against FP64
        main_out = {k: v.clone() for k, v in out.items()}
        main_out["cush"] = cush.clone()
        nfix = ctx.mixed_recomputed(stream) if prec == _capi.MIXED else None
        other = _capi.FP64 if prec == _capi.MIXED else _capi.MIXED
        o_ms, _ = timed(lambda: fwd(other), args.steps, args.warmup, stream)
        rc, bad = ctx.status(stream)
        if rc:
            raise _capi.UwshcuError(rc, bad)
        other_out = {k: v.clone() for k, v in out.items()}
        other_out["cush"] = cush.clone()
        if nfix is None:
            nfix = ctx.mixed_recomputed(stream)
        oms = float(np.mean(o_ms))
        f64_out, mix_out = (main_out, other_out) if prec == _capi.FP64 else (other_out, main_out)
        f64_ms, mix_ms = (ms, oms) if prec == _capi.FP64 else (oms, ms)
        extra["fp64_forward"] = {"ms_per_step": f64_ms, "columns_per_s": args.ncol / (f64_ms * 1e-3)}
        extra["mixed_forward"] = {"ms_per_step": mix_ms, "columns_per_s": args.ncol / (mix_ms * 1e-3),
                                  "step_ms_median": float(np.median(o_ms if prec == _capi.FP64 else per_step_ms))}
        extra["mixed_speedup_vs_fp64"] = f64_ms / mix_ms
        extra["mixed"] = mixed_error(mix_out, f64_out, nfix, args.ncol)
        # the FP32-plume pass, kept as an experiment: its speed and error on the same chunk
        raw_ms, _ = timed(lambda: fwd(_capi.MIXED_UNCHECKED), max(2, args.steps // 2), 1, stream)
        raw = dict(out)
        raw["cush"] = cush
        e = mixed_error(raw, mix_out, 0, args.ncol, brief=True)
        extra["fp32_plume_pass_experimental"] = {"ms_per_step": float(np.mean(raw_ms)),
                                                 "tendency_rel_l2_vs_mixed": e["tendency_rel_l2"]}
        # leave the FP64 result in place for the oracle comparison
        for k, v in f64_out.items():
            (cush if k == "cush" else out[k]).copy_(v)
        torch.cuda.synchronize()
    bpc = algorithmic_bytes_per_column(args.pver, args.ncnst)
    achieved_gbs = bpc * args.ncol / (ms * 1e-3) / 1e9
    mode = "fp64 forward" if prec == _capi.FP64 else "mixed forward"
    if args.adjoint:
        mode = "fp64 forward + adjoint"
    if args.chunks > 1:
        mode += f" over {args.chunks} chunks in one call"
    scheme = "compute_zwshcu_inv (default options, rpen 15, rkm 10)" if zw else "compute_uwshcu_inv"
    extra["transfers"] = transfers(ch, out, args.tile)
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
        "dtype": "f64",
        "data": "synthetic" if args.tile == 1 else f"synthetic, {args.ncol // args.tile} distinct columns tiled {args.tile}x",
        "config": {
            "workload": f"{mode} of {scheme}, {args.ncol} columns x {args.pver} levels x "
                        f"{args.ncnst} constituents, dt={ch.dt:g} s",
            "ncol": args.ncol, "pver": args.pver, "ncnst": args.ncnst,
            "convecting_fraction": conv_frac, "inversion_frac": args.inversion_frac, "seed": args.seed,
            "workspace_gib": ws_gib if ws_gib else 24.0,
        },
        "roofline": None,
        "step_ms_min": float(np.min(per_step_ms)),
        "step_ms_median": float(np.median(per_step_ms)),
        "wall_s_timed": wall,
    }
    res.update(extra)
    hbm = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(mode, args),
           "algorithmic_bytes_per_launch": bpc * args.ncol, "algorithmic_bytes_per_column": bpc}
    ops = None
    if not args.no_cpu:
        leg = cpu_leg(args, ch, out, cush, mix_out)
        ops = leg.pop("op_count", None)
        res.update(leg)
    res["roofline"] = roofline(ops, args.ncol, ms, hbm, args.chunks > 1)
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


ERR_FIELDS = ("sten", "qvten", "qlten", "qiten", "uten", "vten", "trten", "umf", "slflx", "qtflx", "qrten", "qsten",
              "evapc", "precip", "snow", "cufrc")


def _col_rel_err(a, b):
    """Per-column max |a - b| / max |b| (floor 1e-12 of the field maximum), on the device."""
    import torch
    a2, b2 = a.reshape(-1, a.shape[-1]), b.reshape(-1, b.shape[-1])
    floor = torch.clamp(b2.abs().max() * 1e-12, min=1e-300)
    return (a2 - b2).abs().max(dim=0).values / torch.maximum(b2.abs().max(dim=0).values, floor)


def mixed_error(mix, f64, nfix, ncol, brief=False):
    """Error of the mixed-mode outputs against the FP64 path on the same chunk, computed on
    the GPU over every column: relative L2 per output over the chunk (the bound is 1e-4),
    the per-column error (worst output) distribution and the integer-diagnostic matches."""
    import torch
    l2 = {}
    for k in ERR_FIELDS:
        den = torch.linalg.vector_norm(f64[k])
        l2[k] = float(torch.linalg.vector_norm(mix[k] - f64[k]) / den) if float(den) > 0 else 0.0
    if brief:
        return {"tendency_rel_l2": l2}
    colw = torch.zeros(ncol, dtype=torch.float64, device=f64["sten"].device)
    for k in ERR_FIELDS:
        colw = torch.maximum(colw, _col_rel_err(mix[k], f64[k]))
    samp = colw[:: max(1, ncol // 1000000)]
    q = torch.quantile(samp[: 1 << 24], torch.tensor([0.5, 0.99, 0.999], dtype=torch.float64, device=samp.device))
    return {"recomputed_fp64_columns": int(nfix), "recomputed_fraction": nfix / ncol,
            "tendency_rel_l2_vs_fp64": l2, "bound": 1e-4,
            "max_rel_l2": max(l2.values()),
            "column_rel_err_vs_fp64": {"median": float(q[0]), "p99": float(q[1]), "p999": float(q[2]),
                                       "max": float(colw.max()), "columns_above_1e-4": int((colw > 1e-4).sum())},
            "integer_diag_mismatch_columns": int((mix["diag"] != f64["diag"]).any(dim=0).sum())}


def transfers(ch, out, tile):
    """Host-to-device copy of the chunk's inputs (from pageable numpy arrays) and
    device-to-host copy of the outputs, each timed once after a warm copy."""
    import torch
    from uwshcu_amd import device
    nbytes_in = sum(v.nbytes for v in ch.fields.values())
    device.chunk_to_device(ch, "cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    device.chunk_to_device(ch, "cuda:0")
    torch.cuda.synchronize()
    h2d = time.perf_counter() - t0
    nbytes_out = sum(v.numel() * v.element_size() for k, v in out.items() if k != "diag")
    t0 = time.perf_counter()
    for k, v in out.items():
        if k != "diag":
            v.cpu()
    d2h = time.perf_counter() - t0
    return {"h2d_ms": h2d * 1e3, "h2d_bytes": int(nbytes_in), "h2d_gbs": nbytes_in / h2d / 1e9,
            "d2h_ms": d2h * 1e3, "d2h_bytes": int(nbytes_out), "d2h_gbs": nbytes_out / d2h / 1e9,
            "note": "pageable host memory, not part of the timed step" +
                    (f"; inputs are the {ch.ncol} distinct columns" if tile > 1 else "")}


def roofline(ops, ncol, ms, hbm, chunked):
    """The binding axis (FP64 vector ALU) from the counted operations, HBM beside it."""
    r = {"bound": "valu", "achieved": None, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": None,
         "traffic": hbm["traffic"],
         "note": ("FP64 vector ALU, not MFMA: the column physics is scalar per column.  achieved = the reference "
                  "algorithm's FP64 operations (add, mul, div, sqrt; transcendentals as the operations of their "
                  "portable evaluation), counted by the instrumented oracle on a random sample of the chunk, times "
                  "ncol, over the timed step.  The peak counts a fused multiply-add as two operations; divisions "
                  "and square roots cost several instructions each.  traffic = measured HBM bytes per step "
                  "(rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated), against the algorithmic bytes of hbm."),
         "hbm": hbm}
    if ops:
        total = ops["mean_per_column"] * ncol
        achieved = total / (ms * 1e-3) / 1e12
        r.update(achieved=achieved, frac=achieved / FP64_PEAK_TFLOPS, ops_per_step=total, op_count=ops)
    if chunked:
        r["timed"] = "one batched call: table upload, gather, the forward pipeline and scatter"
    return r


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


def cpu_leg(args, ch, out, cush, mix_out=None):
    """The oracle leg: CPU baseline (the oracle on a bounded sample), the GPU results
    against the oracle on that sample, and the FP64 operation count per column."""
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
    n1 = min(4096, n)                                     # one thread, on a smaller prefix
    one = Chunk(n1, ch.pver, ch.ncnst, ch.dt, {k: np.ascontiguousarray(v[..., :n1]) for k, v in ch.fields.items()})
    t1 = time.perf_counter()
    orun(one, 1)
    dt_one = time.perf_counter() - t1
    pm, pmdiag, _, _ = orun(sub, nthr, portable=True)
    got = {k: v[..., :n].cpu().numpy() for k, v in out.items() if k != "diag"}
    got["cush"] = cush[:n].cpu().numpy()
    diag = out["diag"][:, :n].cpu().numpy().T
    errs = parity.compare(got, ref)
    worst_col = np.zeros(n)
    for k in ref:
        worst_col = np.maximum(worst_col, parity.column_rel_err(got[k], ref[k]))
    identical = all(np.array_equal(got[k], pm[k]) for k in pm) and np.array_equal(diag, pmdiag)
    res = {
        "cpu_baseline": {"value": n / dt_cpu, "unit": "columns/s", "cores": nthr, "kind": "port",
                         "sample": f"first {n} columns of the same chunk, FP64 C oracle (host libm), "
                                   f"{nthr} threads over columns ({dt_cpu:.1f} s)",
                         "single_thread_value": n1 / dt_one, "single_thread_sample": f"first {n1} columns, 1 thread",
                         "host_cpu": host_cpu()},
        "error_vs_oracle": {
            "fp64_bit_identical_to_portable_math_oracle": bool(identical),
            "fp64_vs_host_libm_oracle_max_col_rel_err": float(worst_col.max()),
            "fp64_vs_host_libm_oracle_cols_above_1e-10": int((worst_col > 1e-10).sum()),
            "fp64_vs_host_libm_oracle_worst_field": max(errs, key=errs.get),
            "integer_diag_mismatch_columns": int((diag != rdiag).any(axis=1).sum()),
            "sample_columns": n,
        },
    }
    if mix_out is not None:
        mx = {k: v[..., :n].cpu().numpy() for k, v in mix_out.items() if k not in ("diag",)}
        l2 = {k: float(np.linalg.norm(mx[k] - pm[k]) / max(np.linalg.norm(pm[k]), 1e-300)) for k in ERR_FIELDS}
        colw = np.zeros(n)
        for k in ERR_FIELDS:
            colw = np.maximum(colw, parity.column_rel_err(mx[k], pm[k]))
        res["error_vs_oracle"]["mixed_vs_portable_oracle"] = {
            "tendency_rel_l2": l2, "max_rel_l2": max(l2.values()),
            "columns_above_1e-4": int((colw > 1e-4).sum()), "max_col_rel_err": float(colw.max()),
            "integer_diag_mismatch_columns": int((mix_out["diag"][:, :n].cpu().numpy().T != pmdiag).any(axis=1).sum())}
        # the FP64 reference's own sensitivity, for scale: host libm against portable math
        hl = np.zeros(n)
        for k in ERR_FIELDS:
            hl = np.maximum(hl, parity.column_rel_err(ref[k], pm[k]))
        res["error_vs_oracle"]["reference_libm_sensitivity"] = {
            "note": "the FP64 oracle with the host libm against the portable-math oracle on the same sample",
            "max_col_rel_err": float(hl.max()), "columns_above_1e-4": int((hl > 1e-4).sum())}
    if args.scheme == "uw" and args.op_sample > 0:
        rng = np.random.default_rng(args.seed + 17)
        idx = np.sort(rng.choice(ch.ncol, min(args.op_sample, ch.ncol), replace=False))
        ops, _, odiag = oracle.count_ops(Chunk(len(idx), ch.pver, ch.ncnst, ch.dt,
                                               {k: np.ascontiguousarray(v[..., idx]) for k, v in ch.fields.items()}))
        flops = ops[:, :4].sum(axis=1).astype(np.float64)
        conv = odiag[:, 7] > 0
        hist, edges = np.histogram(flops, bins=[0, 2e4, 5e4, 1e5, 2e5, 3e5, 4e5, 5e5, 7e5, 1e6, 1e9])
        res["op_count"] = {
            "mean_per_column": float(flops.mean()), "sample_columns": int(len(idx)),
            "mean_convecting": float(flops[conv].mean()) if conv.any() else 0.0,
            "mean_not_convecting": float(flops[~conv].mean()) if (~conv).any() else 0.0,
            "by_kind_mean": {k: float(v) for k, v in zip(oracle.OP_KINDS, ops.mean(axis=0))},
            "histogram": {"edges": [float(e) for e in edges], "counts": [int(c) for c in hist]}}
    return res


if __name__ == "__main__":
    main()
