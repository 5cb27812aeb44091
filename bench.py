"""Benchmark: pixel-iterations/s of the perturbation loop on one MI355X.

One step = one pass of the GPU engine over the whole frame of the workload
(the seeded 2048x2048 frame at width 1e-300, max_iter 200000, interior
detection on, Xrange deltas -- workload C2 of fractalshades_amd.configs),
with the pixel offsets already resident in HBM.  pixel-iterations =
sum(stop_iter) (BLA-skipped iterations included).

Prints one JSON line.  Also reports:
  * roofline: algorithmic FP64 flop/s of the tile kernel (plain / BLA /
    lookup / rebase counts x flops per unit) against the MI355X FP64 vector
    peak, timed with HIP events on the launch stream;
  * cpu_baseline: the CPU restatement of the reference loop (oracle/) on a
    strided sample of the same frame, all host cores, OpenMP.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np

FP64_VECTOR_PEAK_TFLOPS = 78.6   # MI355X published FP64 vector / matrix peak


def flops_per_unit(fields):
    """Algorithmic flops (add/mul = 1, sqrt = 1) per plain step, per BLA jump,
    per BLA lookup and per rebase, from the model formulas (same count for
    binary64 and Xrange deltas)."""
    kind = fields["kind"]
    dc = bool(fields["calc_dzndc"])
    dz = bool(fields["interior_detect"])
    if kind == 2:    # burning ship: (x, y) 18 + Jacobian 48 + escape 5
        plain = 18 + 5 + (48 if dc else 0)
        bla = 14 + (12 if dc else 0)
        return plain, bla, 4, 4 if dc else 0
    d = fields["degree"]
    zn = 12 if kind == 0 else 12 + 16 * (d - 2)
    plain = zn + 5 + (18 if dc else 0) + (23 if dz else 0)
    bla = 14 + (6 if dc else 0) + (6 if dz else 0)
    reb = 2 * int(dc) + 2 * int(dz)
    return plain, bla, 4, reb


def cpu_baseline(fields, arrays, c_pix, budget_s=15.0):
    """Oracle on a strided sample of the frame: grow the sample until it
    takes about `budget_s` seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    ctx = orc.OracleCtx(fields, arrays)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    stride = 4096
    while True:
        sample = np.ascontiguousarray(c_pix[::stride])
        t0 = time.perf_counter()
        Z, U, sr, si = ctx.cycles(sample, threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 8 or stride == 1:
            break
        stride = max(1, stride // 8)
    return {
        "value": float(si.astype(np.int64).sum() / dt),
        "unit": "pixel-iters/s",
        "cores": threads,
        "kind": "port",
        "sample": f"every {stride}th pixel of the frame ({sample.size} px), {dt:.1f} s, "
                  "C restatement of the reference numba loop, OpenMP dynamic schedule",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; "
                          "multi-GPU sharding of the tile loop is a design note, not measured",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from fractalshades_amd import configs
    from fractalshades_amd._native import work_bytes, work_stats

    dev = torch.device("cuda:0")
    w = configs.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    f = configs.build(w)
    cia = f._calc_data["bench"]["cycle_indep_args"]
    ctx = cia.ctx
    prep_s = time.perf_counter() - t0
    fields = cia.fields
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
    n = c_pix.size
    kind = fields["kind"]
    if kind == 2:
        nz = 2 + (4 if fields["calc_dzndc"] else 0)
        z_words = nz * n
    else:
        nz = 1 + int(bool(fields["calc_dzndc"])) + int(bool(fields["interior_detect"]))
        z_words = 2 * nz * n

    d_pix = torch.from_numpy(c_pix.view(np.float64)).to(dev)
    d_Z = torch.empty(z_words, dtype=torch.float64, device=dev)
    d_U = torch.empty(n, dtype=torch.int32, device=dev)
    d_sr = torch.empty(n, dtype=torch.int8, device=dev)
    d_si = torch.empty(n, dtype=torch.int32, device=dev)
    d_work = torch.zeros(work_bytes() // 8, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        ctx.tile_device(d_pix.data_ptr(), n, d_Z.data_ptr(), d_U.data_ptr(), d_sr.data_ptr(),
                        d_si.data_ptr(), d_work.data_ptr(), stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = ev0.elapsed_time(ev1) / max(args.steps, 1)

    pix_iters = int(d_si.to(torch.int64).sum().item())
    reasons = torch.bincount(d_sr.to(torch.int64) + 1, minlength=4).cpu().tolist()
    st = work_stats(d_work.data_ptr())
    fp, fb, fl, fr = flops_per_unit(fields)
    flops = st["plain_steps"] * fp + st["bla_steps"] * fb + st["bla_lookups"] * fl + st["rebases"] * fr
    achieved = flops / (dev_ms * 1e-3) / 1e12
    ms_per_step = wall * 1e3 / max(args.steps, 1)
    value = pix_iters / (ms_per_step * 1e-3)

    out = {
        "metric": "pixel-iters/s at 1e-300 & 1e-2000 (fp64 vs fp32+xexp); p50/p99 tile latency",
        "value": value,
        "unit": "pixel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+xexp" if fields["xr_mode"] else "f64",
        "data": "synthetic (seeded nucleus-search frame)",
        "config": {
            "workload": f"{w.name}: Perturbation_mandelbrot {f.nx}x{f.ny} px, width {w.width}, "
                        f"max_iter {w.max_iter}, interior_detect {w.interior_detect}, "
                        f"BLA_eps 1e-6, FP64 drop-in (Xrange deltas)",
            "pixels": n, "pixel_iters_per_step": pix_iters,
            "ref_len": fields["ref_len"], "ref_order": f.FP_params["order"],
            "ref_div_iter": f.FP_params["div_iter"],
            "stop_reason_hist": {"-1": reasons[0], "max_iter": reasons[1],
                                 "divergence": reasons[2], "stationnary": reasons[3]},
            "host_prep_s": round(prep_s, 2),
        },
        "roofline": {
            "bound": "valu_fp64",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": None,
            "kernel_ms": dev_ms,
            "counts": st,
            "flops_per_unit": {"plain": fp, "bla": fb, "lookup": fl, "rebase": fr},
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(fields, cia.arrays, c_pix)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
