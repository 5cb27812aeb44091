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
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak

# HBM bytes per launch of the binary64 tile kernel on workload C2, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by the same counters
# on a plain 256 MiB copy (`bench.py --calib-copy`): FETCH x2.0, WRITE x1.0.
# Summary and commands: profiles/k_tile_C2_2048sq_1e-300.md.
MEASURED_TRAFFIC = {("C2", 0): 316.3e6}


def calib_copy():
    """Plain device copy of known size (256 MiB, three times) against which
    the FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import torch
    n = 256 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": a.numel() * 8, "copies": 3}))


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


def cpu_model():
    """The host CPU's model name (lscpu's "Model name")."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"every {stride}th pixel of the frame ({sample.size} px), {dt:.1f} s, "
                  "C restatement of the reference numba loop, OpenMP dynamic schedule",
    }, stride, (Z, U, sr, si)


def glitch_report(res, fields):
    """The glitch criterion on the binary64 results: pixels left unfinished
    (stop reason -1) and non-finite z_n or dz/dc."""
    Z, _, sr, _ = res
    if fields["kind"] == 2:
        rows = [0, 1] + ([2, 3, 4, 5] if fields["calc_dzndc"] else [])
    else:
        rows = [0] + ([1 + int(bool(fields["interior_detect"]))] if fields["calc_dzndc"] else [])
    nonfinite = np.zeros(Z.shape[1], dtype=bool)
    for r in rows:
        nonfinite |= ~np.isfinite(Z[r])
    return {"unfinished": int(np.count_nonzero(sr < 0)), "non_finite": int(np.count_nonzero(nonfinite))}


def serving_leg(n_requests, precision, rate_frac=0.7, seed=0, layers=False, exterior=False):
    """The request stream: seeded requests of mixed size / width / max_iter
    on eight centres (fractalshades_amd.serving), first all at once (the
    makespan gives the capacity), then with Poisson arrivals at `rate_frac`
    of that capacity (latency under sustained load)."""
    from fractalshades_amd import serving
    from fractalshades_amd.postproc import post_desc
    t0 = time.perf_counter()
    centres = serving.serve_centres()
    if exterior:
        centres = [serving.boundary_centre()]
    specs = serving.serve_requests(n_requests, len(centres), seed=seed)
    bases = {}
    reqs = []
    widths = [sp.nx for sp in specs]        # the pixels are rows of nx: 8x8 blocks
    posts = [] if layers else None
    for sp in specs:
        f, cia, c_pix = serving.build_request(sp, centres, precision)
        if layers:
            posts.append(post_desc(f, "serve"))
        key = cia.resident_key
        if key not in bases:
            bases[key] = cia.ctx            # resident orbit + BLA coefficients of this centre
        else:
            cia = cia.shared_with(bases[key])
        _ = cia.ctx
        reqs.append((cia, c_pix))
    prep_s = time.perf_counter() - t0
    tiles = serving.layer_tiles(reqs) if layers else serving.pinned_tiles(reqs)
    srv = serving.TileServer(lanes=3, chunk_px=65536, hbm_budget=8 << 30)
    try:
        burst = serving.run_stream(srv, reqs, None, tiles=tiles, posts=posts, widths=widths)
        t_first = min(r["t_submit_ms"] for r in burst)
        makespan_s = (max(r["t_done_ms"] for r in burst) - t_first) * 1e-3
        rate = rate_frac * n_requests / makespan_s
        steady = serving.run_stream(srv, reqs, rate, seed=seed + 1, tiles=tiles, posts=posts, widths=widths)
    finally:
        srv.close()

    for rs in (burst, steady):
        for r, sp in zip(rs, specs):
            r["shape"] = (sp.nx, sp.ny)

    def lat(rs, shape=None):
        v = np.array([r["latency_ms"] for r in rs if shape is None or r["shape"] == shape])
        if v.size == 0:
            return None
        return {"n": int(v.size), "p50_ms": float(np.percentile(v, 50)),
                "p99_ms": float(np.percentile(v, 99)), "max_ms": float(v.max())}

    pix_iters = sum(r["pixel_iters"] for r in burst)
    span = (max(r["t_done_ms"] for r in steady) - min(r["t_submit_ms"] for r in steady)) * 1e-3
    out = {
        "requests": n_requests, "precision": precision, "host_prep_s": round(prep_s, 1),
        "frames": "one boundary centre (the seeded 1e-2000 location), exterior-dominated" if exterior
                  else "eight nuclei, offset within 0.3 widths",
        "returns": "nu/dem float32 layers + stop reason/iteration (13 B/px)" if layers
                   else "raw Z/U/stop reason/iteration (57 B/px)",
        "pixels": int(sum(r["pixels"] for r in burst)), "pixel_iters": int(pix_iters),
        "burst": {"makespan_s": makespan_s, "pixel_iters_per_s": pix_iters / makespan_s,
                  "latency": lat(burst)},
        "poisson": {"rate_req_per_s": rate, "span_s": span,
                    "pixel_iters_per_s": sum(r["pixel_iters"] for r in steady) / span,
                    "latency": lat(steady), "latency_512sq": lat(steady, (512, 512)),
                    "latency_by_size": {f"{nx}x{ny}": lat(steady, (nx, ny)) for nx, ny in serving.SIZES}},
    }
    for t in tiles:
        t.close()
    for cia, _ in reqs:
        cia.close()
    return out


def movie_leg(n_frames, nx=480, ny=270, max_iter=500000, chunk_px=1 << 21):
    """A zoom sequence through the serving loop: n_frames frames, width
    geometric from 4 down to 1e-2000, onto one 2100-digit nucleus whose orbit
    is computed unwrapped to max_iter and kept resident (one base context per
    arithmetic mode; every frame shares its orbit and BLA coefficients).  The
    frames are submitted at once and return the nu / dem layers; binary64
    and fp32x are run on the same frames and their escape counts compared."""
    from fractalshades_amd import serving
    from fractalshades_amd.postproc import post_desc
    t0 = time.perf_counter()
    centre = serving.movie_centre(max_iter)
    orbit_s = time.perf_counter() - t0
    specs = serving.movie_requests(n_frames, nx, ny, max_iter)
    out = {"workload": f"zoom sequence: {n_frames} of the 2000 frames of the full sequence (same width "
                       f"schedule, every {2000 // max(n_frames, 1)}th frame), {nx}x{ny} px, max_iter "
                       f"{max_iter}, one resident unwrapped orbit of {max_iter + 1} points; fewer frames "
                       "because each frame's derivative paths over the whole orbit are prepared on the host",
           "frames": n_frames, "shape": f"{nx}x{ny}", "max_iter": max_iter,
           "widths": f"{specs[0].width} .. {specs[-1].width}", "orbit_points": max_iter + 1,
           "chunk_px": chunk_px,
           "orbit_host_s": round(orbit_s, 1)}
    results = {}
    for precision in ("fp64", "fp32x"):
        t0 = time.perf_counter()
        bases, reqs, posts = {}, [], []
        for sp in specs:
            f, cia, c_pix = serving.build_request(sp, [centre], precision)
            posts.append(post_desc(f, "serve"))
            key = cia.resident_key
            if key not in bases:
                bases[key] = cia.ctx
            else:
                cia = cia.shared_with(bases[key])
            _ = cia.ctx
            reqs.append((cia, c_pix))
        prep_s = time.perf_counter() - t0
        tiles = serving.layer_tiles(reqs)
        srv = serving.TileServer(lanes=3, chunk_px=chunk_px, hbm_budget=8 << 30)
        try:
            rs = serving.run_stream(srv, reqs, None, tiles=tiles, posts=posts,
                                    widths=[sp.nx for sp in specs])
        finally:
            srv.close()
        span = (max(r["t_done_ms"] for r in rs) - min(r["t_submit_ms"] for r in rs)) * 1e-3
        pix_iters = sum(r["pixel_iters"] for r in rs)
        results[precision] = [np.array(t.stop_iter, copy=True) for t in tiles]
        out[precision] = {"host_prep_s": round(prep_s, 1), "makespan_s": span,
                          "frames_per_s": n_frames / span, "pixel_iters": int(pix_iters),
                          "pixel_iters_per_s": pix_iters / span}
        for t in tiles:
            t.close()
        for cia, _ in reqs:
            cia.close()
    mis = sum(int(np.count_nonzero(a != b)) for a, b in zip(results["fp64"], results["fp32x"]))
    npx = sum(a.size for a in results["fp64"])
    out["fp32x"]["speedup_vs_fp64"] = out["fp64"]["makespan_s"] / out["fp32x"]["makespan_s"]
    out["fp32x"]["escape_count_mismatch_frac"] = mis / npx
    return out


class Leg:
    """One workload x precision on the GPU: device buffers + timed steps."""

    def __init__(self, f, cia, precision, torch, dev, grid=True):
        from fractalshades_amd._native import work_bytes
        self.f = f
        self.grid_w = f.nx if grid else 0     # 8x8 pixel blocks per wave, or index order
        self.cia = cia if precision == cia.fields["precision"] else cia.with_fields(precision=precision)
        self.ctx = self.cia.ctx
        self.torch = torch
        fields = self.cia.fields
        self.fields = fields
        c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
        self.c_pix = c_pix
        n = self.n = c_pix.size
        from fractalshades_amd._native import n_z_fields
        self.nz = n_z_fields(fields)
        z_words = self.nz * n * (1 if fields["kind"] == 2 else 2)
        self.d_pix = torch.from_numpy(c_pix.view(np.float64)).to(dev)
        self.d_Z = torch.empty(z_words, dtype=torch.float64, device=dev)
        self.d_U = torch.empty(n, dtype=torch.int32, device=dev)
        self.d_sr = torch.empty(n, dtype=torch.int8, device=dev)
        self.d_si = torch.empty(n, dtype=torch.int32, device=dev)
        self.d_work = torch.zeros(work_bytes() // 8, dtype=torch.int64, device=dev)
        self.stream = torch.cuda.current_stream(dev)

    def step(self):
        self.ctx.tile_device(self.d_pix.data_ptr(), self.n, self.d_Z.data_ptr(), self.d_U.data_ptr(),
                             self.d_sr.data_ptr(), self.d_si.data_ptr(), self.d_work.data_ptr(),
                             self.stream.cuda_stream, grid_w=self.grid_w)

    def run(self, steps, warmup):
        from fractalshades_amd._native import work_stats
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(self.stream)
        for _ in range(steps):
            self.step()
        ev1.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        self.ms_per_step = wall * 1e3 / max(steps, 1)
        self.kernel_ms = ev0.elapsed_time(ev1) / max(steps, 1)
        self.pix_iters = int(self.d_si.to(torch.int64).sum().item())
        self.reasons = torch.bincount(self.d_sr.to(torch.int64) + 1, minlength=4).cpu().tolist()
        self.stats = work_stats(self.d_work.data_ptr())
        self.value = self.pix_iters / (self.ms_per_step * 1e-3)
        fp, fb, fl, fr = flops_per_unit(self.fields)
        st = self.stats
        self.flops = st["plain_steps"] * fp + st["bla_steps"] * fb + st["bla_lookups"] * fl + st["rebases"] * fr
        self.tflops = self.flops / (self.kernel_ms * 1e-3) / 1e12
        self.fpu = (fp, fb, fl, fr)
        return self

    def host_results(self):
        n = self.n
        Z = self.d_Z.cpu().numpy()
        Z = Z.reshape(self.nz, n) if self.fields["kind"] == 2 else Z.view(np.complex128).reshape(self.nz, n)
        return (Z, self.d_U.cpu().numpy().reshape(1, n), self.d_sr.cpu().numpy().reshape(1, n),
                self.d_si.cpu().numpy().reshape(1, n))

    def summary(self):
        # the binary32 form's flops are priced against the FP32 vector peak
        peak = FP32_VECTOR_PEAK_TFLOPS if self.fields["precision"] == 1 else FP64_VECTOR_PEAK_TFLOPS
        return {"value": self.value, "ms_per_step": self.ms_per_step, "kernel_ms": self.kernel_ms,
                "pixel_iters_per_step": self.pix_iters, "fp64_equiv_tflops": self.tflops,
                "vector_peak_tflops": peak, "vector_peak_frac": self.tflops / peak,
                "counts": self.stats,
                "stop_reason_hist": dict(zip(["-1", "max_iter", "divergence", "stationnary"], self.reasons))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--deep", default="C2_2000", help="second frame (1e-2000); '' to skip")
    ap.add_argument("--extra", default="C2_int,C3_d3,C3_d4,C3_d6,C3_bs",
                    help="further workloads reported under config.frames (fewer steps); '' to skip")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32x", action="store_true", help="skip the binary32 legs")
    ap.add_argument("--serve", type=int, default=512, help="requests of the serving stream (0: skip)")
    ap.add_argument("--movie-chunk", type=int, default=1 << 21,
                    help="pixels per chunk of the zoom sequence's server (default: a 1920x1080 "
                         "frame is one chunk, so its longest pixels run in one launch together)")
    ap.add_argument("--movie", type=int, default=24,
                    help="frames of the zoom-sequence leg at 1920x1080, max_iter 5e6, 5e6-point orbit (0: skip)")
    ap.add_argument("--order", choices=("grid", "linear"), default="grid",
                    help="pixel order of the kernel: 8x8 blocks of the frame, or index order")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 256 MiB device copy (PMC counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        calib_copy()
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; "
                          "multi-GPU sharding of the tile loop is a design note, not measured",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from fractalshades_amd import configs
    from fractalshades_amd.postproc import error_report

    dev = torch.device("cuda:0")
    frames = {}
    main_leg = None
    def cached(name):
        try:
            return configs.center_for(name, compute=False) is not None
        except KeyError:
            return False
    extra = [x for x in args.extra.split(",") if x and cached(x)]
    for wname in [args.workload] + ([args.deep] if args.deep else []) + extra:
        w = configs.WORKLOADS[wname]
        steps, warmup = (args.steps, args.warmup) if wname not in extra else \
            (max(2, args.steps // 5), min(args.warmup, 1))
        t0 = time.perf_counter()
        f = configs.build(w)
        cia = f._calc_data["bench"]["cycle_indep_args"]
        prep_s = time.perf_counter() - t0
        grid = args.order == "grid"
        leg64 = Leg(f, cia, 0, torch, dev, grid).run(steps, warmup)
        res64 = leg64.host_results()
        entry = {"model": w.model + (f" d={w.degree}" if w.model == "MN" else ""), "steps": steps,
                 "width": w.width, "pixels": leg64.n, "host_prep_s": round(prep_s, 2),
                 "ref_len": cia.fields["ref_len"], "ref_order": f.FP_params["order"],
                 "ref_div_iter": f.FP_params["div_iter"], "xr_mode": cia.fields["xr_mode"],
                 "fp64": leg64.summary()}
        entry["fp64"]["glitched"] = glitch_report(res64, cia.fields)
        if not args.no_fp32x:
            leg32 = Leg(f, cia, 1, torch, dev, grid).run(steps, warmup)
            entry["fp32x"] = leg32.summary()
            entry["fp32x"]["speedup_vs_fp64"] = leg32.value / leg64.value
            entry["fp32x"]["error_vs_fp64"] = error_report(
                leg32.host_results(), res64, holomorphic=cia.fields["kind"] != 2,
                d=cia.fields["degree"] if cia.fields["kind"] == 1 else 2)
            del leg32
        frames[wname] = entry
        if main_leg is None:
            main_leg = leg64
        else:
            del leg64
    leg = main_leg
    fp, fb, fl, fr = leg.fpu
    w = configs.WORKLOADS[args.workload]
    out = {
        "metric": "pixel-iters/s at 1e-300 & 1e-2000 (fp64 vs fp32+xexp); p50/p99 tile latency",
        "value": leg.value,
        "unit": "pixel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg.ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+xexp" if leg.fields["xr_mode"] else "f64",
        "data": "synthetic (seeded nucleus-search frame)",
        "config": {
            "workload": f"{w.name}: Perturbation_mandelbrot {leg.f.nx}x{leg.f.ny} px, width {w.width}, "
                        f"max_iter {w.max_iter}, interior_detect {w.interior_detect}, "
                        f"BLA_eps 1e-6, FP64 drop-in (Xrange deltas)",
            "pixels": leg.n, "pixel_iters_per_step": leg.pix_iters,
            "pixel_order": "8x8 blocks" if args.order == "grid" else "index",
            "frames": frames,
        },
        "roofline": {
            "bound": "valu_fp64",
            "achieved": leg.tflops,
            "peak": FP64_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": leg.tflops / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": MEASURED_TRAFFIC.get((args.workload, 0)),
            "kernel_ms": leg.kernel_ms,
            "flops_per_unit": {"plain": fp, "bla": fb, "lookup": fl, "rebase": fr},
        },
    }
    if args.serve > 0:
        serve = {"fp64_layers": serving_leg(args.serve, "fp64", layers=True),
                 "fp64_raw": serving_leg(args.serve, "fp64"),
                 "fp64_layers_exterior": serving_leg(args.serve, "fp64", layers=True, exterior=True)}
        if not args.no_fp32x:
            serve["fp32x_layers"] = serving_leg(args.serve, "fp32x", layers=True)
        out["config"]["serving"] = serve
        lat512 = serve["fp64_layers"]["poisson"]["latency_512sq"]
        if lat512:
            out["tile_latency_512sq_ms"] = {"p50": lat512["p50_ms"], "p99": lat512["p99_ms"]}
    if args.movie > 0:
        out["config"]["zoom_sequence"] = movie_leg(args.movie, nx=1920, ny=1080, max_iter=5_000_000,
                                                   chunk_px=args.movie_chunk)
    if not args.no_cpu:
        base, stride, ref = cpu_baseline(leg.fields, leg.cia.arrays, leg.c_pix)
        # the same sample checked against the GPU's results of the timed frame
        got = [np.ascontiguousarray(a[:, ::stride]) for a in leg.host_results()]
        same = (np.array_equal(got[0].view(np.uint64), np.ascontiguousarray(ref[0]).view(np.uint64))
                and all(np.array_equal(g, r) for g, r in zip(got[1:], ref[1:])))
        far = int(np.count_nonzero(np.abs(got[3].astype(np.int64) - ref[3]) > 1))
        base["oracle_check"] = {"pixels": int(ref[3].size), "bit_exact": bool(same),
                                "escape_count_beyond_pm1": far}
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
