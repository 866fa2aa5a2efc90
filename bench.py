"""Benchmark: W-projection grid + degrid throughput on one MI355X.

A step is one grid pass and one degrid pass (IVisGridder::grid and ::degrid, each a full
TableVisGridder::generic: uvw prep, tile binning, kernel) over the whole W-projection array
workload (36 antennas in a 6 km disc, 300 integrations over 8 h at dec -45 deg, 64 channels
700-1000 MHz, 4 polarisations = 12.1M (row, chan) samples x 4 pols; 4096^2 grid, 129 w planes,
oversample 4, support 64), with every input already resident in HBM.

Metric: CF-tap updates/s (one complex multiply-accumulate of a (visibility, pol) with a CF tap =
8 FP32 FLOP, TableVisGridder.cc:264-267,804), whole job.  One JSON line is printed.

After the timed region, the oracle leg (skipped with --no-cpu-baseline) times the oracle's FP32
restatement of the reference loop on the host on bounded samples of the same workload, on one
thread (``cpu_baseline``) and on all cores as independent gridders (``cpu_baseline_all_cores``,
with the host's core count and CPU model), and measures the GPU path's image and degridded-
visibility error against the FP64 oracle on a row subset (``error_vs_fp64``).

``--workload sphfunc_disc`` measures the plain SphFunc case: 160,000 single-pol visibilities
(uv uniform in a 2 km disc at 1.4 GHz, w = 0), 2048^2 grid, support 3, oversample 128, grid then
degrid; HBM- and launch-bound by construction.

``--workload batched_beams`` measures 36 W-projection beams of uneven channel counts (8..64),
visibility counts (log-uniform 1e4..2e6 per polarisation) and grid sizes (1024^2..4096^2): every
beam gridded and degridded per step by one batch (ygrid_batch_*), tuned by its autotuner and
replayed from a HIP graph.

``--workload learned_wdisc`` measures W-projection with learned CFs: 160,000 single-pol
visibilities (uv uniform in a 2 km disc at 1.4 GHz) with continuous w in [-300, 300] m, 2048^2
grid; a coordinate network (4 x 128 SiLU) is trained on the GPU on a 129-plane W-projection table
and then gives every sample its CF at its own w (FP16 weights and CF storage); the plain
129-plane table gridder runs beside it.

``--workload awproject_feeds`` measures the A-projection mixed-precision case instead: 160,000
rows x 32 channels around 1.4 GHz (uv uniform in a 2 km disc), 9 feeds on a 1 degree grid, a
12 m uniform disk, 2048^2 grid; grid + degrid with the CF taps and gridding coefficients stored
in FP32, FP16 and BF16 (FP32 accumulation), the value being the FP16 throughput.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3
FP16_PEAK_TFLOPS = 2500.0  # dense FP16 / BF16 matrix-core peak of MI355X
HBM_PEAK_GBS = 8000.0


PROFILE = "profiles/k_grid_k_degrid_wproject_array.md"
PROFILE_SPH = "profiles/k_grid_small_k_degrid_small_sphfunc_disc.md"
PROFILE_F16 = "profiles/k_grid16_k_degrid16_wproject_array_fp16.md"


def profiled_traffic(kernel, profile=PROFILE):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated) recorded by
    the rocprofv3 --pmc runs of this benchmark summarised in `profile`; None if not recorded."""
    try:
        with open(os.path.join(ROOT, profile)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel.replace(" ", ""):
                    return float(parts[2])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """The host's core count (all, and those this process may run on) and CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def host_threads() -> int:
    """Worker threads for the all-cores CPU baseline: the cores this process may use, at most
    the thread count the environment allows (OMP_NUM_THREADS) and 16."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, min(16, n))


class _CpuPass:
    """One oracle gridder and one degridder (the FP32 restatement of TableVisGridder::generic)
    over a set of rows, set up outside the timing: run() grids the rows, then degrids the same
    rows from the grid it made, and returns the taps of both passes."""

    def __init__(self, wl, cf, prm, ws, rows):
        from oracle import oracle as O
        uvw = wl.uvw[rows]
        cm = np.zeros((wl.nchan, wl.npol, len(rows)), dtype=np.int32)
        for c, f in enumerate(wl.freq):
            cm[c, :, :] = np.array([ws.plane(w / O.C_LIGHT * f) for w in uvw[:, 2]], dtype=np.int32)[None, :]
        self.ch = O.Chunk(uvw, wl.freq, np.transpose(wl.vis[:, :, rows], (2, 1, 0)),
                          noise=np.transpose(wl.noise[:, :, rows], (2, 1, 0)))
        self.og = O.OracleGridder("WProject", "f32", wproject=prm)
        self.og.cf = cf
        self.og.cmap_override = cm
        self.og.initialise_grid((wl.nx, wl.ny, wl.npol, 1), (wl.cell, wl.cell))
        self.od = O.OracleGridder("WProject", "f32", wproject=prm)
        self.od.cf = cf
        self.od.cmap_override = cm
        self.od._setup((wl.nx, wl.ny, wl.npol, 1), (wl.cell, wl.cell))
        self.od.mode = 3
        self.od.freq_axis = None
        self.od.model_empty = False

    def run(self) -> int:
        self.og.grid(self.ch)
        self.od.grid_ = self.og.grid_
        self.od.degrid(self.ch)
        return self.og.counters.taps_gridded + self.od.counters.taps_degridded


def cpu_baseline(wl, g_grid, target_s: float = 12.0):
    """The oracle's FP32 restatement of TableVisGridder::generic on bounded samples of the same
    workload (same CF table, downloaded from the device), grid + degrid:

    * single thread, as one reference gridder runs (the ``cpu_baseline`` object);
    * all host cores: independent gridders, one per thread, each gridding and degridding its own
      slice of rows into its own grid (as the reference's parallel imager gives each worker its
      own gridder and grid; the merge of the workers' grids is not timed).  The C calls release
      the interpreter lock, so the threads run in parallel.
    Returns (single-thread object, all-cores object)."""
    import threading
    from oracle import oracle as O
    os_ = wl.params["oversample"]
    nz = wl.params["nwplanes"]
    cf = O.CFCache([g_grid.cf_plane(k) for k in range(nz * os_ * os_)], os_)
    prm = O.WProjectParams(wmax=wl.params["wmax"], nwplanes=nz, oversample=os_,
                           maxsupport=wl.params["maxsupport"], limitsupport=wl.params["limitsupport"])
    ws = prm.sampling()

    def timed(passes):
        taps = [0] * len(passes)
        start = threading.Barrier(len(passes) + 1)

        def work(i):
            start.wait()
            taps[i] = passes[i].run()

        th = [threading.Thread(target=work, args=(i,)) for i in range(len(passes))]
        for t in th:
            t.start()
        start.wait()
        t0 = time.perf_counter()
        for t in th:
            t.join()
        return sum(taps), time.perf_counter() - t0

    def rows_of(n):
        return np.linspace(0, wl.nrow - 1, n).astype(int)

    n = 4
    taps, dt = timed([_CpuPass(wl, cf, prm, ws, rows_of(n))])
    for _ in range(3):  # the first sizes are dominated by fixed costs: rescale until near the target
        if dt >= 0.5 * target_s or n >= wl.nrow:
            break
        n = min(wl.nrow, max(n + 1, int(n * target_s / max(dt, 1e-3))))
        taps, dt = timed([_CpuPass(wl, cf, prm, ws, rows_of(n))])
    hc = host_cpu()
    one = {"value": taps / dt, "unit": "CF-tap updates/s", "cores": 1, "kind": "port",
           "sample": f"{n} rows x {wl.nchan} chans x {wl.npol} pols of the same workload, grid + degrid, "
                     f"oracle FP32 restatement (gcc -O2, 1 thread), {dt:.1f} s",
           "host": hc}
    T = host_threads()
    rows = rows_of(min(wl.nrow, n * T))
    passes = [_CpuPass(wl, cf, prm, ws, rows[t::T]) for t in range(T)]
    taps_t, dt_t = timed(passes)
    del passes
    allc = {"value": taps_t / dt_t, "unit": "CF-tap updates/s", "cores": T, "kind": "port",
            "sample": f"{len(rows)} rows x {wl.nchan} chans x {wl.npol} pols of the same workload, grid + degrid, "
                      f"{T} independent oracle FP32 gridders (one per thread, rows interleaved, own grids; "
                      f"grid merge not timed), {dt_t:.1f} s wall",
            "host": hc, "speedup_vs_1_thread": (taps_t / dt_t) / (taps / dt)}
    return one, allc


def fp64_errors(ctx, wl, params, model):
    """Error of the GPU path against the FP64 oracle (the checker) on a row subset of the same
    workload: every 97th row at the band's two edge channels, kept where both channels' w planes
    are among the 8 most used (the oracle transforms only those planes).  Image: max |x - x64| /
    max |x64| over the whole image, where the grid-correction gain is <= 10, and divided by the
    gain; degridded visibilities: max |v - v64| / max |v64| (off-grid samples left unchanged on
    both sides)."""
    import yandasoft_amd as ya
    from oracle import oracle as O
    rows = np.arange(0, wl.nrow, 97)
    chans = np.array([0, wl.nchan - 1])
    p = wl.params
    prm = O.WProjectParams(wmax=p["wmax"], nwplanes=p["nwplanes"], cutoff=p["cutoff"], oversample=p["oversample"],
                           maxsupport=p["maxsupport"], limitsupport=p["limitsupport"])
    ws = prm.sampling()
    freq = wl.freq[chans]
    planes = np.array([[ws.plane(wl.uvw[r, 2] / O.C_LIGHT * f) for f in freq] for r in rows])
    vals, cnt = np.unique(planes, return_counts=True)
    keep = set(vals[np.argsort(-cnt)][:8].tolist())
    sel = rows[[i for i in range(len(rows)) if set(planes[i].tolist()) <= keep]]
    uvw = np.ascontiguousarray(wl.uvw[sel])
    vis = np.ascontiguousarray(wl.vis[:, chans][:, :, sel])
    noise = np.ascontiguousarray(wl.noise[:, chans][:, :, sel])
    img = ya.ImageDesc(wl.nx, wl.ny, wl.npol, 1, wl.cell, wl.cell)
    g = ya.Gridder(ctx, "GPUWProject", params)
    g.initialise_grid(img)
    g.grid(ya.Chunk(uvw=uvw, freq=freq, vis=vis, noise=noise))
    gimg = np.asarray(g.finalise_grid(), np.float64)
    och = O.Chunk(uvw, freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)))
    o = O.OracleGridder("WProject", "f64", wproject=prm)
    o.only_planes = set(keep)
    o.initialise_grid((wl.nx, wl.ny, wl.npol, 1), (wl.cell, wl.cell))
    o.grid(och)
    oimg = o.finalise_grid()
    gain = o.correct(np.ones((1, 1, o.ny, o.nx)))[0, 0]
    peak = float(np.max(np.abs(oimg)))
    err = np.abs(gimg.reshape(oimg.shape) - oimg)
    d = ya.Gridder(ctx, "GPUWProject", params)
    d.initialise_degrid(img, model)
    out = np.zeros_like(vis)
    d.degrid(ya.Chunk(uvw=uvw, freq=freq, vis=vis, noise=noise, vis_out=out))
    od = O.OracleGridder("WProject", "f64", wproject=prm)
    od.only_planes = set(keep)
    od.stale_offgrid = False
    od.initialise_degrid((wl.nx, wl.ny, wl.npol, 1), (wl.cell, wl.cell), np.asarray(model, np.float64))
    ov = od.degrid(och)
    verr = float(np.max(np.abs(np.transpose(out, (2, 1, 0)) - ov)) / np.max(np.abs(ov)))
    return {"image_max_err": float(np.max(err)) / peak,
            "image_max_err_gain_le_10": float(np.max(err[..., gain <= 10.0])) / peak,
            "image_max_err_over_gain": float(np.max(err / gain)) / peak,
            "degrid_max_err": verr,
            "sample": f"{len(sel)} rows x 2 chans (band edges) x {wl.npol} pols, {len(keep)} w planes, "
                      f"FP64 oracle (relative to the FP64 peak)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nint", type=int, default=300, help="integrations (workload rows / 630)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle leg: CPU baseline timings and (wproject_array) the FP64 error check")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (3x), the reference for rocprof byte counters")
    ap.add_argument("--cf-precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="wproject_array: CF tap and coefficient storage (the metric is quoted for fp32)")
    ap.add_argument("--workload", default="wproject_array",
                    choices=["wproject_array", "sphfunc_disc", "awproject_feeds", "batched_beams", "learned_wdisc",
                             "imaging_loop", "wstack_array", "snapshot_array"])
    ap.add_argument("--train-epochs", type=int, default=200, help="learned_wdisc: training epochs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is described in DESIGN.md"}))
        sys.exit(2)

    import torch
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.workload == "awproject_feeds":
        bench_awproject(args)
        return
    if args.workload == "batched_beams":
        bench_batched(args)
        return
    if args.workload == "learned_wdisc":
        bench_learned(args)
        return
    if args.workload == "sphfunc_disc":
        bench_sphfunc(args)
        return
    if args.workload == "imaging_loop":
        bench_imaging(args)
        return
    if args.workload == "wstack_array":
        bench_wstack(args)
        return
    if args.workload == "snapshot_array":
        bench_snapshot(args)
        return
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.wproject_array(nint=args.nint)
    log(f"workload rows={wl.nrow} chans={wl.nchan} pols={wl.npol} grid={wl.nx} wmax={wl.params['wmax']:.1f}")
    dev = torch.device("cuda:0")
    t_uvw = torch.from_numpy(np.ascontiguousarray(wl.uvw)).to(dev)
    t_freq = torch.from_numpy(np.ascontiguousarray(wl.freq)).to(dev)
    t_vis = torch.from_numpy(np.ascontiguousarray(wl.vis).view(np.float32)).to(dev)
    t_noise = torch.from_numpy(np.ascontiguousarray(wl.noise).view(np.float32)).to(dev)
    t_out = torch.zeros_like(t_vis)
    torch.cuda.synchronize()
    chunk = ya.Chunk(uvw=t_uvw, freq=t_freq, vis=t_vis, vis_out=t_out, noise=t_noise, device=True)
    chunk.npol = wl.npol

    ctx = ya.Context(0)
    img = ya.ImageDesc(wl.nx, wl.ny, wl.npol, 1, wl.cell, wl.cell)
    prm = dict(wl.params)
    if args.cf_precision != "fp32":
        prm.update(cfprecision=args.cf_precision, visprecision=args.cf_precision)
    gg = ya.Gridder(ctx, "GPUWProject", prm)
    dg = ya.Gridder(ctx, "GPUWProject", prm)
    rng = np.random.default_rng(3)
    model = np.zeros((1, wl.npol, wl.ny, wl.nx), np.float32)
    for _ in range(50):
        model[0, :, rng.integers(500, 3500), rng.integers(500, 3500)] = rng.normal(size=wl.npol)
    t_model = torch.from_numpy(model).to(dev)
    torch.cuda.synchronize()
    dg.initialise_degrid(img, t_model)

    def step():
        # every step bins the samples once: the gridding pass bins, the degridding pass of the
        # same chunk reuses that binning (as in a major cycle); nothing carries over between steps
        ctx.clear_cache()
        gg.initialise_grid(img)
        gg.grid(chunk)
        dg.degrid(chunk)

    t0 = time.perf_counter()
    for i in range(args.warmup):
        step()
        log(f"warmup {i} done ({time.perf_counter() - t0:.1f} s)")
    ctx.synchronize()
    torch.cuda.synchronize()
    s0g, s0d = gg.stats(), dg.stats()
    gg.set_profiling(True)
    dg.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1g, s1d = gg.stats(), dg.stats()
    taps_g = s1g.taps_gridded - s0g.taps_gridded
    taps_d = s1d.taps_degridded - s0d.taps_degridded
    taps = taps_g + taps_d
    value = taps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    # dominant kernel: average launch duration from HIP events on the gridder's stream
    gk_ms = s1g.ms_grid_kernel / max(1.0, s1g.n_grid_launches)
    dk_ms = s1d.ms_degrid_kernel / max(1.0, s1d.n_degrid_launches)
    per_launch_g = taps_g / args.steps
    per_launch_d = taps_d / args.steps
    # kernel names as the profiler prints them: k_grid<pols, waves per SIMD, pipelined, CF precision>,
    # k_degrid_mm<pols, CF precision> (matrix-core degridding)
    tf = lambda b: "true" if b else "false"
    gname = f"k_grid<{wl.npol}, {s1g.grid_waves}, {tf(s1g.grid_pipeline)}, {s1g.cf_precision}>"
    if s1d.degrid_matrix:
        dname = f"k_degrid_mm<{wl.npol}, {s1d.cf_precision}>"
    else:
        dname = f"k_degrid<{wl.npol}, {s1d.degrid_waves}, {tf(s1d.degrid_pipeline)}, {s1d.cf_precision}>"
    peak, prof = FP32_PEAK_TFLOPS, PROFILE
    if args.cf_precision != "fp32":
        prof = PROFILE_F16
        # 16-bit taps and coefficients (and model grid): the 16-bit matrix-core kernels
        gname = f"k_grid16<{wl.npol}, {s1g.cf_precision}>"
        dname = f"k_degrid16<{wl.npol}, {s1d.cf_precision}>"
        peak = FP16_PEAK_TFLOPS
    if gk_ms >= dk_ms:
        kname, kms, kt = gname, gk_ms, per_launch_g
    else:
        kname, kms, kt = dname, dk_ms, per_launch_d
    achieved = 8.0 * kt / (kms * 1e-3) / 1e12
    result = {
        "metric": "CF-tap updates/s (W-proj grid+degrid)",
        "value": value,
        "unit": "CF-tap updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.cf_precision == "fp32" else f"{args.cf_precision} storage, f32 arithmetic",
        "data": "synthetic",
        "config": {"workload": "wproject_array: 36 ant 6 km disc, 300 x 630 rows, 64 ch 700-1000 MHz, 4 pol, "
                               "4096^2, 129 w-planes, os 4, support 64",
                   "rows": wl.nrow, "chans": wl.nchan, "pols": wl.npol, "grid": wl.nx,
                   "wplanes": wl.params["nwplanes"], "oversample": wl.params["oversample"],
                   "support": s1g.support, "parallelism": "1 GPU"},
        "fp32_pct_of_peak": 100.0 * 8.0 * value / (FP32_PEAK_TFLOPS * 1e12),
        "phases_ms_per_step": {"grid_kernel": gk_ms, "degrid_kernel": dk_ms,
                               "prep": (s1g.ms_prep + s1d.ms_prep) / args.steps,
                               "binning": (s1g.ms_bin + s1d.ms_bin) / args.steps},
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": profiled_traffic(kname, prof),
                     "traffic_source": prof},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"], result["cpu_baseline_all_cores"] = cpu_baseline(wl, gg)
        log("error vs the FP64 oracle (row subset) ...")
        result["error_vs_fp64"] = fp64_errors(ctx, wl, prm, model)
        log(f"error vs FP64: {result['error_vs_fp64']}")
        log(f"cpu baseline: 1 thread {result['cpu_baseline']['value'] / 1e6:.1f} Mtaps/s, "
            f"{result['cpu_baseline_all_cores']['cores']} threads "
            f"{result['cpu_baseline_all_cores']['value'] / 1e6:.1f} Mtaps/s")
    print(json.dumps(result), flush=True)


def aw_cpu_baseline(wl, target_s: float = 12.0):
    """Oracle's single-thread FP32 restatement of the A-projection grid + degrid on a bounded row
    sample of the same workload (its own CF generation is not timed)."""
    from oracle import oracle as O
    p = wl.params
    prm = O.AWProjectParams(wmax=p["wmax"], nwplanes=p["nwplanes"], oversample=p["oversample"],
                            maxsupport=p["maxsupport"], maxfeeds=p["maxfeeds"], cutoff=p["cutoff"], diameter=12.0,
                            blockage=0.0)
    lm = {f: O.direction_lm(tuple(wl.pointing1[f]), wl.centre) for f in range(p["maxfeeds"])}

    def run(nrow_s):
        rows = np.linspace(0, wl.nrow - 1, nrow_s).astype(int)
        ch = O.Chunk(wl.uvw[rows], wl.freq, np.transpose(wl.vis[:, :, rows], (2, 1, 0)),
                     noise=np.transpose(wl.noise[:, :, rows], (2, 1, 0)))
        ch.feed1 = wl.feed1[rows]
        og = O.OracleGridder("AWProject", "f32", wproject=prm)
        og.aw_lm, og.aw_nchan = lm, wl.nchan
        og.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
        og._init_aw(ch, False)  # CF generation outside the timed region
        t0 = time.perf_counter()
        og.grid(ch)
        t1 = time.perf_counter()
        od = O.OracleGridder("AWProject", "f32", wproject=prm)
        od.aw_lm, od.aw_nchan = lm, wl.nchan
        od.initialise_degrid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell), np.ones((1, 1, wl.ny, wl.nx)))
        od.cf, od.aw_state, od.aw_done = og.cf, og.aw_state, og.aw_done
        t2 = time.perf_counter()
        od.degrid(ch)
        t3 = time.perf_counter()
        return og.counters.taps_gridded + od.counters.taps_degridded, (t1 - t0) + (t3 - t2)

    n = 200
    taps, dt = run(n)
    for _ in range(3):  # the first sizes are dominated by fixed costs: rescale until near the target
        if dt >= 0.5 * target_s or n >= wl.nrow:
            break
        n = min(wl.nrow, max(n + 1, int(n * target_s / max(dt, 1e-3))))
        taps, dt = run(n)
    return {"value": taps / dt, "unit": "CF-tap updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} rows x {wl.nchan} chans of the same workload, grid + degrid, oracle FP32 "
                      f"restatement (gcc -O2, 1 thread), {dt:.1f} s"}


def bench_awproject(args):
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.awproject_feeds()
    log(f"workload rows={wl.nrow} chans={wl.nchan} feeds={wl.params['maxfeeds']} grid={wl.nx}")
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    chunk = ya.Chunk(uvw=t(wl.uvw), freq=t(wl.freq), vis=t(wl.vis.view(np.float32)),
                     vis_out=torch.zeros(wl.vis.size * 2, dtype=torch.float32, device=dev),
                     noise=t(wl.noise.view(np.float32)), feed1=t(wl.feed1), pointing1=t(wl.pointing1), device=True)
    chunk.npol = 1
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell, centre=wl.centre)
    rng = np.random.default_rng(3)
    model = np.zeros((1, 1, wl.ny, wl.nx), np.float32)
    for _ in range(30):
        model[0, 0, rng.integers(600, 1400), rng.integers(600, 1400)] = rng.normal()
    t_model = t(model)
    ctx = ya.Context(0)
    modes, images, vis_out = {}, {}, {}
    for prec in ("fp32", "fp16", "bf16"):
        params = dict(wl.params, cfprecision=prec, visprecision=prec)
        gg = ya.Gridder(ctx, "GPUAWProject", params)
        dg = ya.Gridder(ctx, "GPUAWProject", params)
        t0 = time.perf_counter()
        gg.initialise_grid(img)
        gg.grid(chunk)  # first pass: CF generation for the 9 feeds x 32 channels, then gridding
        ctx.synchronize()
        first = time.perf_counter() - t0
        cf_s = gg.stats().time_cf_s
        dg.initialise_degrid(img, t_model)

        def step():
            ctx.clear_cache()
            gg.initialise_grid(img)
            gg.grid(chunk)
            dg.degrid(chunk)

        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        s0g, s0d = gg.stats(), dg.stats()
        gg.set_profiling(True)
        dg.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        s1g, s1d = gg.stats(), dg.stats()
        taps_g = s1g.taps_gridded - s0g.taps_gridded
        taps_d = s1d.taps_degridded - s0d.taps_degridded
        gk = s1g.ms_grid_kernel / max(1.0, s1g.n_grid_launches)
        dk = s1d.ms_degrid_kernel / max(1.0, s1d.n_degrid_launches)
        modes[prec] = {"value": (taps_g + taps_d) / el, "ms_per_step": 1e3 * el / args.steps,
                       "first_pass_s": first, "cf_generation_s": cf_s, "grid_kernel_ms": gk, "degrid_kernel_ms": dk,
                       "taps_per_step": (taps_g + taps_d) / args.steps, "support": s1g.support,
                       "_grid_taps": taps_g / args.steps}
        chunk.vis_out.zero_()
        dg.degrid(chunk)
        vis_out[prec] = chunk.vis_out.cpu().numpy()
        images[prec] = gg.finalise_grid()
        log(f"{prec}: {modes[prec]['value'] / 1e12:.3f} Ttaps/s, {modes[prec]['ms_per_step']:.1f} ms/step")
    ref_i, ref_v = images["fp32"], vis_out["fp32"]
    for prec in ("fp16", "bf16"):
        q = slice(wl.nx // 4, 3 * wl.nx // 4)
        modes[prec]["deviation_vs_fp32_gpu"] = {
            "image_central_half": float(np.max(np.abs(images[prec][..., q, q] - ref_i[..., q, q])) /
                                        np.max(np.abs(ref_i))),
            "degridded_vis": float(np.max(np.abs(vis_out[prec] - ref_v)) / np.max(np.abs(ref_v)))}
    m = modes["fp16"]
    kname, kms = ("k_grid<1, 2, true, 1>", m["grid_kernel_ms"])
    achieved = 8.0 * m["_grid_taps"] / (kms * 1e-3) / 1e12
    for v in modes.values():
        v.pop("_grid_taps")
    result = {
        "metric": "CF-tap updates/s (AW-proj grid+degrid, FP16 CF and visibilities)",
        "value": m["value"], "unit": "CF-tap updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 storage, f32 arithmetic", "data": "synthetic",
        "config": {"workload": "awproject_feeds: 160000 rows x 32 ch 1.384-1.416 GHz, uv uniform in a 2 km disc, "
                               "9 feeds on a 1 deg grid, 12 m disk, 2048^2, os 8",
                   "rows": wl.nrow, "chans": wl.nchan, "feeds": wl.params["maxfeeds"], "grid": wl.nx,
                   "parallelism": "1 GPU"},
        "modes": modes,
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = aw_cpu_baseline(wl)
    print(json.dumps(result), flush=True)


def bench_imaging(args):
    """The imaging equation's data loop (ygrid_imaging_chunk, ImageFFTEquation.cc:437-480) on the
    W-projection array workload: per step one model degridder predicts, the residual is formed on
    the device, and a residual and a PSF gridder grid it (one call per step)."""
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.wproject_array(nint=args.nint)
    dev = torch.device("cuda:0")
    t_uvw = torch.from_numpy(np.ascontiguousarray(wl.uvw)).to(dev)
    t_freq = torch.from_numpy(np.ascontiguousarray(wl.freq)).to(dev)
    t_vis = torch.from_numpy(np.ascontiguousarray(wl.vis).view(np.float32)).to(dev)
    t_noise = torch.from_numpy(np.ascontiguousarray(wl.noise).view(np.float32)).to(dev)
    torch.cuda.synchronize()
    chunk = ya.Chunk(uvw=t_uvw, freq=t_freq, vis=t_vis, noise=t_noise, device=True)
    chunk.npol = wl.npol
    ctx = ya.Context(0)
    img = ya.ImageDesc(wl.nx, wl.ny, wl.npol, 1, wl.cell, wl.cell)
    prm = dict(wl.params)
    dg, res, psf = (ya.Gridder(ctx, "GPUWProject", prm) for _ in range(3))
    rng = np.random.default_rng(3)
    model = np.zeros((1, wl.npol, wl.ny, wl.nx), np.float32)
    for _ in range(50):
        model[0, :, rng.integers(500, 3500), rng.integers(500, 3500)] = rng.normal(size=wl.npol)
    t_model = torch.from_numpy(model).to(dev)
    torch.cuda.synchronize()
    dg.initialise_degrid(img, t_model)

    def step():
        ctx.clear_cache()  # nothing carries over between steps: each step bins once
        res.initialise_grid(img)
        psf.initialise_grid(img, dopsf=True)
        ya.imaging_chunk([dg], [res, psf], chunk)

    t0 = time.perf_counter()
    for i in range(args.warmup):
        step()
        log(f"warmup {i} done ({time.perf_counter() - t0:.1f} s)")
    ctx.synchronize()
    s0 = [g.stats() for g in (dg, res, psf)]
    for g in (dg, res, psf):
        g.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1 = [g.stats() for g in (dg, res, psf)]
    taps_d = s1[0].taps_degridded - s0[0].taps_degridded
    taps_g = sum(s1[k].taps_gridded - s0[k].taps_gridded for k in (1, 2))
    dk_ms = s1[0].ms_degrid_kernel / max(1.0, s1[0].n_degrid_launches)
    gk_ms = sum(s1[k].ms_grid_kernel for k in (1, 2)) / max(1.0, sum(s1[k].n_grid_launches for k in (1, 2)))
    tf = lambda b: "true" if b else "false"
    if dk_ms >= gk_ms:
        kname, kms, kt = f"k_degrid_mm<{wl.npol}, 0>" if s1[0].degrid_matrix else \
            f"k_degrid<{wl.npol}, {s1[0].degrid_waves}, {tf(s1[0].degrid_pipeline)}, 0>", dk_ms, taps_d / args.steps
    else:
        kname, kms, kt = f"k_grid<{wl.npol}, {s1[1].grid_waves}, {tf(s1[1].grid_pipeline)}, 0>", gk_ms, \
            taps_g / (2 * args.steps)
    achieved = 8.0 * kt / (kms * 1e-3) / 1e12
    other_ms = lambda f: sum(getattr(s1[k], f) - getattr(s0[k], f) for k in range(3)) / args.steps
    print(json.dumps({
        "metric": "CF-tap updates/s (imaging loop: degrid model, residual, grid residual + PSF)",
        "value": (taps_d + taps_g) / elapsed, "unit": "CF-tap updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "imaging_loop: wproject_array data (189000 rows x 64 ch x 4 pol, 4096^2, 129 "
                               "w-planes, S = 64), 1 model degridder + residual and PSF gridders, one "
                               "ygrid_imaging_chunk call per step", "parallelism": "1 GPU"},
        "phases_ms_per_step": {"degrid_kernel": dk_ms, "grid_kernel_each": gk_ms,
                               "prep": other_ms("ms_prep"), "binning": other_ms("ms_bin")},
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None},
    }), flush=True)


def bench_wstack(args):
    """W stacking (GPUWStack) on the W-projection array data: 65 w planes of 4096^2 x 4 pol grids
    (35 GB per gridder), SphFunc support 3, oversample 128; per step one grid and one degrid pass
    (the degridding reuses the gridding's binning); the per-plane transforms of initialise_degrid
    and finalise_grid are outside the timed steps."""
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.wproject_array(nint=args.nint)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    t_vis = t(np.ascontiguousarray(wl.vis).view(np.float32))
    chunk = ya.Chunk(uvw=t(wl.uvw), freq=t(wl.freq), vis=t_vis, vis_out=torch.zeros_like(t_vis),
                     noise=t(np.ascontiguousarray(wl.noise).view(np.float32)), device=True)
    chunk.npol = wl.npol
    ctx = ya.Context(0)
    img = ya.ImageDesc(wl.nx, wl.ny, wl.npol, 1, wl.cell, wl.cell)
    prm = {"wmax": wl.params["wmax"], "nwplanes": 65}
    gg, dg = ya.Gridder(ctx, "GPUWStack", prm), ya.Gridder(ctx, "GPUWStack", prm)
    rng = np.random.default_rng(3)
    model = np.zeros((1, wl.npol, wl.ny, wl.nx), np.float32)
    for _ in range(50):
        model[0, :, rng.integers(500, 3500), rng.integers(500, 3500)] = rng.normal(size=wl.npol)
    dg.initialise_degrid(img, t(model))
    torch.cuda.synchronize()

    def step():
        ctx.clear_cache()
        gg.initialise_grid(img)
        gg.grid(chunk)
        dg.degrid(chunk)

    for i in range(args.warmup):
        step()
    ctx.synchronize()
    s0g, s0d = gg.stats(), dg.stats()
    gg.set_profiling(True)
    dg.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1g, s1d = gg.stats(), dg.stats()
    taps_g, taps_d = s1g.taps_gridded - s0g.taps_gridded, s1d.taps_degridded - s0d.taps_degridded
    gk = s1g.ms_grid_kernel / max(1.0, s1g.n_grid_launches)
    dk = s1d.ms_degrid_kernel / max(1.0, s1d.n_degrid_launches)
    tf = lambda b: "true" if b else "false"
    if gk >= dk:
        kname, kms, kt = f"k_grid<{wl.npol}, {s1g.grid_waves}, {tf(s1g.grid_pipeline)}, 0>", gk, taps_g / args.steps
    else:
        kname = (f"k_degrid_mm<{wl.npol}, 0>" if s1d.degrid_matrix else
                 f"k_degrid<{wl.npol}, {s1d.degrid_waves}, {tf(s1d.degrid_pipeline)}, 0>")
        kms, kt = dk, taps_d / args.steps
    achieved = 8.0 * kt / (kms * 1e-3) / 1e12
    print(json.dumps({
        "metric": "CF-tap updates/s (W-stacking grid+degrid)", "value": (taps_g + taps_d) / elapsed,
        "unit": "CF-tap updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "wstack_array: wproject_array data (189000 rows x 64 ch x 4 pol, 4096^2), "
                               "65 w planes, SphFunc support 3, oversample 128", "parallelism": "1 GPU"},
        "phases_ms_per_step": {"grid_kernel": gk, "degrid_kernel": dk,
                               "prep": (s1g.ms_prep + s1d.ms_prep) / args.steps,
                               "binning": (s1g.ms_bin + s1d.ms_bin) / args.steps},
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None},
    }), flush=True)


def bench_sphfunc(args):
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.sphfunc_disc()
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = torch.zeros(wl.vis.size * 2, dtype=torch.float32, device=dev)
    chunk = ya.Chunk(uvw=t(wl.uvw), freq=t(wl.freq), vis=t(wl.vis.view(np.float32)), vis_out=out,
                     noise=t(wl.noise.view(np.float32)), device=True)
    chunk.npol = 1
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell)
    ctx = ya.Context(0)
    model = np.zeros((1, 1, wl.ny, wl.nx), np.float32)
    model[0, 0, wl.ny // 2, wl.nx // 2] = 1.0
    gg = ya.Gridder(ctx, "GPUSphFunc", wl.params)
    dg = ya.Gridder(ctx, "GPUSphFunc", wl.params)
    dg.initialise_degrid(img, t(model))
    # the small-support gridding kernel's LDS band height, tuned on this chunk (kept by the context)
    band_rows, band_ms = gg.autotune_small(img, chunk, trials=3)

    def step():
        ctx.clear_cache()
        gg.initialise_grid(img)
        gg.grid(chunk)
        dg.degrid(chunk)

    for _ in range(max(1, args.warmup)):
        step()
    ctx.synchronize()
    s0g, s0d = gg.stats(), dg.stats()
    gg.set_profiling(True)
    dg.set_profiling(True)
    nsteps = max(args.steps, 20)  # a step is ~1 ms: time enough of them
    t0 = time.perf_counter()
    for _ in range(nsteps):
        step()
    ctx.synchronize()
    el = time.perf_counter() - t0
    s1g, s1d = gg.stats(), dg.stats()
    taps = (s1g.taps_gridded - s0g.taps_gridded) + (s1d.taps_degridded - s0d.taps_degridded)
    gk = s1g.ms_grid_kernel / max(1.0, s1g.n_grid_launches)
    dk = s1d.ms_degrid_kernel / max(1.0, s1d.n_degrid_launches)
    # compulsory bytes of one gridding pass: uvw 24 B + vis 8 B + noise 8 B per sample, the CF
    # table once (128^2 planes x 7 x 7 taps x 8 B), the grid written once (2048^2 x 8 B)
    nbytes = wl.nrow * (24 + 8 + 8) + 128 * 128 * 49 * 8 + wl.nx * wl.ny * 8
    achieved = nbytes / (gk * 1e-3) / 1e9
    small = s1g.small_support_passes > 0
    lanes = 8 if s1g.support <= 3 else (16 if s1g.support <= 7 else 32)
    gname = f"k_grid_small<1, 0, {lanes}>" if small else "k_grid<1, 2, true, 0>"
    traffic = profiled_traffic(gname, PROFILE_SPH)
    result = {
        "metric": "CF-tap updates/s (SphFunc grid+degrid)", "value": taps / el, "unit": "CF-tap updates/s",
        "n_gpus": 1, "steps": nsteps, "warmup": args.warmup, "ms_per_step": 1e3 * el / nsteps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "sphfunc_disc: 160000 vis, uv uniform in a 2 km disc, 1.4 GHz, w = 0, 2048^2, "
                               "support 3, oversample 128", "rows": wl.nrow, "grid": wl.nx, "parallelism": "1 GPU"},
        "phases_ms_per_pass": {"grid_kernel": gk, "degrid_kernel": dk,
                               "prep": s1g.ms_prep / max(1.0, s1g.n_grid_launches),
                               "binning": s1g.ms_bin / max(1.0, s1g.n_grid_launches)},
        "small_support_kernels": small,
        "small_band_rows": {"chosen": band_rows, "tuning_pass_ms": band_ms},
        "roofline": {"bound": "hbm", "kernel": gname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": PROFILE_SPH,
                     "measured_traffic_gbs": None if traffic is None else traffic / (gk * 1e-3) / 1e9,
                     "measured_traffic_frac": None if traffic is None else traffic / (gk * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    print(json.dumps(result), flush=True)


def bench_snapshot(args):
    """Snapshot imaging (gridder.snapshotimaging) around the SphFunc gridder on a 4096^2 x 4 pol
    image: per step three coplanar snapshots of 20,000 rows (w = A u + B v with three planes), so
    three w-plane fits and, at the plane changes and the end, the wrapped gridder's finalisation
    and the regrid of its image and weights into the target frame.  The dominant new kernel is
    k_snapshot_regrid: its compulsory bytes per launch are the source image, the output read and
    the output written (3 x 4 pol x 4096^2 x 4 B)."""
    import yandasoft_amd as ya

    N, cell = 4096, math.radians(5 / 3600)
    rng = np.random.default_rng(20261204)
    F = np.array([1.4e9])
    chunks = []
    for A, B in ((1.5, 0.5), (-1.0, 1.2), (0.3, -2.0)):
        u, v = rng.normal(size=(2, 20000)) * 800
        uvw = np.ascontiguousarray(np.column_stack([u, v, A * u + B * v]))
        vis = np.ascontiguousarray((rng.normal(size=(4, 1, 20000)) + 1j * rng.normal(size=(4, 1, 20000))).astype(np.complex64))
        chunks.append(ya.Chunk(uvw=uvw, freq=F, vis=vis, noise=np.ones_like(vis)))
    ctx = ya.Context(0)
    img = ya.ImageDesc(N, N, 4, 1, cell, cell)
    g = ya.Gridder(ctx, "GPUSphFunc", {})
    g.configure({"snapshotimaging": True, "snapshotimaging.wtolerance": 1.0})

    def step():
        g.initialise_grid(img)
        for ch in chunks:
            g.grid(ch)
        g.finalise_grid()
        g.finalise_weights()

    for _ in range(max(1, args.warmup)):
        step()
    s0 = g.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    el = time.perf_counter() - t0
    s1 = g.stats()
    result = {
        "metric": "snapshot imaging steps/s (3 snapshots, fits, finalisations and regrids)",
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 images, f64 coordinates", "data": "synthetic",
        "config": {"workload": "snapshot_array: SphFunc, 4096^2 x 4 pol, 3 coplanar snapshots of 20000 rows",
                   "parallelism": "1 GPU"},
        "fits_per_step": (s1.snapshot_fits - s0.snapshot_fits) / args.steps,
        "regrids_per_step": (s1.snapshot_regrids - s0.snapshot_regrids) / args.steps,
        "regrid_compulsory_bytes": 3 * 4 * N * N * 4,
    }
    print(json.dumps(result), flush=True)


def bench_learned(args):
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wl = W.learned_wdisc()
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = torch.zeros(wl.vis.size * 2, dtype=torch.float32, device=dev)
    chunk = ya.Chunk(uvw=t(wl.uvw), freq=t(wl.freq), vis=t(wl.vis.view(np.float32)), vis_out=out,
                     noise=t(wl.noise.view(np.float32)), device=True)
    chunk.npol = 1
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell)
    ctx = ya.Context(0)
    model = np.zeros((1, 1, wl.ny, wl.nx), np.float32)
    rng = np.random.default_rng(9)
    for _ in range(20):
        model[0, 0, rng.integers(wl.nx // 4, 3 * wl.nx // 4), rng.integers(wl.nx // 4, 3 * wl.nx // 4)] = 1.0
    t_model = t(model)
    tg = ya.Gridder(ctx, "GPUWProject", wl.params)
    tg.initialise_grid(img)
    tg.grid(chunk)  # generates the 129-plane table
    lcf = ya.LearnedCF(ctx, hidden=128, layers=4, seed=2)
    rep = lcf.train_wproject(tg, epochs=args.train_epochs, batch=1024, lr=3e-3, seed=3, holdout_every=8)
    log(f"trained {rep.steps} steps in {rep.train_s:.1f} s: support {rep.support}, held-out max err "
        f"{rep.holdout_max_err:.2e} rms {rep.holdout_rms_err:.2e}")
    modes = {}
    grids = {}
    # the 1441-plane table: closer to continuous w than the network (5.1e-4 against 1.1e-3 of the
    # 1025-plane FP64 oracle's peak, test_gpu_learned), so it bounds the network's matched-error time
    for name, use_lcf, prec, nw in (("table_fp32", False, "fp32", None), ("learned_fp16", True, "fp16", None),
                                    ("table_1441_planes_fp32", False, "fp32", 1441)):
        params = dict(wl.params, cfprecision=prec, visprecision=prec)
        if nw is not None:
            params["nwplanes"] = nw
        gg = ya.Gridder(ctx, "GPUWProject", params)
        dg = ya.Gridder(ctx, "GPUWProject", params)
        if use_lcf:
            gg.set_learned_cf(lcf, 1.0)
            dg.set_learned_cf(lcf, 1.0)
        dg.initialise_degrid(img, t_model)

        def step():
            # each step starts cold: the binning and the network's planes are built again; the
            # degridding pass then reuses the gridding pass's binning and network evaluation
            ctx.clear_cache()
            lcf.clear_cache()
            gg.initialise_grid(img)
            gg.grid(chunk)
            dg.degrid(chunk)

        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        s0g, s0d = gg.stats(), dg.stats()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        el = time.perf_counter() - t0
        s1g, s1d = gg.stats(), dg.stats()
        taps = (s1g.taps_gridded - s0g.taps_gridded) + (s1d.taps_degridded - s0d.taps_degridded)
        modes[name] = {"value": taps / el, "ms_per_step": 1e3 * el / args.steps, "support": s1g.support,
                       "learned_cf_active": s1g.learned_cf_active}
        grids[name] = gg.grid_array()
        log(f"{name}: {taps / el / 1e9:.2f} Gtaps/s, {1e3 * el / args.steps:.1f} ms/step")
    dev_ = float(np.max(np.abs(grids["learned_fp16"] - grids["table_fp32"])) / np.max(np.abs(grids["table_fp32"])))
    m = modes["learned_fp16"]
    # the network runs once per (sample, tap) and step: layers [5, 128 x 4, 2]; the degridding
    # pass copies the gridding pass's planes (same samples)
    dims = [5, 128, 128, 128, 128, 2]
    net_macs = sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    W = 2 * rep.support + 1
    evals_per_step = wl.nrow * wl.nchan * W * W
    net_tflops = evals_per_step * 2 * net_macs / (m["ms_per_step"] * 1e-3) / 1e12
    result = {
        "metric": "CF-tap updates/s (learned-CF W-proj grid+degrid, continuous w)",
        "value": m["value"], "unit": "CF-tap updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 weights and CF storage, f32 arithmetic", "data": "synthetic",
        "config": {"workload": "learned_wdisc: 160000 vis, uv uniform in a 2 km disc, w uniform in [-300, 300] m, "
                               "1.4 GHz, 2048^2, network 4 x 128 SiLU trained on a 129-plane W-proj table (os 4)",
                   "rows": wl.nrow, "grid": wl.nx, "parallelism": "1 GPU"},
        "training": {"epochs": rep.epochs, "steps": rep.steps, "examples": rep.examples, "cfs": rep.cfs,
                     "holdout_cfs": rep.holdout_cfs, "seconds": rep.train_s, "first_loss": rep.first_loss,
                     "last_loss": rep.last_loss, "holdout_max_err_vs_table": rep.holdout_max_err,
                     "holdout_rms_err_vs_table": rep.holdout_rms_err, "support": rep.support},
        "modes": modes,
        "grid_deviation_learned_vs_129_plane_table": dev_,
        "matched_error": {"learned_ms_per_step": m["ms_per_step"],
                          "table_1441_planes_ms_per_step": modes["table_1441_planes_fp32"]["ms_per_step"],
                          "learned_over_table": m["ms_per_step"] / modes["table_1441_planes_fp32"]["ms_per_step"],
                          "errors_vs_1025_plane_fp64_oracle": "learned 1.10e-3, 1441-plane table 5.1e-4, 129-plane "
                                                              "table 2.71e-3 (test_gpu_learned)"},
        # the network (hidden width 128) runs in k_lcf_fused<true> on the FP16 matrix cores
        "roofline": {"bound": "mfma", "kernel": "k_lcf_fused<true> (every network layer, FP16 MFMA 32x32x16, "
                                               "FP32 sums), over the whole pass time",
                     "achieved": net_tflops, "peak": FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": net_tflops / FP16_PEAK_TFLOPS, "traffic": None,
                     "flop_per_tap_evaluation": 2 * net_macs},
    }
    print(json.dumps(result), flush=True)


def bench_batched(args):
    import torch
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W

    wls = W.batched_beams()
    nsamp = sum(w.nrow * w.nchan for w in wls)
    log(f"{len(wls)} beams, {nsamp} samples x 4 pols, grids {sorted(set(w.nx for w in wls))}")
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ctx = ya.Context(0)
    batch = ya.Batch(ctx)
    grid_g, imgs, keep = [], [], []
    rng = np.random.default_rng(5)
    t0 = time.perf_counter()
    for w in wls:
        img = ya.ImageDesc(w.nx, w.ny, w.npol, 1, w.cell, w.cell)
        ch = ya.Chunk(uvw=t(w.uvw), freq=t(w.freq), vis=t(w.vis.view(np.float32)), noise=t(w.noise.view(np.float32)),
                      device=True)
        ch.npol = w.npol
        out = torch.zeros(w.vis.size * 2, dtype=torch.float32, device=dev)
        chd = ya.Chunk(uvw=ch.uvw, freq=ch.freq, vis=ch.vis, vis_out=out, noise=ch.noise, device=True)
        chd.npol = w.npol
        model = np.zeros((1, w.npol, w.ny, w.nx), np.float32)
        for _ in range(10):
            model[0, :, rng.integers(w.nx // 4, 3 * w.nx // 4), rng.integers(w.nx // 4, 3 * w.nx // 4)] = 1.0
        g = ya.Gridder(ctx, "GPUWProject", w.params)
        d = ya.Gridder(ctx, "GPUWProject", w.params)
        g.initialise_grid(img)
        d.initialise_degrid(img, t(model))
        batch.add(g, ch)
        batch.add(d, chd)
        grid_g.append(g)
        imgs.append(img)
        keep += [g, d, ch, chd, out]
    torch.cuda.synchronize()
    batch.run()  # CF generation, buffers, first run
    setup_s = time.perf_counter() - t0
    log(f"set-up + first run {setup_s:.1f} s")
    tune = batch.autotune(1)
    log(f"autotuned: waves {tune.grid_waves} pipeline {tune.grid_pipeline} item pairs {tune.item_pairs} "
        f"({tune.ms:.1f} ms, {tune.trials} configurations)")
    batch.set_graph(True)

    def step():
        for g, img in zip(grid_g, imgs):
            g.initialise_grid(img)
        batch.run()

    gridders = keep[0::5] + keep[1::5]
    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    s0 = [(x.stats().taps_gridded, x.stats().taps_degridded) for x in gridders]
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i} done")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    s1 = [(x.stats().taps_gridded, x.stats().taps_degridded) for x in gridders]
    taps = sum((b[0] - a[0]) + (b[1] - a[1]) for a, b in zip(s0, s1))
    value = taps / el
    achieved = 8.0 * value / 1e12
    result = {
        "metric": "CF-tap updates/s (batched W-proj grid+degrid, 36 beams)",
        "value": value, "unit": "CF-tap updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched_beams: 36 W-proj beams, 8-64 chans, 1e4-2e6 vis per pol (log-uniform), "
                               "grids 1024^2/2048^2/4096^2, 129 w-planes, os 4, support <= 64, 4 pols",
                   "beams": len(wls), "samples": nsamp, "pols": 4, "parallelism": "1 GPU",
                   "graph": True, "launches_per_step": "per beam prep + binning, one grid and one degrid tile launch"},
        "autotune": {"grid_waves": tune.grid_waves, "grid_pipeline": tune.grid_pipeline,
                     "item_pairs": tune.item_pairs, "best_ms": tune.ms, "configurations": tune.trials},
        "setup_s": setup_s,
        "roofline": {"bound": "mfma", "kernel": "whole batch run (prep, binning and tile kernels)",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None},
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
