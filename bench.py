"""Benchmark of the column biogeophysics soil step on one MI355X.

A "step" is one time step (dtime = 1800 s) of the soil physics over every
column of the batch: SoilTemperature (snow + soil heat solve with phase
change), SetSoilWaterFractions, and the moisture-form Richards solve, in ONE
fused launch (the library's ctsm_soil_step_fused_dev).  The workload is the
uneven-column batch: 65,536 columns, snow layers uniform 0..12, bedrock index
uniform 5..20, one patch per column, FP64, synthetic seeded inputs resident
in HBM.  The same steps are also timed in the plain form (one launch per
reference routine, intermediates through HBM) for the fused/plain ratio.

Prints ONE JSON line (rank 0) with the metric, the roofline of the fused
kernel (algorithmic bytes per launch / measured launch time vs 8 TB/s) and a
CPU baseline (the C oracle, single thread, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(st):
    """Compulsory HBM bytes of one fused soil step, summed over the columns.

    Every array element the step must read is counted once and every element
    it must write once (FP64 = 8 B, int32 = 4 B), over active extents only:
      per column   : 14 FP64 + 4 int32 scalars read, 18 FP64 scalars written;
      snow layer   : 6 FP64 read (t, liq, ice, dz, z, zi), 7 FP64 + 1 int32 written
                     (t, liq, ice, thk, fact, bw + imelt) ;
      all nlevsno snow slots: qflx_snomelt_lyr, qflx_snofrz_lyr written (2 FP64);
      ground level : 12 FP64 read (t, liq, ice, dz, z, zi, watsat, tkmg, tkdry,
                     csol, bsw, sucsat) + zi(0), 6 FP64 + 1 int32 written
                     (t, liq, ice, thk, fact, eflx_fgr + imelt);
      nlevsoi level: eff_porosity, icefrac written (2 FP64);
      Richards layer (nbedrock): hksat, qflx_rootsoi read; smp_l, hk_l, qin,
                     qout written (2 + 4 FP64);
      patch        : 15 FP64 + 1 int32 read, sabg_lyr over snl+1..1 read,
                     3 FP64 written.
    """
    S = -st.col["snl"].astype(np.int64)
    L = st.col["nbedrock"].astype(np.int64)
    P = (st.col["patchf"] - st.col["patchi"] + 1).astype(np.int64)
    NS, NSOI, NG = st.dims.nlevsno, st.dims.nlevsoi, st.dims.nlevgrnd
    per_col = (14 * 8 + 4 * 4 + 18 * 8
               + S * (6 * 8 + 7 * 8 + 4)
               + NS * 2 * 8
               + NG * (12 * 8 + 6 * 8 + 4) + 8
               + NSOI * 2 * 8
               + L * 6 * 8
               + P * (15 * 8 + 4 + 3 * 8) + P * (S + 1) * 8)
    return int(per_col.sum())


def cpu_baseline(st, seconds=15.0):
    """Time the C oracle (single thread) on a bounded slice of the same workload."""
    from oracle import oracle as O
    ncol = st.dims.ncol
    sample = min(ncol, 4096)
    sub = st.subset(np.arange(sample))
    t0 = time.perf_counter()
    steps = 0
    while True:
        rc, fail = O.soil_step(sub)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 48:
            break
    return sample * steps / el, f"{sample} columns x {steps} steps of the same workload ({el:.1f} s)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncol", type=int, default=65536)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-sort", action="store_true", help="keep generation order (no binning)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--lanes", type=int, default=0, help="force lanes per column in the Richards solve")
    ap.add_argument("--single-kernel", action="store_true", help="force the one-kernel fused step")
    ap.add_argument("--order", choices=["bins", "substeps"], default="substeps",
                    help="column order: (bedrock, snow) bins, or also by the sub-steps of the warm-up step")
    ap.add_argument("--breakdown", action="store_true",
                    help="also time SoilTemperature and SoilWater launched separately")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run a streaming copy kernel of this size (MB) after the steps, "
                         "to calibrate the HBM byte counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(columns shard trivially by range, see DESIGN.md)"}))
        sys.exit(2)

    from ctsm_amd import lib, synthetic
    from ctsm_amd.state import DeviceState

    st = synthetic.workload("uneven", ncol=args.ncol)
    ctx = lib.Context(device=0)

    from ctsm_amd import autotune as AT

    filt = None
    if not args.no_sort:
        filt = ctx.device_array(AT.column_order(st.col["snl"], st.col["nbedrock"]))

    tuning, table = None, []
    if args.lanes or args.single_kernel:
        lanes = args.lanes or 1
        ctx.set_tuning(richards_lanes=lanes, richards_block=64 if lanes == 1 else 128,
                       fused_single_kernel=int(args.single_kernel))
    elif not args.no_autotune:
        ds0 = DeviceState(st, ctx)
        tuning, _, table = AT.autotune(ctx, ds0, filter_dev=filt, steps=2)
        ds0.free()

    def run(plain):
        nonlocal filt
        ds = DeviceState(st, ctx)
        structs = ds.structs()
        for _ in range(args.warmup):
            ctx.step_dev(ds, filter_dev=filt, plain=plain, structs=structs)
        ctx.sync()
        if args.order == "substeps" and not args.no_sort:
            # order columns by the sub-steps they took in the last warm-up step
            nsub = np.empty(st.dims.ncol)
            ctx.d2h(nsub, ds.col["num_substeps"])
            ctx.h2d(filt, AT.column_order(st.col["snl"], st.col["nbedrock"], nsub))
        # events on the library's own stream, around each step's launch(es)
        starts, stops = [], []
        t0 = time.perf_counter()
        for k in range(args.steps):
            starts.append(ctx.record())
            ctx.step_dev(ds, filter_dev=filt, plain=plain, structs=structs)
            stops.append(ctx.record())
        ctx.sync()
        wall = time.perf_counter() - t0
        per = [ctx.elapsed_ms(a, b) for a, b in zip(starts, stops)]
        for e in starts + stops:
            ctx.destroy_event(e)
        return wall, float(np.mean(per)), ds

    wall, step_ms, ds = run(plain=False)
    ds.free()
    value = args.ncol * args.steps / wall
    res = {"fused_step_ms": step_ms, "tuning": ctx.get_tuning()}
    if tuning is not None:
        res["autotune_ms_per_step"] = [[c, round(ms, 4)] for c, ms in table]
    if not args.no_plain:
        wall_p, step_ms_p, dsp = run(plain=True)
        dsp.free()
        res["plain_step_ms"] = step_ms_p
        res["plain_column_steps_per_s"] = args.ncol * args.steps / wall_p
        res["fused_over_plain"] = value / res["plain_column_steps_per_s"]

    if args.breakdown:
        ds = DeviceState(st, ctx)
        for name, fn in (("soiltemperature_fused_kernel_ms", lambda: ctx.soiltemperature_dev(ds, filt)),
                         ("soilwater_fused_kernel_ms", lambda: ctx.soilwater_dev(ds, filt))):
            fn()
            ctx.sync()
            e0 = ctx.record()
            for _ in range(3):
                fn()
            e1 = ctx.record()
            ctx.sync()
            res[name] = ctx.elapsed_ms(e0, e1) / 3
            ctx.destroy_event(e0)
            ctx.destroy_event(e1)
        ds.free()
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy(b, a, nb)
        e0 = ctx.record()
        ctx.copy(b, a, nb)
        e1 = ctx.record()
        ctx.sync()
        ms = ctx.elapsed_ms(e0, e1)
        res["copy_calib"] = {"bytes_read": nb, "bytes_written": nb, "ms": ms, "GBps": 2 * nb / ms / 1e6}
        ctx.free(a)
        ctx.free(b)

    bytes_alg = algorithmic_bytes(st)
    achieved = bytes_alg / (step_ms * 1e-3) / 1e9
    out = {
        "metric": "column-steps/sec (FP64, 1 GPU)",
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": "uneven columns: 65,536 soil columns, snow 0-12 layers, bedrock 5-20, "
                               "SoilTemperature + SoilWater, dtime 1800 s, 48-step day",
                   "ncol": args.ncol, "binned": not args.no_sort},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_soil_step", "alg_bytes_per_launch": bytes_alg},
        "detail": res,
    }
    if not args.no_cpu:
        v, sample = cpu_baseline(st, args.cpu_seconds)
        out["cpu_baseline"] = {"value": v, "unit": "column-steps/s", "cores": 1, "kind": "port",
                               "sample": sample}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
