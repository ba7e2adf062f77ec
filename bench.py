"""Benchmark of the column biogeophysics step on one MI355X.

A "step" is one time step (dtime = 1800 s) of the soil physics over every
column of the batch: SoilTemperature (snow + soil heat solve with phase
change) with SetSoilWaterFractions, then the moisture-form Richards solve
(the library's fused soil step; launch shape -- lanes per column, blocks,
streams the columns are split across -- chosen by the autotuner).  The timed
steps are captured once as one HIP graph and launched once.  The headline
workload is the uneven-column batch: 65,536 columns, snow layers uniform
0..12, bedrock index uniform 5..20, one patch per column, FP64, synthetic
seeded inputs resident in HBM.  The same steps are also timed in the plain
form (one launch per reference routine, intermediates through HBM) for the
fused/plain ratio.

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(algorithmic bytes per launch / its measured launch time vs 8 TB/s), a CPU
baseline (the C oracle on a bounded sample, one column chunk per thread) and,
in "detail", the full-column step with CanopyFluxes (262,144 columns, 1-16
patches each).
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

# HBM bytes per launch of the full-batch kernel from rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE (KiB per dispatch), each scaled by the same
# counter's reading on a streaming copy of 1 GiB (FETCH x 2.0, WRITE x 1.0):
# k_soilwater_g<8> on the 65,536-column batch reads 151,589 KiB x 2 and
# writes 276,934 KiB (profiles/richards_heat_solve_uneven65536.md).
TRAFFIC = {"k_soilwater_g<8>": (151589.2 * 2.0 + 276934.4) * 1024.0}
# FP64 VALU utilisation of the same launch from rocprofv3 --pmc SQ_INSTS_VALU_*_F64
# (same profile): 64 lanes x (ADD + MUL + 2 FMA + TRANS) instructions per launch
# over the launch time, against the 78.6 TFLOP/s FP64 vector peak.
FP64_VALU_FRAC = {"k_soilwater_g<8>": 64 * (62.838e6 + 65.686e6 + 2 * 92.273e6 + 15.013e6) / 1.6863e-3 / 78.6e12}


def richards_bytes(st):
    """Compulsory HBM bytes of one Richards launch (k_soilwater*), summed over
    the columns: per Richards layer (nbedrock of them) 10 FP64 read
    (h2osoi_liq, dz, z, watsat, bsw, sucsat, hksat, icefrac, eff_porosity,
    qflx_rootsoi) and 5 written (h2osoi_liq, smp_l, hk_l, qin, qout); per column
    nbedrock + qflx_infl read and qcharge + num_substeps written:
    B_W = 120 L + 28 bytes."""
    L = st.col["nbedrock"].astype(np.int64)
    return int((120 * L + 28).sum())


def soiltemp_bytes(st):
    """Compulsory HBM bytes of one heat-solve launch (SoilTemperature), summed
    over columns:
      per column   : 14 FP64 + 4 int32 scalars read, 18 FP64 scalars written;
      snow layer   : 6 FP64 read (t, liq, ice, dz, z, zi), 7 FP64 + 1 int32
                     written (t, liq, ice, thk, fact, bw + imelt);
      all nlevsno snow slots: qflx_snomelt_lyr, qflx_snofrz_lyr written (2 FP64);
      ground level : 12 FP64 read (t, liq, ice, dz, z, zi, watsat, tkmg, tkdry,
                     csol, bsw, sucsat) + zi(0), 6 FP64 + 1 int32 written
                     (t, liq, ice, thk, fact, eflx_fgr + imelt);
      patch        : 15 FP64 + 1 int32 read, sabg_lyr over snl+1..1 read,
                     3 FP64 written.
    """
    S = -st.col["snl"].astype(np.int64)
    P = (st.col["patchf"] - st.col["patchi"] + 1).astype(np.int64)
    NS, NG = st.dims.nlevsno, st.dims.nlevgrnd
    per_col = (14 * 8 + 4 * 4 + 18 * 8
               + S * (6 * 8 + 7 * 8 + 4)
               + NS * 2 * 8
               + NG * (12 * 8 + 6 * 8 + 4) + 8
               + P * (15 * 8 + 4 + 3 * 8) + P * (S + 1) * 8)
    return int(per_col.sum())


_CANOPY_INOUT = {"displa", "z0mv", "t_veg", "t_stem", "snocan", "liqcan", "cgrnds", "cgrndl"}


def _canopy_const_fields():
    import re
    with open(os.path.join(ROOT, "include", "ctsm_gpu.h")) as f:
        s = f.read()
    body = s[s.index("typedef struct ctsm_canopy_state {"):s.index("} ctsm_canopy_state;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = set()
    for line in body.splitlines():
        line = line.strip()
        if line.startswith("const"):
            for nm in line.split(None, 2)[2].rstrip(";").split(","):
                names.add(nm.strip().lstrip("*").strip())
    return names


def canopy_bytes(st):
    """Compulsory HBM bytes of one CanopyFluxes launch (k_canopy): every patch
    field of ctsm_canopy_state counted once if read and once if written
    (inout both), FP64 8 B / int32 4 B, the three nlevgrnd patch profiles
    (rootfr read, rresis and rootr written) included; per column its ground
    state (23 FP64 + snl, 7 nlevgrnd profiles read, eff_porosity and
    h2osoi_liqvol written) and per gridcell the forcing (9 FP64 + 1 int32)."""
    from ctsm_amd import abi
    const = _canopy_const_fields()
    ins = outs = 0
    for name, t, kind in abi.CANOPY_FIELDS:
        if kind not in ("p", "soip"):
            continue
        b = (4 if t == "i32" else 8) * (st.dims.nlevgrnd if kind == "soip" else 1)
        if name in const:
            ins += b
        elif name in _CANOPY_INOUT:
            ins += b
            outs += b
        else:
            outs += b
    NG = st.dims.nlevgrnd
    per_col = 23 * 8 + 4 + NG * 7 * 8 + NG * 2 * 8
    per_grid = 9 * 8 + 4
    return int(st.dims.npatch * (ins + outs) + st.dims.ncol * (per_col + per_grid))


def cpu_baseline(st, seconds=15.0, threads=None):
    """Time the C oracle on a bounded slice of the same workload, one column
    chunk per thread (the oracle runs outside the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    per = 1024
    ncol = st.dims.ncol
    # a seeded sample of the whole batch (its storage order is binned by work)
    pick = np.random.default_rng(12345).permutation(ncol)
    subs = [st.subset(np.sort(pick[(np.arange(per) + i * per) % ncol])) for i in range(threads)]

    def work(sub):
        n = 0
        t = time.perf_counter()
        while True:
            O.soil_step(sub)
            n += 1
            if time.perf_counter() - t >= seconds or n >= 48 * 40:
                return n

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        steps = list(ex.map(work, subs))
    el = time.perf_counter() - t0
    total = per * sum(steps)
    return total / el, threads, (f"{threads} threads x {per} columns (a seeded random sample of the batch) x "
                                 f"{min(steps)}-{max(steps)} steps ({el:.1f} s)")


def served_ensemble(ctx, args):
    """Perturbed-parameter members arriving as a Poisson stream, run by the
    continuous-batching server in one resident pool; latency = completion -
    arrival (wall clock)."""
    from ctsm_amd import serve, synthetic
    ctx.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
    reqs = serve.make_requests(args.serve_requests, args.serve_rate, seed=0xC75D0004,
                               day_scale=args.serve_day_scale)
    srv, done, wall = serve.serve(ctx, reqs, args.serve_pool_cols, args.serve_pool_cols * 17)
    lat = np.array([q.done - q.arrival for q in done])
    out = {"workload": f"{len(reqs)} members, Poisson {args.serve_rate}/s, 1-20,000 columns (log-uniform) with "
                       f"1-16 patches, 30 parameters x U(0.5,1.5), 1-365 days (log-uniform) x 48 steps x "
                       f"{args.serve_day_scale}; full land step (CanopyFluxes, SoilTemperature, SoilFluxes, "
                       f"surface water, SoilWater, water tables, drainage) and the soil balance check every step",
           "pool_columns": args.serve_pool_cols, "pool_bytes_per_column": srv.bytes_per_column(),
           "completed": len(done), "wall_s": wall,
           "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
           "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
           "column_steps_per_s": srv.column_steps / wall,
           "max_abs_errsoi_Wm2": srv.worst_errsoi}
    srv.close()
    return out


def time_events(ctx, fns, reps, prefill=None, inner=1):
    """Mean ms of each callable launched `reps` x `inner` times, events on the
    library stream.  For kernels shorter than the host's launch overhead,
    `prefill` enqueues long work first so that the GPU is still busy while the
    host enqueues the timed launches back to back: the events then time the
    kernels, not the host."""
    acc = [0.0] * len(fns)
    for _ in range(reps):
        if prefill is not None:
            prefill()
        evs = [ctx.record()]
        for fn in fns:
            for _ in range(inner):
                fn()
            evs.append(ctx.record())
        ctx.sync()
        for i in range(len(fns)):
            acc[i] += ctx.elapsed_ms(evs[i], evs[i + 1]) / inner
        for e in evs:
            ctx.destroy_event(e)
    return [a / reps for a in acc]


def full_column(ctx, args):
    """Full column step with CanopyFluxes (1-16 patches per column): land
    step throughput and the canopy kernel's own launch time."""
    from ctsm_amd import autotune as AT, synthetic
    from ctsm_amd.state import DeviceState
    st = synthetic.canopy_workload("canopy", ncol=args.canopy_ncol)
    ctx.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
    # warm-up steps, then store the batch in binned column order (as the soil bench)
    ds = DeviceState(st, ctx)
    for _ in range(max(args.warmup, 1)):
        ctx.land_step_dev(ds)
    ctx.sync()
    stw = ds.to_host()
    ds.free()
    order = AT.column_order(stw.col["snl"], stw.col["nbedrock"], stw.col["num_substeps"], parts=4) - 1
    st = stw.subset(order)
    ds = DeviceState(st, ctx)
    filt = None
    structs = (*ds.structs(), ds.canopy_struct())
    for _ in range(args.warmup):
        ctx.land_step_dev(ds, filterc_dev=filt, structs=structs)
    ctx.sync()
    split, _, split_table = AT.autotune_canopy(ctx, ds)
    e0 = ctx.record()
    t0 = time.perf_counter()
    for _ in range(args.canopy_steps):
        ctx.land_step_dev(ds, filterc_dev=filt, structs=structs)
    e1 = ctx.record()
    ctx.sync()
    wall = time.perf_counter() - t0
    step_ms = ctx.elapsed_ms(e0, e1) / args.canopy_steps
    ctx.destroy_event(e0)
    ctx.destroy_event(e1)
    can_ms, = time_events(ctx, [lambda: ctx.canopyfluxes_dev(ds, structs=structs)], 3)
    b = canopy_bytes(st)
    nit = np.empty(st.dims.npatch)
    ctx.d2h(nit, ds.can["num_iter"])
    out = {"workload": f"{st.dims.ncol:,} columns, {st.dims.npatch:,} patches (1-16 per column): CanopyFluxes, "
                       "transpiration sink, SoilTemperature, SoilWater; no snow, 20 layers",
           "column_steps_per_s": st.dims.ncol * args.canopy_steps / wall,
           "patch_steps_per_s": st.dims.npatch * args.canopy_steps / wall,
           "step_ms": step_ms, "steps": args.canopy_steps,
           "canopy_kernel_ms": can_ms,
           "mean_canopy_iterations": float(nit.mean()),
           "canopy_rounds": {"first_iterations": split[0], "later_iterations": split[1],
                             "candidates_ms": {f"{a}/{c}": round(ms, 3) for (a, c), ms in split_table}},
           "canopy_roofline": {"kernel": "k_canopy_round (all rounds of one call)", "alg_bytes_per_launch": b,
                               "achieved_GBps": b / (can_ms * 1e-3) / 1e9,
                               "frac": b / (can_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    out["with_soilfluxes_and_surface_water"] = _full_land_step(ctx, args, ds, structs)
    ds.free()
    out["phs_soil_side"] = _phs_soil_side(ctx, st.dims.ncol)
    if args.canopy_parts > 1:
        out["concurrent_parts"] = _full_column_parts(ctx, args, st)
    return out


def surface_water_bytes(st):
    """Algorithmic bytes of one surface water launch: per column 22 FP64 inputs
    (14 of the chain, frac_sno_eff, frac_h2osfc, icefrac and hksat of layers
    1-3), snl, and 16 FP64 outputs (h2osfc included)."""
    return st.dims.ncol * (22 * 8 + 4 + 16 * 8)


def _full_land_step(ctx, args, ds, structs):
    """The land step with the reference's bookkeeping (before-step save,
    SoilFluxes, the surface water input chain feeding SoilWater's
    infiltration, the water tables and RenewCondensation after it, the
    drainage of HydrologyDrainage) on the
    same resident batch, and the launch times of the SoilFluxes and surface
    water kernels."""
    from ctsm_amd import synthetic
    hydro = synthetic.hydro_params()
    full = (*structs, ds.soilflux_struct())
    for _ in range(2):
        ctx.land_step_full_dev(ds, structs=full, hydro=hydro, water_table=True, drainage=True)
    ctx.sync()
    e0 = ctx.record()
    t0 = time.perf_counter()
    for _ in range(args.canopy_steps):
        ctx.land_step_full_dev(ds, structs=full, hydro=hydro, water_table=True, drainage=True)
    e1 = ctx.record()
    ctx.sync()
    wall = time.perf_counter() - t0
    step_ms = ctx.elapsed_ms(e0, e1) / args.canopy_steps
    ctx.destroy_event(e0)
    ctx.destroy_event(e1)
    # the same steps captured once and launched as one graph
    g = ctx.graph_capture_full(ds, nsteps=args.canopy_steps, hydro=hydro, water_table=True, drainage=True)
    ctx.graph_launch(g)                      # first launch uploads the graph
    ctx.sync()
    e0 = ctx.record()
    t0 = time.perf_counter()
    ctx.graph_launch(g)
    e1 = ctx.record()
    ctx.sync()
    graph_wall = time.perf_counter() - t0
    graph_step_ms = ctx.elapsed_ms(e0, e1) / args.canopy_steps
    ctx.destroy_event(e0)
    ctx.destroy_event(e1)
    ctx.graph_destroy(g)
    worst, _ = ctx.soil_balance_check_dev(ds)
    err = np.empty(ds.dims.ncol)
    ctx.d2h(err, ds.sf["errsoi_col"])
    liq = np.empty((ds.dims.nlevsno + ds.dims.nlevgrnd, ds.dims.ncol))
    ctx.d2h(liq, ds.col["h2osoi_liq"])
    # (the water-table kernel is not re-run here: RenewCondensation applied
    # twice would take the step's sublimation twice)
    sw_ms, sf_ms = time_events(ctx, [lambda: ctx.surface_water_dev(ds, hydro), lambda: ctx.soilfluxes_dev(ds)], 3,
                               prefill=lambda: ctx.canopyfluxes_dev(ds, structs=structs), inner=10)
    b = surface_water_bytes(ds)
    out = {"history": _history_and_checkpoint(ctx, ds, structs)}
    return {**out, "column_steps_per_s": ds.dims.ncol * args.canopy_steps / wall, "step_ms": step_ms,
            "graph_column_steps_per_s": ds.dims.ncol * args.canopy_steps / graph_wall,
            "graph_step_ms": graph_step_ms,
            "max_abs_errsoi_Wm2": worst,
            "columns_errsoi_over_1e-5_Wm2": int(np.sum(~(np.abs(err) <= 1e-5))),
            "columns_negative_soil_liquid": int(np.sum((liq[ds.dims.nlevsno:] < 0).any(axis=0))),
            "surface_water_kernel_ms": sw_ms, "soilfluxes_kernels_ms": sf_ms,
            "surface_water_roofline": {"kernel": "k_surface_water", "alg_bytes_per_launch": b,
                                       "achieved_GBps": b / (sw_ms * 1e-3) / 1e9,
                                       "frac": b / (sw_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}


def _phs_soil_side(ctx, ncol):
    """The two plant-hydraulic-stress soil kernels (root-soil conductances per
    patch, the vertical sink per column) on a synthetic case of ncol columns
    with 0-4 patches each, device-resident, with their HBM rooflines."""
    from ctsm_amd import phs as PH
    A = PH.synthetic(ncol, max_patches=4, seed=3)
    ctx.set_phs_params(PH.default_params())
    D = PH.DevicePhs(ctx, A)
    s = D.struct()
    lay = A.layout()
    nl = A.dims["nlevsoi"]
    cond_ms, sink_ms = time_events(ctx, [lambda: ctx.phs_soil_root_conductance(lay, s, A.npatch, None, dev=True),
                                         lambda: ctx.phs_root_sink(lay, s, A.ncol, None, dev=True)], 3, inner=10)
    D.free()
    # unique bytes: per patch its scalars, rootfr and the three outputs per layer;
    # per column z, dz, hksat, hk_l, smp_l per layer
    cb = A.npatch * (32 + nl * 32) + A.ncol * nl * 40
    used = (A["active"] != 0) & (A["frac_veg_nosno"] > 0) & (A["wtcol"] > 0)
    # per column patch range, z, smp_l and the output per layer, qflx_phs_neg; per
    # patch the three tests and the output, and for contributing patches
    # vegwp(root) and k_soil_root per layer
    sb = A.ncol * (16 + nl * 24) + A.npatch * 24 + int(used.sum()) * (8 + nl * 8)
    roof = lambda b, ms: {"alg_bytes_per_launch": b, "achieved_GBps": b / (ms * 1e-3) / 1e9,
                          "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out = {"workload": f"{A.ncol:,} columns, {A.npatch:,} patches (0-4 per column), {nl} soil layers",
           "conductance_kernel_ms": cond_ms, "root_sink_kernel_ms": sink_ms,
           "conductance_roofline": dict(kernel="k_phs_conductance", **roof(cb, cond_ms)),
           "root_sink_roofline": dict(kernel="k_phs_root_sink", **roof(sb, sink_ms))}
    # calcstress: the Newton solve of the vegetation water potential, every patch
    # once per call (the input is restored before each timed call)
    B = PH.synthetic_stress(ncol, max_patches=4, seed=4)
    D = PH.DevicePhs(ctx, B)
    s2 = D.struct()
    lay2 = B.layout()
    vw = D.ptr["vegwp"]
    ctx.phs_calcstress(lay2, s2, B.npatch, None, dev=True)
    stress_ms, = time_events(ctx, [lambda: ctx.phs_calcstress(lay2, s2, B.npatch, None, dev=True)], 5,
                             prefill=lambda: ctx.h2d(vw, B["vegwp"]))
    D.free()
    out["calcstress"] = {"patches": B.npatch, "kernel_ms": stress_ms,
                         "patch_solves_per_s": B.npatch / (stress_ms * 1e-3),
                         "bound": "FP64 latency (Newton iterations, 12 pow per iteration)"}
    return out


def _history_and_checkpoint(ctx, ds, structs):
    """The default history tape's per-step update (one launch over every field,
    buffers in HBM) with its HBM roofline, and a checkpoint of the whole
    resident state written and read back through the library."""
    import os
    import shutil
    import tempfile
    from ctsm_amd import restart
    from ctsm_amd.history import HistoryTape
    tape = HistoryTape(ctx, ds)
    elems = sum(f.n1d * f.num2d for f in tape.desc)
    upd_ms, = time_events(ctx, [tape.update], 3, prefill=lambda: ctx.canopyfluxes_dev(ds, structs=structs), inner=10)
    tape.free()
    hb = elems * (8 + 8 + 4 + 8 + 4)        # field read, hbuf and nacs read and written
    out = {"fields": len(tape.specs), "elements": elems, "update_kernel_ms": upd_ms,
           "update_roofline": {"kernel": "k_hist_update", "alg_bytes_per_launch": hb,
                               "achieved_GBps": hb / (upd_ms * 1e-3) / 1e9,
                               "frac": hb / (upd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    tmp = tempfile.gettempdir()
    need = ds.nbytes()
    if shutil.disk_usage(tmp).free < 3 * need:
        out["checkpoint"] = f"skipped: {need / 1e9:.1f} GB state, too little space in {tmp}"
        return out
    fd, path = tempfile.mkstemp(suffix=".ckpt")
    os.close(fd)
    try:
        t0 = time.perf_counter()
        restart.checkpoint(ctx, ds, path)
        t_write = time.perf_counter() - t0
        nbytes = os.path.getsize(path)
        t0 = time.perf_counter()
        ds2, _ = restart.restore(ctx, path)
        t_read = time.perf_counter() - t0
        ds2.free()
        out.update({"checkpoint_bytes": nbytes, "checkpoint_write_GBps": nbytes / t_write / 1e9,
                    "checkpoint_restore_GBps": nbytes / t_read / 1e9})
    except Exception as e:      # a full or slow file system must not end the benchmark
        out["checkpoint"] = f"failed: {e}"
    finally:
        os.remove(path)
    return out


def _full_column_parts(ctx, args, st):
    """The same batch as concurrent sub-batches: contiguous parts of the
    binned order (balanced, as the bins are dealt over the parts), each
    resident in its own context and stream and stepped without waiting for
    the others, so one part's long canopy rounds overlap another's work."""
    from ctsm_amd import synthetic
    from ctsm_amd.lib import Context
    from ctsm_amd.state import DeviceState
    P = args.canopy_parts
    n = st.dims.ncol
    bounds = [n * k // P for k in range(P + 1)]
    tun = dict(ctx.get_tuning(), step_streams=1)
    ctxs = [ctx] + [Context() for _ in range(P - 1)]
    before = ctx.get_tuning()
    for c in ctxs:
        c.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
        c.set_tuning(**tun)
    dss = [DeviceState(st.subset(np.arange(bounds[k], bounds[k + 1])), c) for k, c in enumerate(ctxs)]
    sts = [(*d.structs(), d.canopy_struct()) for d in dss]
    for c, d, s in zip(ctxs, dss, sts):
        c.land_step_dev(d, structs=s)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for c, d, s in zip(ctxs, dss, sts):
        for _ in range(args.canopy_steps):
            c.land_step_dev(d, structs=s)
    for c in ctxs:
        c.sync()
    wall = time.perf_counter() - t0
    for d in dss:
        d.free()
    for c in ctxs[1:]:
        c.close()
    ctx.set_tuning(**before)
    return {"parts": P, "column_steps_per_s": n * args.canopy_steps / wall,
            "patch_steps_per_s": st.dims.npatch * args.canopy_steps / wall,
            "step_ms": wall / args.canopy_steps * 1e3}


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
    ap.add_argument("--filter-sort", action="store_true",
                    help="bin through a sorted column filter instead of storing the batch in binned order")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--streams", type=int, default=0,
                    help="with --lanes: streams the step's columns are split across (default 4)")
    ap.add_argument("--no-graph", action="store_true", help="launch each step eagerly instead of replaying "
                                                            "the captured step")
    ap.add_argument("--lanes", type=int, default=0, help="force lanes per column in the Richards solve")
    ap.add_argument("--single-kernel", action="store_true", help="force the one-kernel fused step")
    ap.add_argument("--order", choices=["bins", "substeps"], default="substeps",
                    help="column order: (bedrock, snow) bins, or also by the sub-steps of the warm-up step")
    ap.add_argument("--canopy-ncol", type=int, default=262144,
                    help="columns of the full-column (CanopyFluxes) measurement; 0 skips it")
    ap.add_argument("--canopy-steps", type=int, default=6)
    ap.add_argument("--canopy-parts", type=int, default=1,
                    help="also run the full-column batch as this many concurrent sub-batches (1 = no)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--serve-requests", type=int, default=80,
                    help="served-ensemble measurement: number of requests (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-day-scale", type=float, default=0.1,
                    help="run lengths are log-uniform 1-365 days x 48 steps x this factor")
    ap.add_argument("--serve-pool-cols", type=int, default=131072)
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

    # Column binning.  Columns are independent, so the batch may be stored in
    # any order: the device-resident batch is stored grouped by (snow layers,
    # Richards sub-steps of the warm-up steps, bedrock depth), which keeps a
    # wavefront's columns on similar work AND its memory accesses coalesced
    # (a sorted index filter gives the first but not the second).
    filt = None
    binning = "none"
    if not args.no_sort:
        ds = DeviceState(st, ctx)
        for _ in range(max(args.warmup, 1)):
            ctx.step_dev(ds)
        ctx.sync()
        stw = ds.to_host()
        ds.free()
        order = AT.column_order(stw.col["snl"], stw.col["nbedrock"],
                                stw.col["num_substeps"] if args.order == "substeps" else None, parts=4) - 1
        if args.filter_sort:
            filt = ctx.device_array(order + 1)
            st = stw
            binning = "sorted index filter"
        else:
            st = stw.subset(order)
            binning = "batch stored in binned column order"

    tuning, table = None, []
    if args.lanes or args.single_kernel:
        lanes = args.lanes or 1
        ctx.set_tuning(richards_lanes=lanes, richards_block=64 if lanes == 1 else 128,
                       fused_single_kernel=int(args.single_kernel),
                       step_streams=args.streams or (1 if args.single_kernel else 4))
    elif not args.no_autotune:
        ds0 = DeviceState(st, ctx)
        tuning, _, table = AT.autotune(ctx, ds0, filter_dev=filt, steps=8, reps=2)
        ds0.free()

    def run(plain):
        nonlocal filt
        ds = DeviceState(st, ctx)
        structs = ds.structs()
        for _ in range(args.warmup):
            ctx.step_dev(ds, filter_dev=filt, plain=plain, structs=structs)
        ctx.sync()
        # the timed steps are captured once as one graph (each part of the
        # split step runs all steps on its own stream) and launched once
        graph = None
        if not plain and not args.no_graph:
            graph = ctx.graph_capture(ds, nsteps=args.steps, canopy=False, filterc_dev=filt,
                                      structs=(*structs, None))
        # events on the library's own stream, around the launch(es)
        starts, stops = [], []
        t0 = time.perf_counter()
        if graph is not None:
            starts.append(ctx.record())
            ctx.graph_launch(graph)
            stops.append(ctx.record())
        else:
            for k in range(args.steps):
                starts.append(ctx.record())
                ctx.step_dev(ds, filter_dev=filt, plain=plain, structs=structs)
                stops.append(ctx.record())
        ctx.sync()
        wall = time.perf_counter() - t0
        per = [ctx.elapsed_ms(a, b) / (args.steps if graph is not None else 1) for a, b in zip(starts, stops)]
        for e in starts + stops:
            ctx.destroy_event(e)
        if graph is not None:
            ctx.graph_destroy(graph)
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

    # per-kernel launch times over the same trajectory as the timed steps
    # (events on the library stream): heat solve, water fractions, Richards --
    # the step's work launched as separate kernels, averaged over the steps
    ds = DeviceState(st, ctx)
    for _ in range(args.warmup):
        ctx.step_dev(ds, filter_dev=filt)
    t_heat, t_frac, t_rich = time_events(ctx, [lambda: ctx.soiltemperature_dev(ds, filt),
                                               lambda: ctx.soilwater_fractions_dev(ds, filt),
                                               lambda: ctx.soilwater_dev(ds, filt)], args.steps)
    ds.free()
    res["kernel_ms"] = {"soiltemp (k_soiltemp)": t_heat, "fractions (k_soilwater_fractions)": t_frac,
                        "richards (k_soilwater_g)": t_rich}
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

    b_rich, b_heat = richards_bytes(st), soiltemp_bytes(st)
    lanes = ctx.get_tuning()["richards_lanes"]
    kern = {"richards": (t_rich, b_rich, "k_soilwater_g<%d>" % lanes if lanes > 1 else "k_soilwater"),
            "soiltemp": (t_heat, b_heat, "k_soiltemp")}
    dom = max(kern, key=lambda k: kern[k][0])
    kms, kbytes, kname = kern[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    res["roofline_other"] = {k: {"kernel": v[2], "ms": v[0], "alg_bytes_per_launch": v[1],
                                 "achieved_GBps": v[1] / (v[0] * 1e-3) / 1e9} for k, v in kern.items() if k != dom}
    if args.canopy_ncol > 0:
        res["full_column_step"] = full_column(ctx, args)
    if args.serve_requests > 0:
        res["served_ensemble"] = served_ensemble(ctx, args)
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
                   "ncol": args.ncol, "binning": binning,
                   "launch": "eager, one call per step" if args.no_graph else
                             "the timed steps captured as one graph, launched once"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC.get(kname),
                     "kernel": kname, "launch_ms": kms, "alg_bytes_per_launch": kbytes,
                     "fp64_valu_frac": FP64_VALU_FRAC.get(kname)},
        "detail": res,
    }
    if not args.no_cpu:
        v, threads, sample = cpu_baseline(st, args.cpu_seconds, args.cpu_threads or None)
        out["cpu_baseline"] = {"value": v, "unit": "column-steps/s", "cores": threads, "kind": "port",
                               "sample": sample}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
